// gfx950 kernels of the reachable-pore labelling.
//
// The reference labels 6-connected components of "spans" (fastlib/src/cuda/VolumeCCL.cu:86-375):
// scanning each x-row from a `background` previous value, every change of value toggles a
// span open/closed; spans in neighbouring rows that overlap in x are connected; a component's
// label is the minimum span-start linear index, then labels are renumbered 1..K in increasing
// order of that index.  Two in-span voxels adjacent along x always share a span, so the
// components are exactly the 6-connected components of the in-span voxel set, and the
// renumbering orders components by their minimum linear voxel index.
//
// Here: (1) one thread per row computes the span set and points every in-span voxel at its
// span start; (2) a lock-free union-find links rows across y and z, always towards the lower
// index, so each component's root is its minimum index; (3) paths are flattened; (4) a
// deterministic prefix count of roots gives every root its rank, i.e. the reference label.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace ftc {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanChunk = kScanThreads * kScanPer;

namespace {

__device__ __forceinline__ uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t find_root(const uint32_t* P, uint32_t x) {
    uint32_t p = ld(P + x);
    while (p != x) {
        x = p;
        p = ld(P + x);
    }
    return x;
}

__device__ __forceinline__ void merge(uint32_t* P, uint32_t a, uint32_t b) {
    for (;;) {
        a = find_root(P, a);
        b = find_root(P, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t old = atomicMin(P + b, a);
            if (old == b) return;
            b = old;
        } else {
            const uint32_t old = atomicMin(P + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

// Span starts, one wave per x-row in chunks of 64 voxels; the waves stride over the rows, so
// the grid stays within the launch limit however many rows there are.  A voxel is in a span
// when the number of value changes up to it (from `background` before the row) is odd, and
// its span starts at the last change at or before it; both come from a ballot of the change
// flags, with the parity and the last change carried from chunk to chunk.
__global__ __launch_bounds__(256) void k_spans(const uint8_t* __restrict__ mask, uint32_t* __restrict__ P, int nx,
                                               int rows, uint8_t background) {
    const int lane = threadIdx.x & 63;
    const unsigned long long upto = (2ull << lane) - 1ull;   // lanes 0..lane
    for (long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); row < rows; row += (long long)gridDim.x * 4) {
    const long long base = row * nx;
    int prev_last = background;   // value before this chunk
    unsigned parity = 0;          // number of changes before this chunk, mod 2
    uint32_t last_change = kNone; // index of the last change before this chunk
    for (int x0 = 0; x0 < nx; x0 += 64) {
        const int x = x0 + lane;
        const bool in = x < nx;
        const int v = in ? (int)mask[base + x] : 0;
        int pv = __shfl_up(v, 1, 64);
        if (lane == 0) pv = prev_last;
        const unsigned long long changes = __ballot(in && v != pv);
        const unsigned long long mine = changes & upto;
        const bool open = ((parity + (unsigned)__popcll(mine)) & 1u) != 0u;
        const uint32_t start = mine ? (uint32_t)(base + x0 + (63 - __clzll((long long)mine))) : last_change;
        if (in) P[base + x] = open ? start : kNone;
        parity = (parity + (unsigned)__popcll(changes)) & 1u;
        if (changes) last_change = (uint32_t)(base + x0 + (63 - __clzll((long long)changes)));
        prev_last = __shfl(v, 63, 64);
    }
    }
}

// Link every span to the spans it overlaps in the rows at y-1 and z-1.  A voxel whose byte
// equals its x-1 neighbour's lies in the same span as that neighbour (spans change only where
// the value changes), so where neither this row nor the other row changes value between x-1
// and x, the pair of spans was already linked at x-1 and the voxel is skipped.  Span starts are
// the only entries the merges write; every other in-span voxel keeps pointing at its start.
__global__ __launch_bounds__(256) void k_union(const uint8_t* __restrict__ mask, uint32_t* P, int nx, int ny, int nz) {
    const long long n = (long long)nx * ny * nz;
    const long long plane = (long long)nx * ny;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (ld(P + i) == kNone) continue;
        const unsigned row = (unsigned)i / (unsigned)nx;   // labels are 32-bit, so i < 2^32
        const int x = (int)((unsigned)i - row * (unsigned)nx);
        const int y = (int)(row % (unsigned)ny);
        const int z = (int)(row / (unsigned)ny);
        const bool same_here = x > 0 && mask[i] == mask[i - 1];
        if (y > 0 && ld(P + i - nx) != kNone && !(same_here && mask[i - nx] == mask[i - nx - 1]))
            merge(P, (uint32_t)i, (uint32_t)(i - nx));
        if (z > 0 && ld(P + i - plane) != kNone && !(same_here && mask[i - plane] == mask[i - plane - 1]))
            merge(P, (uint32_t)i, (uint32_t)(i - plane));
    }
}

__global__ __launch_bounds__(256) void k_flatten(uint32_t* P, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t p = P[i];
        if (p == kNone) continue;
        P[i] = find_root(P, (uint32_t)i);
    }
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive block scan of one value per thread (kScanThreads threads); returns block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* excl) {
    __shared__ uint32_t wtot[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
        if (i < w) before += wtot[i];
        total += wtot[i];
    }
    *excl = before + inc - v;
    __syncthreads();
    return total;
}

__device__ __forceinline__ bool is_root(const uint32_t* P, long long i) { return P[i] == (uint32_t)i; }

__global__ __launch_bounds__(kScanThreads) void k_count_roots(const uint32_t* P, long long n, uint32_t* blocksum) {
    const long long b0 = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanPer;
    uint32_t c = 0;
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = b0 + k;
        if (i < n && is_root(P, i)) ++c;
    }
    uint32_t excl;
    const uint32_t tot = block_excl_scan(c, &excl);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = tot;
}

// exclusive scan of nb block sums in one block; total number of roots to *total
__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* blocksum, int nb, uint32_t* total) {
    __shared__ uint32_t part[1024];
    const int per = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nb) s += blocksum[b0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            const uint32_t t = part[i];
            part[i] = run;
            run += t;
        }
        *total = run;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int k = 0; k < per; ++k) {
        if (b0 + k < nb) {
            const uint32_t t = blocksum[b0 + k];
            blocksum[b0 + k] = run;
            run += t;
        }
    }
}

// roots get label = 1 + number of roots before them in linear order
__global__ __launch_bounds__(kScanThreads) void k_assign_roots(const uint32_t* P, long long n, const uint32_t* blockoff,
                                                               uint32_t* L) {
    const long long b0 = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanPer;
    uint32_t c = 0;
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = b0 + k;
        if (i < n && is_root(P, i)) ++c;
    }
    uint32_t excl;
    (void)block_excl_scan(c, &excl);
    uint32_t lab = blockoff[blockIdx.x] + excl;
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = b0 + k;
        if (i < n && is_root(P, i)) L[i] = ++lab;
    }
}

__global__ __launch_bounds__(256) void k_final_labels(const uint32_t* P, long long n, uint32_t* L) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t p = P[i];
        if (p == kNone)
            L[i] = 0;
        else if (p != (uint32_t)i)
            L[i] = L[p];   // roots were labelled by the previous kernel and are not written here
    }
}

// face table: row f, label l present on face f (Dir order X+,X-,Y+,Y-,Z+,Z-)
__global__ __launch_bounds__(256) void k_face_table(const uint32_t* L, int nx, int ny, int nz, uint8_t* table,
                                                    uint32_t numLabels) {
    const int f = blockIdx.y;
    const int k = f >> 1;
    const int dims[3] = {nx, ny, nz};
    const int fixed = (f & 1) ? 0 : dims[k] - 1;
    const int s0 = (k + 1) % 3, s1 = (k + 2) % 3;
    const long long m = (long long)dims[s0] * dims[s1];
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (long long)gridDim.x * blockDim.x) {
        int pos[3];
        pos[k] = fixed;
        pos[s0] = (int)(q % dims[s0]);
        pos[s1] = (int)(q / dims[s0]);
        const uint32_t l = L[(long long)pos[0] + (long long)nx * ((long long)pos[1] + (long long)ny * pos[2])];
        if (l > 0) table[(size_t)f * numLabels + l] = 1;
    }
}

__global__ __launch_bounds__(256) void k_face_all(uint8_t* table, uint32_t numLabels) {
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < numLabels; l += gridDim.x * blockDim.x) {
        uint8_t a = 1;
        for (int f = 0; f < 6; ++f) a &= table[(size_t)f * numLabels + l];
        table[(size_t)6 * numLabels + l] = a;
    }
}

__global__ __launch_bounds__(256) void k_reachable(const uint32_t* L, long long n, const uint8_t* table, uint32_t numLabels,
                                                   int fa, int fb, uint8_t* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t l = L[i];
        out[i] = (table[(size_t)fa * numLabels + l] && table[(size_t)fb * numLabels + l]) ? 255 : 0;
    }
}

// Face bits of every voxel: bit f set when its label touches face f (rows 0..5 of the face
// table; label 0, the background, is on no row).  One byte per voxel for the six
// boundary-connected volumes of generateBoundaryConnectedVolume (VolumeSegmentation.cpp:66-90).
__global__ __launch_bounds__(256) void k_face_codes(const uint32_t* L, long long n, const uint8_t* table,
                                                    uint32_t numLabels, uint8_t* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t l = L[i];
        uint32_t bits = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) bits |= (uint32_t)(table[(size_t)f * numLabels + l] != 0) << f;
        out[i] = (uint8_t)bits;
    }
}

// Per-label statistics (getCCLSegmentInfo, VolumeSegmentation.cpp:135-178): voxel count and
// bounding box.  info[7 l ..]: count, min x, y, z, max x, y, z (inclusive).  Every x-run of
// equal labels contributes once, from its first voxel, with integer atomics (the result does
// not depend on their order).
__global__ __launch_bounds__(256) void k_seg_init(uint32_t* info, uint32_t numLabels, int nx, int ny, int nz) {
    for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < numLabels;
         l += (long long)gridDim.x * blockDim.x) {
        uint32_t* r = info + 7 * l;
        r[0] = 0;
        r[1] = (uint32_t)nx;
        r[2] = (uint32_t)ny;
        r[3] = (uint32_t)nz;
        r[4] = r[5] = r[6] = 0;
    }
}

__global__ __launch_bounds__(256) void k_seg_runs(const uint32_t* L, int nx, int ny, int nz, uint32_t* info) {
    const long long n = (long long)nx * ny * nz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t l = L[i];
        const unsigned row = (unsigned)i / (unsigned)nx;
        const int x = (int)((unsigned)i - row * (unsigned)nx);
        if (x > 0 && L[i - 1] == l) continue;   // not the first voxel of its run
        int e = x;
        while (e + 1 < nx && L[i + (e + 1 - x)] == l) ++e;
        const int y = (int)(row % (unsigned)ny);
        const int z = (int)(row / (unsigned)ny);
        uint32_t* r = info + 7 * (size_t)l;
        atomicAdd(r + 0, (uint32_t)(e - x + 1));
        atomicMin(r + 1, (uint32_t)x);
        atomicMin(r + 2, (uint32_t)y);
        atomicMin(r + 3, (uint32_t)z);
        atomicMax(r + 4, (uint32_t)e);
        atomicMax(r + 5, (uint32_t)y);
        atomicMax(r + 6, (uint32_t)z);
    }
}

inline unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

// Labels (0 = background, 1..K canonical) into L; returns K via *num_roots (device, 1 uint32).
// P, L: n uint32 each; blocksum: ceil(n / 4096) uint32.
hipError_t ccl_label(const uint8_t* mask, int nx, int ny, int nz, uint8_t background, uint32_t* P, uint32_t* L,
                     uint32_t* blocksum, uint32_t* num_roots, hipStream_t s) {
    const long long n = (long long)nx * ny * nz;
    const int rows = ny * nz;   // < 2^31: volumes hold fewer than 2^31 voxels
    const int span_blocks = std::min((rows + 3) / 4, 65536);   // 4 rows per block and sweep
    hipLaunchKernelGGL(k_spans, dim3(span_blocks), dim3(256), 0, s, mask, P, nx, rows, background);
    hipLaunchKernelGGL(k_union, dim3(grid_for(n)), dim3(256), 0, s, mask, P, nx, ny, nz);
    hipLaunchKernelGGL(k_flatten, dim3(grid_for(n)), dim3(256), 0, s, P, n);
    const int nb = (int)((n + kScanChunk - 1) / kScanChunk);
    hipLaunchKernelGGL(k_count_roots, dim3(nb), dim3(kScanThreads), 0, s, (const uint32_t*)P, n, blocksum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, blocksum, nb, num_roots);
    hipLaunchKernelGGL(k_assign_roots, dim3(nb), dim3(kScanThreads), 0, s, (const uint32_t*)P, n,
                       (const uint32_t*)blocksum, L);
    hipLaunchKernelGGL(k_final_labels, dim3(grid_for(n)), dim3(256), 0, s, (const uint32_t*)P, n, L);
    return hipGetLastError();
}

// 7 x numLabels face table (must be zeroed by the caller).
hipError_t ccl_face_table(const uint32_t* L, int nx, int ny, int nz, uint8_t* table, uint32_t numLabels, hipStream_t s) {
    const long long maxface = (long long)std::max(nx, ny) * std::max(std::max(ny, nz), nx) + (long long)nz * nx;
    hipLaunchKernelGGL(k_face_table, dim3(grid_for(maxface), 6), dim3(256), 0, s, L, nx, ny, nz, table, numLabels);
    hipLaunchKernelGGL(k_face_all, dim3(grid_for(numLabels)), dim3(256), 0, s, table, numLabels);
    return hipGetLastError();
}

hipError_t ccl_reachable(const uint32_t* L, long long n, const uint8_t* table, uint32_t numLabels, int dir, uint8_t* out,
                         hipStream_t s) {
    const int k = dir / 2;
    hipLaunchKernelGGL(k_reachable, dim3(grid_for(n)), dim3(256), 0, s, L, n, table, numLabels, 2 * k, 2 * k + 1, out);
    return hipGetLastError();
}

hipError_t ccl_segment_stats(const uint32_t* L, int nx, int ny, int nz, uint32_t numLabels, uint32_t* info,
                             hipStream_t s) {
    const long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_seg_init, dim3(grid_for(numLabels)), dim3(256), 0, s, info, numLabels, nx, ny, nz);
    hipLaunchKernelGGL(k_seg_runs, dim3(grid_for(n)), dim3(256), 0, s, L, nx, ny, nz, info);
    return hipGetLastError();
}

hipError_t ccl_face_codes(const uint32_t* L, long long n, const uint8_t* table, uint32_t numLabels, uint8_t* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_face_codes, dim3(grid_for(n)), dim3(256), 0, s, L, n, table, numLabels, out);
    return hipGetLastError();
}

}  // namespace ftc
