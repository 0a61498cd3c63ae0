// Reactive area density: marching-cubes area of the isovalue level set of a volume.
//
// Reference (`fast alpha`): getReactiveAreaDensity / getReactiveAreaDensityTensor
// (fastlib/src/volume/VolumeMeasures.cpp:294-333) -> getVolumeArea
// (fastlib/src/volume/VolumeSurface.cpp:9-36) -> ___generateTrianglesSurfaceArea
// (fastlib/src/cuda/VolumeSurface.cu:203-258), then Volume::sum of the per-voxel areas.
//
// The reference samples the volume as a linearly filtered, clamped, normalized-coordinate
// uchar texture (fastlib/src/utility/DataPtr.cpp:565-592) at the corners of every voxel
// (VolumeSurface.cu:83-103, positions i/N).  A corner sits on the midpoint between texel
// centres in every axis, so its value is the mean of the 8 voxels around it (indices clamped
// into the volume), byte/255: (sum of the 8 bytes) / 2040, which for a 0/255 volume is k/8
// exactly.  Here those corner values are built from voxel sums instead of texture reads:
// a block keeps, for its 64x4 tile and the plane being marched, the 2x2 in-plane voxel sums
// of every lattice point (LDS) and adds two consecutive planes of them into corner counts;
// each thread classifies its cube from the 8 counts, and the cubes that cross the level set
// are queued and triangulated by the whole block.
//
// Two modes share the march.  MODE 0 reads the volume's bytes (one surface).  MODE 1 reads
// a byte of face bits per voxel (bit f: a pore voxel whose component touches face f; the
// boundary-connected volumes of getReactiveAreaDensityTensor) and carries the six per-face
// counts packed in nibbles, so one pass yields all six surfaces.
//
// Per cube the arithmetic is the reference's float arithmetic, operation by operation and
// without contraction (the file is built with -ffp-contract=off): positions (1/N)*i and
// (1/N)*i + 1/N, vertexInterp / lerp (VolumeSurface.cu:171-178, CudaMath.h:1134-1137),
// triangleArea = 0.5 * length(cross(v1 - v0, v2 - v0)) (VolumeSurface.cu:187-189), the
// cube's triangle areas summed in double in table order (VolumeSurface.cu:240-255).  The
// volume sum is double-double with one final rounding, so it does not depend on the
// reduction tree.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <mutex>

#include "mc_table.h"

namespace fta {

constexpr int kTX = 64, kTY = 4, kTZ = 16, kNT = kTX * kTY;
constexpr int kLX = kTX + 1, kLY = kTY + 1, kNL = kLX * kLY;   // lattice points of a tile
constexpr int kVX = kTX + 2, kVY = kTY + 2, kNV = kVX * kVY;   // voxels of a tile plane with halo

__constant__ uint64_t c_tri[256] = FAST_MC_TRI_WORDS;

// Double-double sum of doubles (TwoSum, then renormalise).
struct DdSum {
    double hi, lo;
    __device__ __forceinline__ void init() { hi = 0.0; lo = 0.0; }
    __device__ __forceinline__ void add2(double p, double e) {
        const double s = hi + p;
        const double v = s - hi;
        double t = (hi - (s - v)) + (p - v);
        t += lo + e;
        hi = s + t;
        lo = t - (hi - s);
    }
};

__device__ __forceinline__ void wave_combine(DdSum& a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double h = __shfl_xor(a.hi, off, 64);
        const double l = __shfl_xor(a.lo, off, 64);
        a.add2(h, l);
    }
}

// Face bits 0..5 of a code byte, one per nibble.
__device__ __forceinline__ uint32_t spread6(uint32_t b) {
    return (b & 1u) | ((b & 2u) << 3) | ((b & 4u) << 6) | ((b & 8u) << 9) | ((b & 16u) << 12) | ((b & 32u) << 15);
}

// Corner coordinates of one cube: X0 = (1/nx) i, X1 = X0 + 1/nx, likewise y and z.
struct CubeGeo {
    float X0, X1, Y0, Y1, Z0, Z1;
};

// vertexInterp along one axis: the other two coordinates of the edge's end points are equal,
// so p0 + t (p1 - p0) leaves them unchanged.
__device__ __forceinline__ float lerp_axis(float iso, float fa, float fb, float A, float B) {
    if (fb == fa) return A;
    const float t = (iso - fa) / (fb - fa);
    return A + t * (B - A);
}

// Edge e (corners a -> b of the table's edge list) as 5 bits at 5e: bits 0-1 the axis along
// which it runs, bits 2-4 whether its fixed x, y, z coordinates are the cube's upper ones.
constexpr uint64_t edge_desc(int axis, int x1, int y1, int z1) {
    return (uint64_t)(axis | (x1 << 2) | (y1 << 3) | (z1 << 4));
}
constexpr uint64_t kEdgeDesc =
    edge_desc(0, 0, 0, 0) << 0 | edge_desc(1, 1, 0, 0) << 5 | edge_desc(0, 0, 1, 0) << 10 | edge_desc(1, 0, 0, 0) << 15 |
    edge_desc(0, 0, 0, 1) << 20 | edge_desc(1, 1, 0, 1) << 25 | edge_desc(0, 0, 1, 1) << 30 | edge_desc(1, 0, 0, 1) << 35 |
    edge_desc(2, 0, 0, 0) << 40 | edge_desc(2, 1, 0, 0) << 45 | edge_desc(2, 1, 1, 0) << 50 | edge_desc(2, 0, 1, 0) << 55;

// Vertex on edge e (corners a -> b as in the table's edge list), by cases: used where a
// thread triangulates its own cube (the mesh export).
__device__ __forceinline__ void edge_vertex_switch(unsigned e, const float (&f)[8], const CubeGeo& g, float iso,
                                                   float& x, float& y, float& z) {
    switch (e) {
        case 0: x = lerp_axis(iso, f[0], f[1], g.X0, g.X1); y = g.Y0; z = g.Z0; break;
        case 1: x = g.X1; y = lerp_axis(iso, f[1], f[2], g.Y0, g.Y1); z = g.Z0; break;
        case 2: x = lerp_axis(iso, f[2], f[3], g.X1, g.X0); y = g.Y1; z = g.Z0; break;
        case 3: x = g.X0; y = lerp_axis(iso, f[3], f[0], g.Y1, g.Y0); z = g.Z0; break;
        case 4: x = lerp_axis(iso, f[4], f[5], g.X0, g.X1); y = g.Y0; z = g.Z1; break;
        case 5: x = g.X1; y = lerp_axis(iso, f[5], f[6], g.Y0, g.Y1); z = g.Z1; break;
        case 6: x = lerp_axis(iso, f[6], f[7], g.X1, g.X0); y = g.Y1; z = g.Z1; break;
        case 7: x = g.X0; y = lerp_axis(iso, f[7], f[4], g.Y1, g.Y0); z = g.Z1; break;
        case 8: x = g.X0; y = g.Y0; z = lerp_axis(iso, f[0], f[4], g.Z0, g.Z1); break;
        case 9: x = g.X1; y = g.Y0; z = lerp_axis(iso, f[1], f[5], g.Z0, g.Z1); break;
        case 10: x = g.X1; y = g.Y1; z = lerp_axis(iso, f[2], f[6], g.Z0, g.Z1); break;
        default: x = g.X0; y = g.Y1; z = lerp_axis(iso, f[3], f[7], g.Z0, g.Z1); break;
    }
}

// Marching-cubes area of one cube with corner values f (table order of the triangles).  The
// interpolated coordinate of every edge is computed once, without branching on the edge, and
// kept in this thread's LDS slots ev[12 * kNT] (stride kNT); a triangle vertex is then its
// edge's slot plus the two fixed coordinates, so lanes of a wave never diverge on edge ids.
__device__ __forceinline__ double cube_area(const float (&f)[8], const CubeGeo& g, float iso, float* ev) {
    unsigned ci = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) ci |= (f[m] < iso ? 1u : 0u) << m;
    ev[0 * kNT] = lerp_axis(iso, f[0], f[1], g.X0, g.X1);
    ev[1 * kNT] = lerp_axis(iso, f[1], f[2], g.Y0, g.Y1);
    ev[2 * kNT] = lerp_axis(iso, f[2], f[3], g.X1, g.X0);
    ev[3 * kNT] = lerp_axis(iso, f[3], f[0], g.Y1, g.Y0);
    ev[4 * kNT] = lerp_axis(iso, f[4], f[5], g.X0, g.X1);
    ev[5 * kNT] = lerp_axis(iso, f[5], f[6], g.Y0, g.Y1);
    ev[6 * kNT] = lerp_axis(iso, f[6], f[7], g.X1, g.X0);
    ev[7 * kNT] = lerp_axis(iso, f[7], f[4], g.Y1, g.Y0);
    ev[8 * kNT] = lerp_axis(iso, f[0], f[4], g.Z0, g.Z1);
    ev[9 * kNT] = lerp_axis(iso, f[1], f[5], g.Z0, g.Z1);
    ev[10 * kNT] = lerp_axis(iso, f[2], f[6], g.Z0, g.Z1);
    ev[11 * kNT] = lerp_axis(iso, f[3], f[7], g.Z0, g.Z1);
    const uint64_t w = c_tri[ci];
    const int ntri = (int)(w >> 60);
    auto vertex = [&](unsigned e, float& x, float& y, float& z) {
        const unsigned d = (unsigned)(kEdgeDesc >> (5 * e)) & 31u;
        const float v = ev[e * kNT];
        const unsigned axis = d & 3u;
        x = axis == 0u ? v : ((d & 4u) ? g.X1 : g.X0);
        y = axis == 1u ? v : ((d & 8u) ? g.Y1 : g.Y0);
        z = axis == 2u ? v : ((d & 16u) ? g.Z1 : g.Z0);
    };
    double area = 0.0;
    for (int j = 0; j < ntri; ++j) {
        float ax, ay, az, bx, by, bz, cx, cy, cz;
        vertex((unsigned)(w >> (12 * j)) & 15u, ax, ay, az);
        vertex((unsigned)(w >> (12 * j + 4)) & 15u, bx, by, bz);
        vertex((unsigned)(w >> (12 * j + 8)) & 15u, cx, cy, cz);
        const float ux = bx - ax, uy = by - ay, uz = bz - az;
        const float vx = cx - ax, vy = cy - ay, vz = cz - az;
        const float nx = uy * vz - uz * vy;
        const float ny = uz * vx - ux * vz;
        const float nz = ux * vy - uy * vx;
        area += (double)(sqrtf(nx * nx + ny * ny + nz * nz) * 0.5f);
    }
    return area;
}

// One block: cubes [x0, x0+64) x [y0, y0+4) x [k0, k0+16).  Writes NS double-double
// partials (hi, lo) per block to part[(block * NS + s) * 2].
// thr: the smallest corner count whose value (count / 2040, or / 8 for face bits) is not
// below the isovalue.  The value is monotone in the count, so "corner inside" is count < thr.
// Per layer every thread classifies its cube with integer work only; cubes with corners on
// both sides (per face in MODE 1) are queued in LDS (wave ballot + one LDS atomic per wave).
// Once the queue holds a block's worth of cubes (or at the last layer) the whole block
// triangulates it densely, so the float work runs on full waves instead of on the few
// surface lanes of every wave.
template <int MODE>
__global__ __launch_bounds__(kNT) void k_mc_area(const uint8_t* __restrict__ src, int nx, int ny, int nz, float iso,
                                                 uint32_t thr, double* __restrict__ part) {
    constexpr int NS = MODE ? 6 : 1;
    constexpr int QW = MODE ? 1 : 4;          // words of corner counts per queued cube
    constexpr int KQ = kNT + NS * kNT;        // flush at >= kNT items; a layer adds <= NS per thread
    __shared__ uint32_t V[kNV];
    __shared__ uint32_t S[kNL];
    __shared__ uint32_t CP[2][kNL];
    __shared__ double red[kNT / 64][NS][2];
    __shared__ float EV[12 * kNT];            // per-thread edge coordinates of the cube being triangulated
    __shared__ uint32_t QC[QW][KQ];           // queued corner counts (MODE 0: 2 x 16 bits, MODE 1: 8 nibbles)
    __shared__ uint32_t QH[KQ];               // queued cube: thread of the tile | surface mask << 8 | layer << 14
    __shared__ int qn;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY, k0 = blockIdx.z * kTZ;
    const int k1 = min(k0 + kTZ, nz);
    const size_t sxy = (size_t)nx * (size_t)ny;
    const float sx = 1.0f / (float)nx, sy = 1.0f / (float)ny, sz = 1.0f / (float)nz;

    auto load_plane = [&](int z) {
        const uint8_t* plane = src + sxy * (size_t)z;
        for (int v = tid; v < kNV; v += kNT) {
            const int a = min(max(x0 - 1 + v % kVX, 0), nx - 1);
            const int b = min(max(y0 - 1 + v / kVX, 0), ny - 1);
            const uint32_t byte = plane[(size_t)a + (size_t)nx * (size_t)b];
            V[v] = MODE ? spread6(byte) : byte;
        }
    };
    auto lattice_sum = [&](int l) {
        const int lx = l % kLX, ly = l / kLX;
        return V[ly * kVX + lx] + V[ly * kVX + lx + 1] + V[(ly + 1) * kVX + lx] + V[(ly + 1) * kVX + lx + 1];
    };
    // Append this lane's item when `pred`; called by every lane of the wave.
    auto append = [&](bool pred, const uint32_t (&words)[QW], uint32_t head) {
        const unsigned long long mask = __ballot(pred);
        if (mask == 0ull) return;
        int base = 0;
        if (lane == 0) base = atomicAdd(&qn, (int)__popcll(mask));
        base = __shfl(base, 0, 64);
        if (pred) {
            const int slot = base + (int)__popcll(mask & ((1ull << lane) - 1ull));
#pragma unroll
            for (int w = 0; w < QW; ++w) QC[w][slot] = words[w];
            QH[slot] = head;
        }
    };

    // Corner counts of lattice plane k0: voxel planes k0 - 1 and k0 (clamped).
    load_plane(max(k0 - 1, 0));
    __syncthreads();
    for (int l = tid; l < kNL; l += kNT) S[l] = lattice_sum(l);
    __syncthreads();
    load_plane(k0);
    __syncthreads();
    int cur = 0;
    for (int l = tid; l < kNL; l += kNT) {
        const uint32_t s = lattice_sum(l);
        CP[0][l] = S[l] + s;
        S[l] = s;
    }

    const int tx = tid % kTX, ty = tid / kTX;
    const bool active = x0 + tx < nx && y0 + ty < ny;
    DdSum acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s].init();

    if (tid == 0) qn = 0;
    bool flushed = false;
    for (int k = k0; k < k1; ++k) {
        __syncthreads();   // CP[cur] complete; V free; a flushed queue is no longer read
        if (tid == 0 && flushed) qn = 0;
        load_plane(min(k + 1, nz - 1));
        __syncthreads();
        for (int l = tid; l < kNL; l += kNT) {
            const uint32_t s = lattice_sum(l);
            CP[cur ^ 1][l] = S[l] + s;
            S[l] = s;
        }
        __syncthreads();
        // classify this thread's cube (integer work only) and queue its surfaces
        uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (active) {
            const uint32_t* lo = CP[cur];
            const uint32_t* hi = CP[cur ^ 1];
            const int b = ty * kLX + tx;
            c[0] = lo[b];
            c[1] = lo[b + 1];
            c[2] = lo[b + kLX + 1];
            c[3] = lo[b + kLX];
            c[4] = hi[b];
            c[5] = hi[b + 1];
            c[6] = hi[b + kLX + 1];
            c[7] = hi[b + kLX];
        }
        if constexpr (MODE == 0) {
            unsigned in = 0;
#pragma unroll
            for (int m = 0; m < 8; ++m) in |= (c[m] < thr ? 1u : 0u) << m;
            const uint32_t words[QW] = {c[0] | c[1] << 16, c[2] | c[3] << 16, c[4] | c[5] << 16, c[6] | c[7] << 16};
            append(active && in != 0u && in != 255u, words, (uint32_t)tid | (uint32_t)(k - k0) << 14);
        } else {
            uint32_t diff = 0;
#pragma unroll
            for (int m = 1; m < 8; ++m) diff |= c[m] ^ c[0];
            if (__ballot(active && diff != 0u) != 0ull) {   // some cube of the wave has differing corners
                // per face: the 8 counts as nibbles, and whether the face's surface crosses the cube
                uint32_t nib[NS];
                unsigned surf = 0;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    nib[s] = 0;
                    unsigned in = 0;
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        const uint32_t q = (c[m] >> (4 * s)) & 15u;
                        nib[s] |= q << (4 * m);
                        in |= (q < thr ? 1u : 0u) << m;
                    }
                    if (active && in != 0u && in != 255u) surf |= 1u << s;
                }
                // faces with equal counts have the same triangles: queue the cube once, with
                // the mask of those faces (usually all six share the pore space here)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    bool lead = ((surf >> s) & 1u) != 0u;
                    unsigned fm = 0;
#pragma unroll
                    for (int t = 0; t < NS; ++t) {
                        const bool same = ((surf >> t) & 1u) != 0u && nib[t] == nib[s];
                        if (t < s && same) lead = false;
                        if (t >= s && same) fm |= 1u << t;
                    }
                    const uint32_t words[QW] = {nib[s]};
                    append(lead, words, (uint32_t)tid | fm << 8 | (uint32_t)(k - k0) << 14);
                }
            }
        }
        __syncthreads();   // queue complete
        const int n = qn;
        flushed = n >= kNT || k == k1 - 1;   // the same for every thread of the block
        CubeGeo g;
        for (int q = flushed ? tid : n; q < n; q += kNT) {
            const uint32_t h = QH[q];
            const int t = (int)(h & 255u);
            g.Z0 = sz * (float)(k0 + (int)(h >> 14));
            g.Z1 = g.Z0 + sz;
            g.X0 = sx * (float)(x0 + t % kTX);
            g.X1 = g.X0 + sx;
            g.Y0 = sy * (float)(y0 + t / kTX);
            g.Y1 = g.Y0 + sy;
            float f[8];
            if constexpr (MODE == 0) {
#pragma unroll
                for (int m = 0; m < 8; ++m) f[m] = (float)((QC[m >> 1][q] >> (16 * (m & 1))) & 0xFFFFu) / 2040.0f;
            } else {
                const uint32_t w = QC[0][q];
#pragma unroll
                for (int m = 0; m < 8; ++m) f[m] = (float)((w >> (4 * m)) & 15u) / 8.0f;
            }
            const double a = cube_area(f, g, iso, EV + tid);
            if constexpr (MODE == 0) {
                acc[0].add2(a, 0.0);
            } else {
                const unsigned faces = (h >> 8) & 63u;
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    if ((faces >> s) & 1u) acc[s].add2(a, 0.0);
            }
        }
        cur ^= 1;
    }

    // Fixed-order block reduction.
    const int w = tid >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        DdSum a = acc[s];
        wave_combine(a);
        if (lane == 0) {
            red[w][s][0] = a.hi;
            red[w][s][1] = a.lo;
        }
    }
    __syncthreads();
    if (tid < NS) {
        DdSum a;
        a.init();
        for (int q = 0; q < kNT / 64; ++q) a.add2(red[q][tid][0], red[q][tid][1]);
        const size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part[(blk * NS + tid) * 2] = a.hi;
        part[(blk * NS + tid) * 2 + 1] = a.lo;
    }
}

// One block per surface: sum that surface's nblk block partials in a fixed order.
__global__ __launch_bounds__(256) void k_mc_final(const double* __restrict__ part, size_t nblk, int ns,
                                                  double* __restrict__ out) {
    __shared__ double red[4][2];
    const int s = blockIdx.x, tid = threadIdx.x;
    DdSum a;
    a.init();
    for (size_t b = tid; b < nblk; b += 256) a.add2(part[(b * ns + s) * 2], part[(b * ns + s) * 2 + 1]);
    wave_combine(a);
    if ((tid & 63) == 0) {
        red[tid >> 6][0] = a.hi;
        red[tid >> 6][1] = a.lo;
    }
    __syncthreads();
    if (tid == 0) {
        DdSum t;
        t.init();
        for (int q = 0; q < 4; ++q) t.add2(red[q][0], red[q][1]);
        out[s] = t.hi + t.lo;
    }
}

// Scratch of the allocation-free whole-volume area (fast_alpha_area_noalloc): the block
// partials (2 doubles per block of k_mc_area<0>) followed by the result (1 double).  Every
// block writes its own partial and k_mc_final writes the result, so nothing in it needs to
// start at zero.  One copy per device, held until the process ends; calls that overlap on one
// device share it.
constexpr size_t kScratchDoubles = (size_t)1 << 17;   // 1 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

size_t mc_scratch_doubles() { return kScratchDoubles; }

// Device address of `scratch` on `device` (the current device): looked up with
// hipGetSymbolAddress on the first call for that device and kept.  Host code uses only this
// address; the array's name is not a device address in host code.
hipError_t mc_scratch_address(int device, double** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = (double*)p;
    }
    *out = addr[device];
    return hipSuccess;
}

// Grid of the area kernel for an nx x ny x nz volume.
dim3 mc_grid(int nx, int ny, int nz) {
    return dim3((unsigned)((nx + kTX - 1) / kTX), (unsigned)((ny + kTY - 1) / kTY), (unsigned)((nz + kTZ - 1) / kTZ));
}

// The grid's y and z extents stay within the launch limit (65535).
bool mc_fits(int nx, int ny, int nz) {
    const dim3 g = mc_grid(nx, ny, nz);
    return g.y <= 65535u && g.z <= 65535u;
}

size_t mc_blocks(int nx, int ny, int nz) {
    const dim3 g = mc_grid(nx, ny, nz);
    return (size_t)g.x * g.y * g.z;
}

// Areas of `src` (an nx x ny x nz volume): MODE 0 one surface of the bytes, MODE 1 six
// surfaces of face bits.  part: mc_blocks() * NS * 2 doubles; out: NS doubles (device).
hipError_t mc_area(const uint8_t* src, int nx, int ny, int nz, float iso, int faces, double* part, double* out,
                   hipStream_t s) {
    const dim3 grid = mc_grid(nx, ny, nz);
    const size_t nblk = mc_blocks(nx, ny, nz);
    // first count whose value is not below iso (the same float division as the kernels)
    const uint32_t maxc = faces ? 8u : 2040u;
    const float div = faces ? 8.0f : 2040.0f;
    uint32_t thr = 0;
    while (thr <= maxc && (float)thr / div < iso) ++thr;
    if (faces) {
        hipLaunchKernelGGL(k_mc_area<1>, grid, dim3(kNT), 0, s, src, nx, ny, nz, iso, thr, part);
        hipLaunchKernelGGL(k_mc_final, dim3(6), dim3(256), 0, s, part, nblk, 6, out);
    } else {
        hipLaunchKernelGGL(k_mc_area<0>, grid, dim3(kNT), 0, s, src, nx, ny, nz, iso, thr, part);
        hipLaunchKernelGGL(k_mc_final, dim3(1), dim3(256), 0, s, part, nblk, 1, out);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Triangle mesh for the OBJ export (`fast alpha --mesh`, AlphaModule.h:145-167): the volume
// padded with one zero voxel per side (Volume::withZeroPadding), marching cubes over the
// padded grid with output positions in a box centred at the origin, spacing 1/maxDim
// (getVolumeAreaMesh, VolumeSurface.cpp:39-74; ___generateTriangles, VolumeSurface.cu:260-357).
// Triangles are emitted in linear cube order, then table order, as the reference's compaction
// by exclusive scan orders them.
// ---------------------------------------------------------------------------------------
namespace {

__global__ __launch_bounds__(256) void k_pad(const uint8_t* __restrict__ src, int nx, int ny, int nz,
                                             uint8_t* __restrict__ dst) {
    const int px = nx + 2, py = ny + 2, pz = nz + 2;
    const long long n = (long long)px * py * pz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % px), y = (int)((i / px) % py), z = (int)(i / ((long long)px * py));
        const bool inside = x >= 1 && x <= nx && y >= 1 && y <= ny && z >= 1 && z <= nz;
        dst[i] = inside ? src[(size_t)(x - 1) + (size_t)nx * ((size_t)(y - 1) + (size_t)ny * (size_t)(z - 1))] : 0;
    }
}

// Byte sums of the 8 clamped voxels around each corner of cube (x, y, z) of an r0 x r1 x r2 volume.
__device__ __forceinline__ void corner_sums(const uint8_t* v, int r0, int r1, int r2, int x, int y, int z,
                                            uint32_t (&c)[8]) {
    const int dx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        uint32_t s = 0;
#pragma unroll
        for (int k = -1; k <= 0; ++k)
#pragma unroll
            for (int j = -1; j <= 0; ++j)
#pragma unroll
                for (int i = -1; i <= 0; ++i) {
                    const int a = min(max(x + dx[m] + i, 0), r0 - 1);
                    const int b = min(max(y + dy[m] + j, 0), r1 - 1);
                    const int cc = min(max(z + dz[m] + k, 0), r2 - 1);
                    s += v[(size_t)a + (size_t)r0 * ((size_t)b + (size_t)r1 * (size_t)cc)];
                }
        c[m] = s;
    }
}

__global__ __launch_bounds__(256) void k_mesh_count(const uint8_t* __restrict__ v, int r0, int r1, int r2,
                                                    uint32_t thr, uint32_t* __restrict__ count) {
    const long long n = (long long)r0 * r1 * r2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % r0), y = (int)((i / r0) % r1), z = (int)(i / ((long long)r0 * r1));
        uint32_t c[8];
        corner_sums(v, r0, r1, r2, x, y, z, c);
        unsigned in = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) in |= (c[m] < thr ? 1u : 0u) << m;
        count[i] = 3u * (uint32_t)(c_tri[in] >> 60);
    }
}

// out: 6 floats per vertex (position, then the triangle's normal), at 6 * offset[cube].
__global__ __launch_bounds__(256) void k_mesh_emit(const uint8_t* __restrict__ v, int r0, int r1, int r2, float iso,
                                                   const uint32_t* __restrict__ count,
                                                   const uint64_t* __restrict__ offset, float* __restrict__ out) {
    const long long n = (long long)r0 * r1 * r2;
    const float tvs = 1.0f / (float)max(r0, max(r1, r2));   // params.dx = dy = dz = 1 / maxDim
    const float hx = 0.5f * (tvs * (float)r0), hy = 0.5f * (tvs * (float)r1), hz = 0.5f * (tvs * (float)r2);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (count[i] == 0) continue;
        const int x = (int)(i % r0), y = (int)((i / r0) % r1), z = (int)(i / ((long long)r0 * r1));
        uint32_t c[8];
        corner_sums(v, r0, r1, r2, x, y, z, c);
        float f[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) f[m] = (float)c[m] / 2040.0f;
        unsigned ci = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) ci |= (f[m] < iso ? 1u : 0u) << m;
        CubeGeo g;
        g.X0 = tvs * (float)x - hx;
        g.X1 = g.X0 + tvs;
        g.Y0 = tvs * (float)y - hy;
        g.Y1 = g.Y0 + tvs;
        g.Z0 = tvs * (float)z - hz;
        g.Z1 = g.Z0 + tvs;
        const uint64_t w = c_tri[ci];
        const int ntri = min((int)(w >> 60), (int)(count[i] / 3u));   // never beyond the space counted
        float* o = out + 6 * (size_t)offset[i];
        for (int t = 0; t < ntri; ++t) {
            float p[3][3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                edge_vertex_switch((unsigned)(w >> (4 * (3 * t + q))) & 15u, f, g, iso, p[q][0], p[q][1], p[q][2]);
            const float ux = p[1][0] - p[0][0], uy = p[1][1] - p[0][1], uz = p[1][2] - p[0][2];
            const float vx = p[2][0] - p[0][0], vy = p[2][1] - p[0][1], vz = p[2][2] - p[0][2];
            const float nxv = uy * vz - uz * vy, nyv = uz * vx - ux * vz, nzv = ux * vy - uy * vx;
            const float inv = 1.0f / sqrtf(nxv * nxv + nyv * nyv + nzv * nzv);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                float* d = o + 6 * (3 * t + q);
                d[0] = p[q][0];
                d[1] = p[q][1];
                d[2] = p[q][2];
                d[3] = nxv * inv;
                d[4] = nyv * inv;
                d[5] = nzv * inv;
            }
        }
    }
}

inline unsigned mesh_grid(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

// Padded volume and per-cube vertex counts.  padded: (nx+2)(ny+2)(nz+2) bytes; count: as many
// uint32.  thr as for the area kernels.
hipError_t mesh_count(const uint8_t* src, int nx, int ny, int nz, float iso, uint8_t* padded, uint32_t* count,
                      hipStream_t s) {
    const long long n = (long long)(nx + 2) * (ny + 2) * (nz + 2);
    uint32_t thr = 0;
    while (thr <= 2040u && (float)thr / 2040.0f < iso) ++thr;
    hipLaunchKernelGGL(k_pad, dim3(mesh_grid(n)), dim3(256), 0, s, src, nx, ny, nz, padded);
    hipLaunchKernelGGL(k_mesh_count, dim3(mesh_grid(n)), dim3(256), 0, s, padded, nx + 2, ny + 2, nz + 2, thr, count);
    return hipGetLastError();
}

namespace {
struct Widen {
    __host__ __device__ uint64_t operator()(uint32_t v) const { return v; }
};
}  // namespace

// Exclusive prefix sum of n per-cube counts into 64-bit offsets (hipCUB, summing in 64 bits).
// Call with tmp = nullptr for *tmp_bytes.
hipError_t mesh_scan(const uint32_t* count, uint64_t* offset, int n, void* tmp, size_t* tmp_bytes, hipStream_t s) {
    hipcub::TransformInputIterator<uint64_t, Widen, const uint32_t*> wide(count, Widen());
    return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, wide, offset, n, s);
}

hipError_t mesh_emit(const uint8_t* padded, int nx, int ny, int nz, float iso, const uint32_t* count,
                     const uint64_t* offset, float* out, hipStream_t s) {
    const long long n = (long long)(nx + 2) * (ny + 2) * (nz + 2);
    hipLaunchKernelGGL(k_mesh_emit, dim3(mesh_grid(n)), dim3(256), 0, s, padded, nx + 2, ny + 2, nz + 2, iso, count,
                       offset, out);
    return hipGetLastError();
}

}  // namespace fta
