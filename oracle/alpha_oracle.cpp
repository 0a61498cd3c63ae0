/*
 * alpha_oracle.cpp -- CPU restatement of the reference's `fast alpha` area measure.
 *
 * TEST INFRASTRUCTURE ONLY, on the same terms as tau_oracle.cpp: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * never links it.
 *
 * What it restates (citations relative to the reference tree):
 *   - texture read     fastlib/src/utility/DataPtr.cpp:565-592 (normalized coordinates,
 *                      clamp addressing, linear filtering, uchar read as byte/255) sampled
 *                      at voxel corners i/N (fastlib/src/cuda/VolumeSurface.cu:83-103,
 *                      219-224).  A corner lies on the texel-centre midpoint in every axis,
 *                      so it reads the mean of its 8 clamped voxels: sum / 2040.  Computed
 *                      here straight from the 8 voxels of every corner (no plane sums).
 *   - cube index       VolumeSurface.cu:116-141 (bit m set when corner m < isovalue)
 *   - vertices         vertexInterp VolumeSurface.cu:171-178, lerp CudaMath.h:1134-1137
 *   - triangle area    VolumeSurface.cu:187-189, cross/dot/length CudaMath.h:1416,1244,1287
 *   - per-cube sum     VolumeSurface.cu:240-255 (double, table order)
 *   - volume sum       Volume::sum (order unspecified): here double-double, rounded once
 *   - mesh (--mesh)    AlphaModule.h:145-167, getVolumeAreaMesh (VolumeSurface.cpp:39-74),
 *                      ___generateTriangles (VolumeSurface.cu:260-357); the triangle normal
 *                      is n / sqrt(n.n) where the reference uses CUDA's approximate rsqrtf,
 *                      so normals agree with the reference to a few ulp only
 *   - tensor           getReactiveAreaDensityTensor, VolumeMeasures.cpp:303-333: face f's
 *                      volume is 255 on pore voxels whose label is on face f
 *                      (generateBoundaryConnectedVolume, VolumeSegmentation.cpp:66-90), from
 *                      this oracle's own CCL (oracle_ccl, oracle_boundary_table)
 *
 * Float arithmetic is evaluated as written, each operation rounded (-ffp-contract=off).
 *
 * Pins (tests/test_alpha_oracle.py): the triangle table against mesh invariants (every case
 * cuts exactly its sign-change edges; the mesh of any zero-padded volume is closed and
 * consistently oriented) and, where the reference tree is present, entry by entry against its
 * table; closed forms (half-spaces and slabs give planes of area 1 per plane, a solid box
 * gives the chamfered box of the formula in the test, per-face tensor values of split
 * volumes); the sphere area against 4 pi r^2.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../fast_amd/csrc/mc_table.h"

namespace {

const uint64_t kTri[256] = FAST_MC_TRI_WORDS;
const int kEdgeA[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
const int kEdgeB[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct DDs {
    double hi = 0.0, lo = 0.0;
    void add2(double p, double e) {
        const double s = hi + p;
        const double v = s - hi;
        double t = (hi - (s - v)) + (p - v);
        t += lo + e;
        hi = s + t;
        lo = t - (hi - s);
    }
};

struct F3 {
    float x, y, z;
};

F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// p0 + t (p1 - p0), componentwise (CudaMath.h lerp)
F3 vertex_interp(float iso, F3 p0, F3 p1, float f0, float f1) {
    if (f1 == f0) return p0;
    const float t = (iso - f0) / (f1 - f0);
    return {p0.x + t * (p1.x - p0.x), p0.y + t * (p1.y - p0.y), p0.z + t * (p1.z - p0.z)};
}

float triangle_area(F3 v0, F3 v1, F3 v2) {
    const F3 a = sub(v1, v0), b = sub(v2, v0);
    const F3 c = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z) * 0.5f;
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Texture value at lattice point (i, j, k): mean of the 8 clamped voxels around it.
float corner_value(const uint8_t* vol, int nx, int ny, int nz, int i, int j, int k) {
    unsigned sum = 0;
    for (int c = k - 1; c <= k; ++c)
        for (int b = j - 1; b <= j; ++b)
            for (int a = i - 1; a <= i; ++a) {
                const size_t idx = (size_t)clampi(a, 0, nx - 1) +
                                   (size_t)nx * ((size_t)clampi(b, 0, ny - 1) + (size_t)ny * (size_t)clampi(c, 0, nz - 1));
                sum += vol[idx];
            }
    return (float)sum / 2040.0f;
}

double cube_area(const uint8_t* vol, int nx, int ny, int nz, int x, int y, int z, float iso) {
    const float vsx = 1.0f / (float)nx, vsy = 1.0f / (float)ny, vsz = 1.0f / (float)nz;
    const F3 pos = {vsx * (float)x, vsy * (float)y, vsz * (float)z};
    float f[8];
    F3 p[8];
    unsigned ci = 0;
    for (int m = 0; m < 8; ++m) {
        f[m] = corner_value(vol, nx, ny, nz, x + kCorner[m][0], y + kCorner[m][1], z + kCorner[m][2]);
        // pos + make_float3(dx ? vs.x : 0, dy ? vs.y : 0, dz ? vs.z : 0)
        p[m] = {pos.x + (kCorner[m][0] ? vsx : 0.0f), pos.y + (kCorner[m][1] ? vsy : 0.0f),
                pos.z + (kCorner[m][2] ? vsz : 0.0f)};
        if (f[m] < iso) ci |= 1u << m;
    }
    const uint64_t w = kTri[ci];
    const int ntri = (int)(w >> 60);
    double area = 0.0;
    for (int t = 0; t < ntri; ++t) {
        F3 v[3];
        for (int q = 0; q < 3; ++q) {
            const int e = (int)((w >> (4 * (3 * t + q))) & 15u);
            v[q] = vertex_interp(iso, p[kEdgeA[e]], p[kEdgeB[e]], f[kEdgeA[e]], f[kEdgeB[e]]);
        }
        area += (double)triangle_area(v[0], v[1], v[2]);
    }
    return area;
}

}  // namespace

extern "C" {

uint32_t oracle_ccl(const uint8_t* mask, int nx, int ny, int nz, uint8_t background, uint32_t* labels);
void oracle_boundary_table(const uint32_t* labels, int nx, int ny, int nz, uint32_t numLabels, uint8_t* table);

/* The triangle table (256 words, encoding of fast_amd/csrc/mc_table.h). */
void oracle_mc_table(uint64_t* out) { std::memcpy(out, kTri, sizeof(kTri)); }

/* Marching-cubes area of one cube (x, y, z) of `vol` (for the tests). */
double oracle_mc_cube_area(const uint8_t* vol, int nx, int ny, int nz, int x, int y, int z, float iso) {
    return cube_area(vol, nx, ny, nz, x, y, z, iso);
}

/* getReactiveAreaDensity: total area of the volume (double-double sum of the per-cube areas). */
double oracle_mc_area(const uint8_t* vol, int nx, int ny, int nz, float iso) {
    const int nt = omp_get_max_threads();
    std::vector<DDs> part(nt);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        DDs acc;
        for (int z = t; z < nz; z += nt)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) acc.add2(cube_area(vol, nx, ny, nz, x, y, z, iso), 0.0);
        part[t] = acc;
    }
    DDs total;
    for (const DDs& p : part) total.add2(p.hi, p.lo);
    return total.hi + total.lo;
}

/* The `--mesh` triangles (AlphaModule.h:145-167; getVolumeAreaMesh, VolumeSurface.cpp:39-74;
 * ___generateTriangles, VolumeSurface.cu:260-357): the volume padded with one zero voxel per
 * side, cubes in linear order, output positions targetPos = (1/maxDim) * ivox - 0.5 * (1/maxDim)
 * * res plus the corner offsets, three vertices per triangle with the triangle's normal
 * normalize(cross(v1 - v0, v2 - v0)), here n * (1 / sqrt(n.n)).  Writes up to `cap` vertices
 * (6 floats each) and returns the vertex count. */
size_t oracle_mc_mesh(const uint8_t* vol, int nx, int ny, int nz, float iso, float* out, size_t cap) {
    const int rx = nx + 2, ry = ny + 2, rz = nz + 2;
    std::vector<uint8_t> pad((size_t)rx * ry * rz, 0);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                pad[(size_t)(x + 1) + (size_t)rx * ((size_t)(y + 1) + (size_t)ry * (size_t)(z + 1))] =
                    vol[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z)];
    const int maxDim = std::max(rx, std::max(ry, rz));
    const float tvs = 1.0f / (float)maxDim;
    const F3 half = {0.5f * (tvs * (float)rx), 0.5f * (tvs * (float)ry), 0.5f * (tvs * (float)rz)};
    size_t nv = 0;
    for (int z = 0; z < rz; ++z)
        for (int y = 0; y < ry; ++y)
            for (int x = 0; x < rx; ++x) {
                float f[8];
                F3 p[8];
                unsigned ci = 0;
                const F3 tp = sub({tvs * (float)x, tvs * (float)y, tvs * (float)z}, half);
                for (int m = 0; m < 8; ++m) {
                    f[m] = corner_value(pad.data(), rx, ry, rz, x + kCorner[m][0], y + kCorner[m][1], z + kCorner[m][2]);
                    p[m] = {tp.x + (kCorner[m][0] ? tvs : 0.0f), tp.y + (kCorner[m][1] ? tvs : 0.0f),
                            tp.z + (kCorner[m][2] ? tvs : 0.0f)};
                    if (f[m] < iso) ci |= 1u << m;
                }
                const uint64_t w = kTri[ci];
                const int ntri = (int)(w >> 60);
                for (int t = 0; t < ntri; ++t) {
                    F3 v[3];
                    for (int q = 0; q < 3; ++q) {
                        const int e = (int)((w >> (4 * (3 * t + q))) & 15u);
                        v[q] = vertex_interp(iso, p[kEdgeA[e]], p[kEdgeB[e]], f[kEdgeA[e]], f[kEdgeB[e]]);
                    }
                    const F3 a = sub(v[1], v[0]), b = sub(v[2], v[0]);
                    const F3 c = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
                    const float inv = 1.0f / std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
                    for (int q = 0; q < 3; ++q) {
                        if (nv < cap) {
                            float* d = out + 6 * nv;
                            d[0] = v[q].x;
                            d[1] = v[q].y;
                            d[2] = v[q].z;
                            d[3] = c.x * inv;
                            d[4] = c.y * inv;
                            d[5] = c.z * inv;
                        }
                        ++nv;
                    }
                }
            }
    return nv;
}

/* getReactiveAreaDensityTensor over getVolumeCCL(mask, 255): out[f] for the six faces. */
void oracle_mc_area_tensor(const uint8_t* mask, int nx, int ny, int nz, float iso, double* out) {
    const size_t n = (size_t)nx * ny * nz;
    std::vector<uint32_t> lab(n);
    const uint32_t nl = oracle_ccl(mask, nx, ny, nz, 255, lab.data());
    std::vector<uint8_t> table((size_t)7 * nl);
    oracle_boundary_table(lab.data(), nx, ny, nz, nl, table.data());
    std::vector<uint8_t> bvol(n);
    for (int f = 0; f < 6; ++f) {
        for (size_t i = 0; i < n; ++i) bvol[i] = table[(size_t)f * nl + lab[i]] ? 255 : 0;
        out[f] = oracle_mc_area(bvol.data(), nx, ny, nz, iso);
    }
}

}  // extern "C"
