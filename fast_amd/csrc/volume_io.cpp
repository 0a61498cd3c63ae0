// Host-side volume utilities of libfasttau: the reference `.bin` format and the seeded
// synthetic-volume generators used for the benchmark workloads.
#include "fast_tau.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

// .bin layout (fastlib/src/volume/VolumeIO.cpp:256-305): int32 PrimitiveType, int32 dims[3],
// one byte (the `deprecated` flag), then the raw voxels.  TYPE_UCHAR = 2
// (fastlib/include/fastlib/utility/PrimitiveTypes.h:11-24).
constexpr int32_t kTypeUchar = 2;
constexpr size_t kHeader = 4 + 12 + 1;

// xoshiro256** seeded through splitmix64: portable, reproducible streams.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
};

// Fill the voxels whose centres satisfy d^T Q d <= 1 (d = centre - c); returns the number
// of voxels that turned solid.  Rows are filled from the exact x-interval of the quadric.
size_t fill_quadric(uint8_t* m, int nx, int ny, int nz, const double c[3], const double Q[3][3], const double ext[3]) {
    size_t added = 0;
    const int zlo = std::max(0, (int)std::ceil(c[2] - ext[2] - 0.5));
    const int zhi = std::min(nz - 1, (int)std::floor(c[2] + ext[2] - 0.5));
    const int ylo = std::max(0, (int)std::ceil(c[1] - ext[1] - 0.5));
    const int yhi = std::min(ny - 1, (int)std::floor(c[1] + ext[1] - 0.5));
    for (int z = zlo; z <= zhi; ++z) {
        const double dz = z + 0.5 - c[2];
        for (int y = ylo; y <= yhi; ++y) {
            const double dy = y + 0.5 - c[1];
            const double B = Q[0][1] * dy + Q[0][2] * dz;
            const double C = Q[1][1] * dy * dy + 2.0 * Q[1][2] * dy * dz + Q[2][2] * dz * dz - 1.0;
            const double disc = B * B - Q[0][0] * C;
            if (disc < 0) continue;
            const double sq = std::sqrt(disc);
            const double lo = c[0] + (-B - sq) / Q[0][0];
            const double hi = c[0] + (-B + sq) / Q[0][0];
            const int xlo = std::max(0, (int)std::ceil(lo - 0.5));
            const int xhi = std::min(nx - 1, (int)std::floor(hi - 0.5));
            uint8_t* row = m + ((size_t)z * ny + y) * nx;
            for (int x = xlo; x <= xhi; ++x) {
                added += (row[x] == 0);
                row[x] = 255;
            }
        }
    }
    return added;
}

}  // namespace

extern "C" {

int fast_load_bin_dims(const char* path, int* nx, int* ny, int* nz) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return FAST_E_IO;
    int32_t hdr[4];
    const bool ok = std::fread(hdr, 4, 4, f) == 4;
    std::fclose(f);
    if (!ok || hdr[0] != kTypeUchar || hdr[1] <= 0 || hdr[2] <= 0 || hdr[3] <= 0) return FAST_E_IO;
    *nx = hdr[1];
    *ny = hdr[2];
    *nz = hdr[3];
    return FAST_OK;
}

int fast_load_bin(const char* path, uint8_t* mask, size_t capacity) {
    int nx, ny, nz;
    int rc = fast_load_bin_dims(path, &nx, &ny, &nz);
    if (rc) return rc;
    const size_t n = (size_t)nx * ny * nz;
    if (!mask || capacity < n) return FAST_E_BADARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return FAST_E_IO;
    bool ok = std::fseek(f, (long)kHeader, SEEK_SET) == 0 && std::fread(mask, 1, n, f) == n;
    std::fclose(f);
    return ok ? FAST_OK : FAST_E_IO;
}

int fast_save_bin(const char* path, const uint8_t* mask, int nx, int ny, int nz) {
    if (!mask || nx <= 0 || ny <= 0 || nz <= 0) return FAST_E_BADARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return FAST_E_IO;
    const int32_t hdr[4] = {kTypeUchar, nx, ny, nz};
    const uint8_t flag = 0;
    const size_t n = (size_t)nx * ny * nz;
    bool ok = std::fwrite(hdr, 4, 4, f) == 4 && std::fwrite(&flag, 1, 1, f) == 1 && std::fwrite(mask, 1, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? FAST_OK : FAST_E_IO;
}

long fast_generate_spheres(uint8_t* mask, int nx, int ny, int nz, double radius, double solid_fraction,
                           uint64_t seed) {
    if (!mask || nx <= 0 || ny <= 0 || nz <= 0 || !(radius > 0) || solid_fraction < 0 || solid_fraction > 1)
        return -1;
    const size_t n = (size_t)nx * ny * nz;
    std::memset(mask, 0, n);
    Rng rng(seed);
    const double target = solid_fraction * (double)n;
    const double inv = 1.0 / (radius * radius);
    const double Q[3][3] = {{inv, 0, 0}, {0, inv, 0}, {0, 0, inv}};
    const double ext[3] = {radius, radius, radius};
    size_t solid = 0;
    long placed = 0;
    while ((double)solid < target) {
        const double c[3] = {rng.uniform() * nx, rng.uniform() * ny, rng.uniform() * nz};
        solid += fill_quadric(mask, nx, ny, nz, c, Q, ext);
        ++placed;
    }
    return placed;
}

long fast_generate_ellipsoids(uint8_t* mask, int nx, int ny, int nz, double a, double r, double solid_fraction,
                              uint64_t seed) {
    if (!mask || nx <= 0 || ny <= 0 || nz <= 0 || !(a > 0) || !(r > 0) || solid_fraction < 0 || solid_fraction > 1)
        return -1;
    const size_t n = (size_t)nx * ny * nz;
    std::memset(mask, 0, n);
    Rng rng(seed);
    const double target = solid_fraction * (double)n;
    const double s[3] = {a, a * r, a * r};
    const double kTwoPi = 6.283185307179586476925286766559;
    size_t solid = 0;
    long placed = 0;
    while ((double)solid < target) {
        const double c[3] = {rng.uniform() * nx, rng.uniform() * ny, rng.uniform() * nz};
        // uniform rotation: random unit quaternion (Shoemake 1992)
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double q0 = std::sqrt(1 - u1) * std::sin(kTwoPi * u2), q1 = std::sqrt(1 - u1) * std::cos(kTwoPi * u2);
        const double q2 = std::sqrt(u1) * std::sin(kTwoPi * u3), q3 = std::sqrt(u1) * std::cos(kTwoPi * u3);
        // rotation matrix of quaternion (w = q3, v = (q0, q1, q2))
        const double w = q3, xq = q0, yq = q1, zq = q2;
        const double R[3][3] = {
            {1 - 2 * (yq * yq + zq * zq), 2 * (xq * yq - zq * w), 2 * (xq * zq + yq * w)},
            {2 * (xq * yq + zq * w), 1 - 2 * (xq * xq + zq * zq), 2 * (yq * zq - xq * w)},
            {2 * (xq * zq - yq * w), 2 * (yq * zq + xq * w), 1 - 2 * (xq * xq + yq * yq)},
        };
        // Q = R diag(1/s^2) R^T ; bounding half-extent along i = sqrt(sum_k R_ik^2 s_k^2)
        double Q[3][3], ext[3];
        for (int i = 0; i < 3; ++i) {
            double e = 0;
            for (int k = 0; k < 3; ++k) e += R[i][k] * R[i][k] * s[k] * s[k];
            ext[i] = std::sqrt(e);
            for (int j = 0; j < 3; ++j) {
                double q = 0;
                for (int k = 0; k < 3; ++k) q += R[i][k] * R[j][k] / (s[k] * s[k]);
                Q[i][j] = q;
            }
        }
        solid += fill_quadric(mask, nx, ny, nz, c, Q, ext);
        ++placed;
    }
    return placed;
}

}  // extern "C"
