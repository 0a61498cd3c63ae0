// Host-side volume utilities of libfasttau: the reference's input formats (`.bin`, TIFF
// folders, `.sph` sphere lists), binarization and subvolumes, and the seeded
// synthetic-volume generators used for the benchmark workloads.
#include "fast_tau.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <new>
#include <string>
#include <sys/stat.h>
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

// ---- TIFF (baseline, uncompressed, 8-bit grayscale strips) ---------------------------------
// The reference reads slices with TinyTIFF (fastlib/src/volume/VolumeIO.cpp:33-77): the first
// image of each file, raw sample data.  This reader takes the same images and refuses what it
// cannot read byte for byte as uint8: compression, more than one sample per pixel, sample
// sizes other than 8 bits, tiles, BigTIFF.
struct TiffInfo {
    uint32_t w = 0, h = 0;
    std::vector<uint64_t> off, cnt;   // strips
};

bool read_file(const std::string& path, std::vector<uint8_t>& buf) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n < 8) return false;
    buf.resize((size_t)n);
    f.seekg(0);
    return (bool)f.read((char*)buf.data(), n);
}

bool tiff_parse(const std::vector<uint8_t>& b, TiffInfo& t) {
    const bool be = b[0] == 'M' && b[1] == 'M';
    if (!be && !(b[0] == 'I' && b[1] == 'I')) return false;
    auto u16 = [&](size_t o) -> uint32_t {
        if (o + 2 > b.size()) return 0;
        return be ? (uint32_t)(b[o] << 8 | b[o + 1]) : (uint32_t)(b[o + 1] << 8 | b[o]);
    };
    auto u32 = [&](size_t o) -> uint32_t {
        if (o + 4 > b.size()) return 0;
        return be ? ((uint32_t)b[o] << 24 | (uint32_t)b[o + 1] << 16 | (uint32_t)b[o + 2] << 8 | b[o + 3])
                  : ((uint32_t)b[o + 3] << 24 | (uint32_t)b[o + 2] << 16 | (uint32_t)b[o + 1] << 8 | b[o]);
    };
    if (u16(2) != 42) return false;   // 43 = BigTIFF: not read
    const size_t ifd = u32(4);
    const uint32_t ne = u16(ifd);
    if (ifd + 2 + 12ull * ne > b.size()) return false;
    uint32_t bps = 8, spp = 1, comp = 1, planar = 1, fmt = 1, rps = 0xFFFFFFFFu;
    bool tiled = false;
    // values of an entry: SHORT (3) or LONG (4), inline when they fit in 4 bytes
    auto values = [&](size_t e, std::vector<uint64_t>& out) -> bool {
        const uint32_t type = u16(e + 2), count = u32(e + 4);
        const uint32_t sz = type == 3 ? 2 : (type == 4 ? 4 : 0);
        if (!sz || count == 0 || count > (1u << 28)) return false;
        const size_t base = (size_t)sz * count <= 4 ? e + 8 : u32(e + 8);
        if (base + (size_t)sz * count > b.size()) return false;
        out.resize(count);
        for (uint32_t i = 0; i < count; ++i) out[i] = sz == 2 ? u16(base + 2 * i) : u32(base + 4 * i);
        return true;
    };
    for (uint32_t k = 0; k < ne; ++k) {
        const size_t e = ifd + 2 + 12ull * k;
        const uint32_t tag = u16(e);
        std::vector<uint64_t> v;
        switch (tag) {
            case 256: if (!values(e, v)) return false; t.w = (uint32_t)v[0]; break;
            case 257: if (!values(e, v)) return false; t.h = (uint32_t)v[0]; break;
            case 258: if (!values(e, v)) return false; bps = (uint32_t)v[0]; break;
            case 259: if (!values(e, v)) return false; comp = (uint32_t)v[0]; break;
            case 273: if (!values(e, t.off)) return false; break;
            case 277: if (!values(e, v)) return false; spp = (uint32_t)v[0]; break;
            case 278: if (!values(e, v)) return false; rps = (uint32_t)v[0]; break;
            case 279: if (!values(e, t.cnt)) return false; break;
            case 284: if (!values(e, v)) return false; planar = (uint32_t)v[0]; break;
            case 339: if (!values(e, v)) return false; fmt = (uint32_t)v[0]; break;
            case 322: case 323: case 324: case 325: tiled = true; break;
            default: break;
        }
    }
    (void)planar;
    if (tiled || comp != 1 || spp != 1 || bps != 8 || fmt != 1 || t.w == 0 || t.h == 0) return false;
    if (t.off.empty() || t.off.size() != t.cnt.size()) return false;
    (void)rps;
    return true;
}

// One slice: the strips in order, w * h bytes.
bool tiff_read(const std::string& path, uint32_t* w, uint32_t* h, uint8_t* dst, size_t cap) {
    std::vector<uint8_t> b;
    TiffInfo t;
    if (!read_file(path, b) || !tiff_parse(b, t)) return false;
    *w = t.w;
    *h = t.h;
    if (!dst) return true;
    const size_t need = (size_t)t.w * t.h;
    if (cap < need) return false;
    size_t pos = 0;
    for (size_t i = 0; i < t.off.size() && pos < need; ++i) {
        const size_t n = std::min<size_t>((size_t)t.cnt[i], need - pos);
        if (t.off[i] + n > b.size()) return false;
        std::memcpy(dst + pos, b.data() + t.off[i], n);
        pos += n;
    }
    return pos == need;
}

bool is_dir(const char* p) {
    struct stat st;
    return p && stat(p, &st) == 0 && S_ISDIR(st.st_mode);
}

// listDir + checkExtension (VolumeIO.cpp:79-111,307-310): names sorted, extension after the
// last '.' equal to "tiff" or "tif" (case-sensitive), directories skipped.
std::vector<std::string> tiff_slices(const char* folder) {
    std::vector<std::string> out;
    DIR* d = opendir(folder);
    if (!d) return out;
    while (dirent* e = readdir(d)) {
        const std::string name = e->d_name;
        const size_t dot = name.find_last_of('.');
        const std::string ext = dot == std::string::npos ? name : name.substr(dot + 1);
        if (ext != "tiff" && ext != "tif") continue;
        const std::string full = std::string(folder) + "/" + name;
        if (is_dir(full.c_str())) continue;
        out.push_back(name);
    }
    closedir(d);
    std::sort(out.begin(), out.end());
    return out;
}

}  // namespace

extern "C" {

static int tiff_folder_dims(const char* folder, int* nx, int* ny, int* nz) {
    if (!folder || !nx || !ny || !nz) return FAST_E_BADARG;
    if (!is_dir(folder)) return FAST_E_IO;
    const std::vector<std::string> sl = tiff_slices(folder);
    if (sl.empty()) return FAST_E_IO;
    uint32_t w = 0, h = 0;
    if (!tiff_read(std::string(folder) + "/" + sl[0], &w, &h, nullptr, 0)) return FAST_E_IO;
    if ((uint64_t)w * h * sl.size() > ((uint64_t)1 << 40) || w > 0x7FFFFFFF || h > 0x7FFFFFFF) return FAST_E_IO;
    *nx = (int)w;
    *ny = (int)h;
    *nz = (int)sl.size();
    return FAST_OK;
}

static int load_tiff_folder(const char* folder, uint8_t* mask, size_t capacity) {
    int nx, ny, nz;
    int rc = tiff_folder_dims(folder, &nx, &ny, &nz);
    if (rc) return rc;
    const size_t plane = (size_t)nx * ny;
    if (!mask || capacity < plane * nz) return FAST_E_BADARG;
    const std::vector<std::string> sl = tiff_slices(folder);
    for (size_t k = 0; k < sl.size(); ++k) {
        uint32_t w, h;
        if (!tiff_read(std::string(folder) + "/" + sl[k], &w, &h, mask + k * plane, plane)) return FAST_E_IO;
        if ((int)w != nx || (int)h != ny) return FAST_E_IO;   // slices of another size
    }
    return FAST_OK;
}

static int rasterize_sph(const char* path, int nx, int ny, int nz, uint8_t* mask);

// The readers allocate host buffers: a failed allocation is FAST_E_OOM, never an exception
// through the C ABI.
int fast_tiff_folder_dims(const char* folder, int* nx, int* ny, int* nz) {
    try {
        return tiff_folder_dims(folder, nx, ny, nz);
    } catch (const std::bad_alloc&) {
        return FAST_E_OOM;
    }
}

int fast_load_tiff_folder(const char* folder, uint8_t* mask, size_t capacity) {
    try {
        return load_tiff_folder(folder, mask, capacity);
    } catch (const std::bad_alloc&) {
        return FAST_E_OOM;
    }
}

int fast_rasterize_sph(const char* path, int nx, int ny, int nz, uint8_t* mask) {
    try {
        return rasterize_sph(path, nx, ny, nz, mask);
    } catch (const std::bad_alloc&) {
        return FAST_E_OOM;
    }
}

void fast_binarize(uint8_t* mask, size_t n, float threshold) {
    if (!mask) return;
    const uint8_t t = (uint8_t)(threshold * 255);   // uchar(threshold * 255), Volume.cu:180
    for (size_t i = 0; i < n; ++i) mask[i] = mask[i] < t ? 0 : 255;
}

static int rasterize_sph(const char* path, int nx, int ny, int nz, uint8_t* mask) {
    if (!path || !mask || nx <= 0 || ny <= 0 || nz <= 0) return FAST_E_BADARG;
    std::ifstream f(path);
    if (!f) return FAST_E_IO;
    long count = 0;
    if (!(f >> count) || count < 0) return FAST_E_IO;
    struct Sph {
        float x, y, z, r, r2;
    };
    std::vector<Sph> sp;   // grown as spheres are read: a bad count cannot reserve memory
    for (long i = 0; i < count; ++i) {
        Sph s;
        if (!(f >> s.x >> s.y >> s.z >> s.r)) return FAST_E_IO;
        s.r2 = s.r * s.r;
        sp.push_back(s);
    }
    std::memset(mask, 0, (size_t)nx * ny * nz);
    const int res[3] = {nx, ny, nz};
    for (const Sph& s : sp) {
        const float c[3] = {s.x, s.y, s.z};
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            // normToDiscrete<true>: int((dim - 1) * p), then clamped to [0, dim - 1]
            lo[a] = std::min(std::max((int)((float)(res[a] - 1) * (c[a] - s.r)), 0), res[a] - 1);
            hi[a] = std::min(std::max((int)((float)(res[a] - 1) * (c[a] + s.r)), 0), res[a] - 1);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    // discreteToNorm: float(i) / dim; glm::length2 summed x, y, z
                    const float dx = (float)x / (float)nx - s.x;
                    const float dy = (float)y / (float)ny - s.y;
                    const float dz = (float)z / (float)nz - s.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < s.r2) mask[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] = 255;
                }
    }
    return FAST_OK;
}

int fast_crop(const uint8_t* src, int nx, int ny, int nz, int ox, int oy, int oz, int sx, int sy, int sz, uint8_t* dst) {
    if (!src || !dst || nx <= 0 || ny <= 0 || nz <= 0 || sx <= 0 || sy <= 0 || sz <= 0 || ox < 0 || oy < 0 || oz < 0 ||
        ox + sx > nx || oy + sy > ny || oz + sz > nz)
        return FAST_E_BADARG;
    for (int z = 0; z < sz; ++z)
        for (int y = 0; y < sy; ++y)
            std::memcpy(dst + ((size_t)z * sy + y) * sx, src + (((size_t)(z + oz) * ny + (y + oy)) * nx + ox), (size_t)sx);
    return FAST_OK;
}


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
