// Reactive area density (`fast alpha`): C ABI over the marching-cubes area kernels
// (alpha_kernels.hip) and the CCL (ccl_api.cpp).
// Reference: getReactiveAreaDensity, getReactiveAreaDensityTensor
// (fastlib/src/volume/VolumeMeasures.cpp:294-333), getVolumeArea
// (fastlib/src/volume/VolumeSurface.cpp:9-36), generateBoundaryConnectedVolume
// (fastlib/src/volume/VolumeSegmentation.cpp:66-90), AlphaModule::getAlpha (fast/src/AlphaModule.h:28-174).
#include <cmath>
#include <cstring>
#include <fstream>
#include <unordered_map>
#include <vector>

#include "fast_tau.h"
#include "tau_host.h"

namespace fta {
bool mc_fits(int nx, int ny, int nz);
size_t mc_blocks(int nx, int ny, int nz);
hipError_t mc_area(const uint8_t* src, int nx, int ny, int nz, float iso, int faces, double* part, double* out,
                   hipStream_t s);
size_t mc_scratch_doubles();
hipError_t mesh_count(const uint8_t* src, int nx, int ny, int nz, float iso, uint8_t* padded, uint32_t* count,
                      hipStream_t s);
hipError_t mesh_scan(const uint32_t* count, uint64_t* offset, int n, void* tmp, size_t* tmp_bytes, hipStream_t s);
hipError_t mesh_emit(const uint8_t* padded, int nx, int ny, int nz, float iso, const uint32_t* count,
                     const uint64_t* offset, float* out, hipStream_t s);
}  // namespace fta
namespace ftc {
hipError_t ccl_face_codes(const uint32_t* L, long long n, const uint8_t* table, uint32_t numLabels, uint8_t* out,
                          hipStream_t s);
}  // namespace ftc

using namespace fth;

namespace {

// Areas of `src` on the device: one surface (faces = 0) or six (faces = 1) into out[].
int run_area(fast_volume* v, const uint8_t* src, float iso, int faces, double* out) {
    fast_ctx* ctx = v->ctx;
    const int ns = faces ? 6 : 1;
    if (!fta::mc_fits(v->nx, v->ny, v->nz)) {
        set_error(ctx, "volume %dx%dx%d exceeds the area kernel's launch grid", v->nx, v->ny, v->nz);
        return FAST_E_BADARG;
    }
    const size_t nblk = fta::mc_blocks(v->nx, v->ny, v->nz);
    double* d = (double*)dalloc(ctx, (nblk * ns * 2 + ns) * sizeof(double));
    if (!d) return FAST_E_OOM;
    double* part = d;
    double* d_out = d + nblk * ns * 2;
    hipError_t e = fta::mc_area(src, v->nx, v->ny, v->nz, iso, faces, part, d_out, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, ns * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, d);
    return hip_ok(ctx, e, "marching-cubes area") ? FAST_OK : FAST_E_HIP;
}

// Face codes of the volume's cached labels (background 255) into a new device buffer.
int face_codes(fast_volume* v, uint8_t** out) {
    fast_ctx* ctx = v->ctx;
    uint32_t nl = 0;
    int rc = ensure_labels(v, 255, &nl);   // getVolumeCCL(volume, 255): solid is the background
    if (rc) return rc;
    uint8_t* table = nullptr;
    rc = face_table(v, nl, &table);
    if (rc) return rc;
    const long long n = (long long)v->nx * v->ny * v->nz;
    uint8_t* codes = (uint8_t*)dalloc(ctx, (size_t)n);
    if (!codes) {
        dfree(ctx, table);
        return FAST_E_OOM;
    }
    hipError_t e = ftc::ccl_face_codes(v->d_labels, n, table, nl, codes, ctx->stream);
    dfree(ctx, table);   // stream-ordered reuse only
    if (!hip_ok(ctx, e, "face codes")) {
        dfree(ctx, codes);
        return FAST_E_HIP;
    }
    *out = codes;
    return FAST_OK;
}

// Mesh of the zero-padded volume: the vertex count, and when `host` is non-null (capacity in
// vertices), the vertices (6 floats each).  Positions and cube order as getVolumeAreaMesh.
int run_mesh(fast_volume* v, float iso, std::vector<float>* host, size_t capacity, float* out, size_t* nverts) {
    fast_ctx* ctx = v->ctx;
    const long long n = (long long)(v->nx + 2) * (v->ny + 2) * (v->nz + 2);
    if (n > 0x7FFFFFFFLL) {   // the scan counts cubes in an int
        set_error(ctx, "volume %dx%dx%d is too large for the mesh export (2^31 padded cubes or more)", v->nx,
                  v->ny, v->nz);
        return FAST_E_BADARG;
    }
    size_t tmp_bytes = 0;
    hipError_t e = fta::mesh_scan(nullptr, nullptr, (int)n, nullptr, &tmp_bytes, ctx->stream);
    if (!hip_ok(ctx, e, "mesh scan size")) return FAST_E_HIP;
    uint8_t* padded = (uint8_t*)dalloc(ctx, (size_t)n);
    uint32_t* count = (uint32_t*)dalloc(ctx, (size_t)n * 4);
    uint64_t* offset = (uint64_t*)dalloc(ctx, (size_t)n * 8);
    void* tmp = dalloc(ctx, tmp_bytes);
    float* d_out = nullptr;
    int rc = (padded && count && offset && tmp) ? FAST_OK : FAST_E_OOM;
    uint64_t last_off = 0;
    uint32_t last_count = 0;
    if (!rc) {
        e = fta::mesh_count(v->d_mask, v->nx, v->ny, v->nz, iso, padded, count, ctx->stream);
        if (e == hipSuccess) e = fta::mesh_scan(count, offset, (int)n, tmp, &tmp_bytes, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&last_off, offset + n - 1, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&last_count, count + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (!hip_ok(ctx, e, "mesh count")) rc = FAST_E_HIP;
    }
    const size_t total = (size_t)(last_off + last_count);
    // the reference numbers vertices with unsigned ints (VolumeSurface.cu:260-357)
    if (!rc && total >= ((size_t)1 << 32)) {
        set_error(ctx, "mesh of %zu vertices exceeds the 2^32 - 1 vertices of the reference's OBJ export", total);
        rc = FAST_E_BADARG;
    }
    if (!rc) *nverts = total;
    const bool want = !rc && total > 0 && (host || (out && capacity >= total));
    if (want) {
        d_out = (float*)dalloc(ctx, total * 6 * sizeof(float));
        if (!d_out) rc = FAST_E_OOM;
        float* dst = out;
        if (!rc && host) {
            try {
                host->resize(total * 6);
            } catch (const std::bad_alloc&) {
                rc = FAST_E_OOM;
            }
            dst = host->data();
        }
        if (!rc) {
            e = fta::mesh_emit(padded, v->nx, v->ny, v->nz, iso, count, offset, d_out, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(dst, d_out, total * 6 * sizeof(float), hipMemcpyDeviceToHost,
                                                    ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (!hip_ok(ctx, e, "mesh emit")) rc = FAST_E_HIP;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx, d_out);
    dfree(ctx, tmp);
    dfree(ctx, offset);
    dfree(ctx, count);
    dfree(ctx, padded);
    return rc;
}

// CUDA_VBO::saveObj (DataPtr.cpp:684-740): vertices merged by equal position in order of first
// appearance, their triangle normals summed and normalised (n / sqrt(n.n)), then `v`, `vn` and
// `f` lines with default stream formatting.
struct PosKey {
    float x, y, z;
    bool operator==(const PosKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct PosHash {
    size_t operator()(const PosKey& k) const {
        auto bits = [](float f) {
            if (f == 0.0f) f = 0.0f;   // -0 and +0 are the same position
            uint32_t u;
            std::memcpy(&u, &f, 4);
            return (size_t)u;
        };
        size_t h = bits(k.x);
        h = h * 0x9E3779B97F4A7C15ull + bits(k.y);
        h = h * 0x9E3779B97F4A7C15ull + bits(k.z);
        return h;
    }
};

bool write_obj(const std::vector<float>& vtx, size_t nverts, const char* path) {
    std::vector<float> pos, nrm;
    std::vector<size_t> idx(nverts);
    std::unordered_map<PosKey, size_t, PosHash> seen;
    for (size_t i = 0; i < nverts; ++i) {
        const float* d = &vtx[6 * i];
        const PosKey k{d[0], d[1], d[2]};
        auto it = seen.find(k);
        if (it == seen.end()) {
            const size_t id = pos.size() / 3;
            seen.emplace(k, id);
            pos.insert(pos.end(), {d[0], d[1], d[2]});
            nrm.insert(nrm.end(), {d[3], d[4], d[5]});
            idx[i] = id;
        } else {
            const size_t id = it->second;
            nrm[3 * id] += d[3];
            nrm[3 * id + 1] += d[4];
            nrm[3 * id + 2] += d[5];
            idx[i] = id;
        }
    }
    std::ofstream f(path);
    if (!f.good()) return false;
    const size_t nv = pos.size() / 3;
    for (size_t v = 0; v < nv; ++v) f << "v " << pos[3 * v] << " " << pos[3 * v + 1] << " " << pos[3 * v + 2] << "\n";
    for (size_t v = 0; v < nv; ++v) {
        const float x = nrm[3 * v], y = nrm[3 * v + 1], z = nrm[3 * v + 2];
        const float inv = 1.0f / std::sqrt(x * x + y * y + z * z);
        f << "vn " << x * inv << " " << y * inv << " " << z * inv << "\n";
    }
    for (size_t t = 0; t + 2 < nverts; t += 3)
        f << "f " << idx[t] + 1 << " " << idx[t + 1] + 1 << " " << idx[t + 2] + 1 << "\n";
    return f.good();
}

bool iso_ok(fast_ctx* ctx, float iso) {
    if (std::isfinite(iso)) return true;
    set_error(ctx, "isovalue must be finite");
    return false;
}

}  // namespace

extern "C" {

int fast_alpha_area(fast_volume* v, float isovalue, double* area) {
    if (!v || !area) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!iso_ok(ctx, isovalue)) return FAST_E_BADARG;
    return run_area(v, v->d_mask, isovalue, 0, area);
}

int fast_alpha_area_noalloc(fast_volume* v, float isovalue, double* area) {
    if (!v || !area) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!iso_ok(ctx, isovalue)) return FAST_E_BADARG;
    const size_t cap = fta::mc_scratch_doubles();
    if (!fta::mc_fits(v->nx, v->ny, v->nz) || fta::mc_blocks(v->nx, v->ny, v->nz) * 2 + 1 > cap) {
        set_error(ctx, "volume %dx%dx%d needs more than the %zu area partials of the static scratch "
                  "(at most %zu blocks of 64x4x16 voxels)", v->nx, v->ny, v->nz, (cap - 1) / 2, (cap - 1) / 2);
        return FAST_E_BADARG;
    }
    // block partials, then the result, in this device's static scratch (set up with the context)
    double* part = ctx->alpha_scratch;
    double* d_out = part + fta::mc_blocks(v->nx, v->ny, v->nz) * 2;
    hipError_t e = fta::mc_area(v->d_mask, v->nx, v->ny, v->nz, isovalue, 0, part, d_out, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(area, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return hip_ok(ctx, e, "marching-cubes area (static scratch)") ? FAST_OK : FAST_E_HIP;
}

int fast_alpha_area_tensor(fast_volume* v, float isovalue, double area[6]) {
    if (!v || !area) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!iso_ok(ctx, isovalue)) return FAST_E_BADARG;
    uint8_t* codes = nullptr;
    int rc = face_codes(v, &codes);
    if (rc) return rc;
    rc = run_area(v, codes, isovalue, 1, area);
    dfree(ctx, codes);
    return rc;
}

int fast_alpha_mesh(fast_volume* v, float isovalue, float* verts, size_t capacity, size_t* nverts) {
    if (!v || !nverts) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!iso_ok(ctx, isovalue)) return FAST_E_BADARG;
    return run_mesh(v, isovalue, nullptr, capacity, verts, nverts);
}

int fast_alpha_save_obj(fast_volume* v, float isovalue, const char* path, size_t* nverts) {
    if (!v || !path || !nverts) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!iso_ok(ctx, isovalue)) return FAST_E_BADARG;
    std::vector<float> vtx;
    int rc = run_mesh(v, isovalue, &vtx, 0, nullptr, nverts);
    if (rc) return rc;
    if (*nverts == 0) return FAST_OK;   // the reference writes no file for an empty mesh
    try {
        if (!write_obj(vtx, *nverts, path)) {
            set_error(ctx, "cannot write %s", path);
            return FAST_E_IO;
        }
    } catch (const std::bad_alloc&) {
        return FAST_E_OOM;
    }
    return FAST_OK;
}

int fast_debug_time_alpha(fast_volume* v, int iters, double* ms) {
    if (!v || !ms || iters < 1) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!fta::mc_fits(v->nx, v->ny, v->nz)) return FAST_E_BADARG;
    const size_t nblk = fta::mc_blocks(v->nx, v->ny, v->nz);
    double* d = (double*)dalloc(ctx, (nblk * 12 + 6) * sizeof(double));
    if (!d) return FAST_E_OOM;
    hipEvent_t ev[2];
    bool ok = hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    int rc = ok ? FAST_OK : FAST_E_HIP;
    for (int which = 0; which < 3 && rc == FAST_OK; ++which) {
        hipError_t e = hipSuccess;
        // one untimed run first: the first launch of a kernel also loads its code object
        for (int it = -1; it < iters && e == hipSuccess && rc == FAST_OK; ++it) {
            if (it == 0) {
                e = hipEventRecord(ev[0], ctx->stream);
                if (e != hipSuccess) break;
            }
            if (which == 0) {
                e = fta::mc_area(v->d_mask, v->nx, v->ny, v->nz, 0.5f, 0, d, d + nblk * 12, ctx->stream);
                continue;
            }
            if (which == 2) v->labels_bg = -1;   // relabel from scratch
            uint8_t* codes = nullptr;
            rc = face_codes(v, &codes);
            if (rc) break;
            e = fta::mc_area(codes, v->nx, v->ny, v->nz, 0.5f, 1, d, d + nblk * 12, ctx->stream);
            dfree(ctx, codes);
        }
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(ev[1], ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
        if (!hip_ok(ctx, e, "alpha timing")) rc = FAST_E_HIP;
        ms[which] = (double)t / iters;
    }
    if (ok) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx, d);
    return rc;
}

}  // extern "C"
