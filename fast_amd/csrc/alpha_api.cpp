// Reactive area density (`fast alpha`): C ABI over the marching-cubes area kernels
// (alpha_kernels.hip) and the CCL (ccl_api.cpp).
// Reference: getReactiveAreaDensity, getReactiveAreaDensityTensor
// (fastlib/src/volume/VolumeMeasures.cpp:294-333), getVolumeArea
// (fastlib/src/volume/VolumeSurface.cpp:9-36), generateBoundaryConnectedVolume
// (fastlib/src/volume/VolumeSegmentation.cpp:66-90), AlphaModule::getAlpha (fast/src/AlphaModule.h:28-174).
#include <cmath>

#include "fast_tau.h"
#include "tau_host.h"

namespace fta {
bool mc_fits(int nx, int ny, int nz);
size_t mc_blocks(int nx, int ny, int nz);
hipError_t mc_area(const uint8_t* src, int nx, int ny, int nz, float iso, int faces, double* part, double* out,
                   hipStream_t s);
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
        hipError_t e = hipEventRecord(ev[0], ctx->stream);
        for (int it = 0; it < iters && e == hipSuccess && rc == FAST_OK; ++it) {
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
