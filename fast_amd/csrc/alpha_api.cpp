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
    if (!hip_ok(ctx, e, "face codes")) rc = FAST_E_HIP;
    if (!rc) rc = run_area(v, codes, isovalue, 1, area);
    // run_area synchronised the stream (or failed), so the buffers are idle
    if (rc) (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx, codes);
    dfree(ctx, table);
    return rc;
}

}  // extern "C"
