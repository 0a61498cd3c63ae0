// Reachable-pore labels: C ABI over the CCL kernels (ccl_kernels.hip).
// Reference: getVolumeCCL (fastlib/src/volume/VolumeSegmentation.cpp:33-64),
// VolumeCCL_BoundaryLabels (fastlib/src/cuda/VolumeCCL.cu:753-818),
// VolumeCCL::getDirMask(begin, end) (VolumeSegmentation.cpp:14-24).
#include "fast_tau.h"
#include "tau_host.h"

#include <cstring>
#include <vector>

namespace ftc {
hipError_t ccl_label(const uint8_t* mask, int nx, int ny, int nz, uint8_t background, uint32_t* P, uint32_t* L,
                     uint32_t* blocksum, uint32_t* num_roots, hipStream_t s);
hipError_t ccl_face_table(const uint32_t* L, int nx, int ny, int nz, uint8_t* table, uint32_t numLabels, hipStream_t s);
hipError_t ccl_segment_stats(const uint32_t* L, int nx, int ny, int nz, uint32_t numLabels, uint32_t* info,
                             hipStream_t s);
hipError_t ccl_reachable(const uint32_t* L, long long n, const uint8_t* table, uint32_t numLabels, int dir, uint8_t* out,
                         hipStream_t s);
}  // namespace ftc

namespace fth {

// Label the volume (cached on the volume per background value).  Returns numLabels = K + 1.
int ensure_labels(fast_volume* v, uint8_t background, uint32_t* num_labels) {
    fast_ctx* ctx = v->ctx;
    if (v->d_labels && v->labels_bg == background) {
        *num_labels = v->num_labels;
        return FAST_OK;
    }
    // labels and span starts are 32-bit voxel indices (all-ones marks "no span"); volumes hold
    // fewer than 2^31 voxels (fast_volume_create refuses larger ones), so every index fits
    const long long n = (long long)v->nx * v->ny * v->nz;
    if (!v->d_labels) v->d_labels = (uint32_t*)dalloc(ctx, (size_t)n * 4);
    uint32_t* P = (uint32_t*)dalloc(ctx, (size_t)n * 4);
    const long long nb = (n + 4095) / 4096;
    uint32_t* aux = (uint32_t*)dalloc(ctx, (size_t)(nb + 64) * 4);
    if (!v->d_labels || !P || !aux) {
        dfree(ctx, P);
        dfree(ctx, aux);
        return FAST_E_OOM;
    }
    uint32_t roots = 0;
    hipError_t e = ftc::ccl_label(v->d_mask, v->nx, v->ny, v->nz, background, P, v->d_labels, aux, aux + nb, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&roots, aux + nb, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, P);
    dfree(ctx, aux);
    if (!hip_ok(ctx, e, "ccl")) return FAST_E_HIP;
    v->labels_bg = background;
    v->num_labels = roots + 1;
    *num_labels = roots + 1;
    return FAST_OK;
}

int face_table(fast_volume* v, uint32_t numLabels, uint8_t** out) {
    fast_ctx* ctx = v->ctx;
    uint8_t* t = (uint8_t*)dalloc(ctx, (size_t)7 * numLabels);
    if (!t) return FAST_E_OOM;
    hipError_t e = hipMemsetAsync(t, 0, (size_t)7 * numLabels, ctx->stream);
    if (e == hipSuccess) e = ftc::ccl_face_table(v->d_labels, v->nx, v->ny, v->nz, t, numLabels, ctx->stream);
    if (!hip_ok(ctx, e, "face table")) {
        dfree(ctx, t);
        return FAST_E_HIP;
    }
    *out = t;
    return FAST_OK;
}

}  // namespace fth

using namespace fth;

extern "C" {

int fast_ccl_labels(fast_volume* v, uint8_t background, uint32_t* labels_out, uint32_t* num_labels) {
    if (!v || !num_labels) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    int rc = ensure_labels(v, background, num_labels);
    if (rc) return rc;
    if (labels_out) {
        const size_t n = (size_t)v->nx * v->ny * v->nz;
        hipError_t e = hipMemcpy(labels_out, v->d_labels, n * 4, hipMemcpyDeviceToHost);
        if (!hip_ok(ctx, e, "labels readback")) return FAST_E_HIP;
    }
    return FAST_OK;
}

int fast_ccl_boundary_table(fast_volume* v, uint8_t background, uint8_t* table, uint32_t table_labels) {
    if (!v || !table) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    uint32_t nl = 0;
    int rc = ensure_labels(v, background, &nl);
    if (rc) return rc;
    if (table_labels != nl) {
        set_error(ctx, "boundary table needs %u labels, caller sized it for %u", nl, table_labels);
        return FAST_E_BADARG;
    }
    uint8_t* t = nullptr;
    rc = face_table(v, nl, &t);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(table, t, (size_t)7 * nl, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, t);
    return hip_ok(ctx, e, "table readback") ? FAST_OK : FAST_E_HIP;
}

int fast_ccl_segment_info(fast_volume* v, uint8_t background, fast_segment_info* out, uint32_t num_labels) {
    if (!v || !out) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    uint32_t nl = 0;
    int rc = ensure_labels(v, background, &nl);
    if (rc) return rc;
    if (num_labels != nl) {
        set_error(ctx, "segment info has %u labels, caller sized it for %u", nl, num_labels);
        return FAST_E_BADARG;
    }
    uint32_t* d = (uint32_t*)dalloc(ctx, (size_t)7 * nl * sizeof(uint32_t));
    if (!d) return FAST_E_OOM;
    std::vector<uint32_t> h;
    try {
        h.resize((size_t)7 * nl);
    } catch (const std::bad_alloc&) {
        dfree(ctx, d);
        return FAST_E_OOM;
    }
    hipError_t e = ftc::ccl_segment_stats(v->d_labels, v->nx, v->ny, v->nz, nl, d, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, d);
    if (!hip_ok(ctx, e, "segment info")) return FAST_E_HIP;
    const int dim[3] = {v->nx, v->ny, v->nz};
    for (uint32_t l = 0; l < nl; ++l) {
        const uint32_t* r = &h[(size_t)7 * l];
        fast_segment_info& o = out[l];
        o.voxel_num = (int64_t)r[0];
        bool at = false;
        for (int a = 0; a < 3; ++a) {
            o.min_bb[a] = (int32_t)r[1 + a];
            o.max_bb[a] = (int32_t)r[4 + a] + 1;   // non-inclusive, as the reference
            at = at || o.max_bb[a] == dim[a] || o.min_bb[a] == 0;
        }
        o.at_boundary = at ? 1 : 0;
        o.label_id = (int32_t)l;
    }
    return FAST_OK;
}

int fast_reachable_mask(fast_volume* v, int dir, uint8_t background, uint8_t* out, uint32_t* num_labels) {
    if (!v || !out || dir < 0 || dir > 5) return FAST_E_BADARG;
    fast_ctx* ctx = v->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    uint32_t nl = 0;
    int rc = ensure_labels(v, background, &nl);
    if (rc) return rc;
    if (num_labels) *num_labels = nl;
    uint8_t* t = nullptr;
    rc = face_table(v, nl, &t);
    if (rc) return rc;
    const long long n = (long long)v->nx * v->ny * v->nz;
    uint8_t* d_out = (uint8_t*)dalloc(ctx, (size_t)n);
    if (!d_out) {
        dfree(ctx, t);
        return FAST_E_OOM;
    }
    hipError_t e = ftc::ccl_reachable(v->d_labels, n, t, nl, dir, d_out, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, t);
    dfree(ctx, d_out);
    return hip_ok(ctx, e, "reachable") ? FAST_OK : FAST_E_HIP;
}

}  // extern "C"
