// Host side of libfasttau: contexts, device memory, resident volumes, the batched solve
// driver and the C ABI declared in include/fast_tau.h.
//
// The reference driver is getTortuosity<T> (fastlib/src/volume/VolumeMeasures.cpp:42-285):
// assemble -> solve -> read x -> flux -> tau, returning 0 on any failure.  Here assembly is
// replaced by a per-voxel stencil code built once per uploaded volume, the Krylov scalars stay
// on the device (the reference copies five reductions to the host per BiCGSTAB iteration,
// fastlib/src/cuda/LinearSystemDevice.cuh:60-112), and the flux integral runs on the device.
#include "fast_tau.h"
#include "tau_internal.h"
#include "tau_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <dlfcn.h>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <tuple>
#include <vector>

using namespace ft;

namespace fta {
hipError_t mc_scratch_address(int device, double** out);   // alpha_kernels.hip
}  // namespace fta

// ------------------------------------------------------------------------------------------
// context and device memory
// ------------------------------------------------------------------------------------------
namespace fth {

void set_error(fast_ctx* ctx, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    if (ctx && ctx->verbose) std::fprintf(stderr, "fasttau: %s\n", buf);
}

static size_t round_alloc(size_t b) {
    if (b <= (1u << 20)) {
        size_t r = 4096;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t g = 2u << 20;
    return (b + g - 1) / g * g;
}

void* dalloc(fast_ctx* ctx, size_t bytes) {
    const size_t sz = round_alloc(bytes ? bytes : 1);
    auto it = ctx->freelist.find(sz);
    if (it != ctx->freelist.end()) {
        void* p = it->second;
        ctx->freelist.erase(it);
        ctx->live[p] = sz;
        return p;
    }
    if (ctx->held + sz > ctx->budget) {
        // release cached blocks before giving up
        for (auto& kv : ctx->freelist) {
            (void)hipFree(kv.second);
            ctx->held -= kv.first;
        }
        ctx->freelist.clear();
        if (ctx->held + sz > ctx->budget) {
            set_error(ctx, "device budget exhausted: need %zu bytes, hold %zu of %zu", sz, ctx->held, ctx->budget);
            return nullptr;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        for (auto& kv : ctx->freelist) {
            (void)hipFree(kv.second);
            ctx->held -= kv.first;
        }
        ctx->freelist.clear();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            set_error(ctx, "hipMalloc(%zu) failed", sz);
            return nullptr;
        }
    }
    ctx->held += sz;
    ctx->live[p] = sz;
    return p;
}

void dfree(fast_ctx* ctx, void* p) {
    if (!p) return;
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) return;
    ctx->freelist.emplace(it->second, p);
    ctx->live.erase(it);
}

void* pinned(fast_ctx* ctx, int slot, size_t bytes) {
    auto& pb = ctx->pin[slot];
    if (pb.size < bytes) {
        if (pb.ptr) (void)hipHostFree(pb.ptr);
        pb.ptr = nullptr;
        pb.size = 0;
        size_t sz = std::max<size_t>(bytes, 1 << 16);
        if (hipHostMalloc(&pb.ptr, sz, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        pb.size = sz;
    }
    return pb.ptr;
}

bool hip_ok(fast_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(ctx, "%s: %s", what, hipGetErrorString(e));
    return false;
}

// ------------------------------------------------------------------------------------------
// coefficients: restated from fastlib/src/cuda/Diffusion.cuh:85-105 with the values the
// reference feeds in: h = float(1/maxDim) (VolumeMeasures.cpp:68), faceArea = h*h in double
// (Diffusion.cpp:26-28), D = double(float coeffs) (VolumeMeasures.cpp:66-67, TauModule.h:70).
// ------------------------------------------------------------------------------------------
template <class T>
Coef<T> make_coef(int nx, int ny, int nz, int dir, double d_pore, double d_solid) {
    const int maxDim = std::max(nx, std::max(ny, nz));
    const float hf = 1.0f / (float)maxDim;
    const double h = (double)hf;
    const double faceArea = h * h;
    const double d0 = (double)(float)d_pore, d1 = (double)(float)d_solid;
    const T D0 = T(d0), D1 = T(d1);
    auto coef = [&](T Dface, bool boundary) -> T {
        T cellDist = T(h);
        if (boundary) cellDist = h * T(0.5);
        return (Dface * faceArea) / cellDist;
    };
    Coef<T> c;
    c.ci0 = coef((D0 + D0) * T(0.5), false);
    c.ci1 = coef((D1 + D0) * T(0.5), false);
    c.ci2 = coef((D1 + D1) * T(0.5), false);
    c.cb0 = coef((D0 + D0) * T(0.5), true);
    c.cb1 = coef((D1 + D1) * T(0.5), true);
    const bool pos = (dir % 2) == 0;
    c.cbeg = pos ? T(1.0) : T(0.0);
    c.cend = pos ? T(0.0) : T(1.0);
    c.dflux0 = T((float)d_pore);
    c.dflux1 = T((float)d_solid);
    return c;
}
template Coef<double> make_coef<double>(int, int, int, int, double, double);
template Coef<float> make_coef<float>(int, int, int, int, double, double);

// Tile shape: the block is TX x TY threads; TX follows the x extent so short rows do not
// leave lanes idle.  tz (planes per tile) is chosen from the volume alone so that a large
// volume yields ~2048+ tiles.  Both depend only on the volume, which keeps per-volume
// reductions independent of the batch.
static bool shape_compiled(int tx, int ty) {
    return (tx == 64 && (ty == 4 || ty == 8 || ty == 16)) || (tx == 32 && (ty == 8 || ty == 16 || ty == 32)) ||
           (tx == 16 && ty == 16);
}

Shape choose_shape(int nx, int ny, int nz, int precision, const fast_ctx* ctx) {
    Shape s{64, 8, 16};
    if (ctx->forced.tx) return ctx->forced;
    const auto key = shape_key(nx, ny, nz, precision);
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) return it->second;
    if (nx <= 16)
        s = {16, 16, 16};
    else if (nx <= 32)
        s = {32, 8, 16};
    const long long txy = (long long)((nx + s.tx - 1) / s.tx) * ((ny + s.ty - 1) / s.ty);
    long long tz = (long long)nz * txy / 2048;
    tz = std::max<long long>(8, std::min<long long>(32, tz));
    s.tz = (int)tz;
    return s;
}

// Tuned tile shapes: text lines "<size-class key> <tx> <ty> <tz>" written by the autotuner
// (fast_amd/autotune.py), read from $FAST_TAU_TUNING, else from tuning_gfx950.txt in the
// directory of this library.
void load_tuning(fast_ctx* ctx) {
    std::string path;
    if (const char* e = std::getenv("FAST_TAU_TUNING")) path = e;
    if (path.empty()) {
        Dl_info info;
        if (dladdr((void*)&load_tuning, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.find_last_of('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning_gfx950.txt";
        }
    }
    FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return;
    char key[64];
    int a, b, c;   // full grid: tx ty tz; red-black ("rb_" keys): ty tz ahead
    while (std::fscanf(f, "%63s %d %d %d", key, &a, &b, &c) == 4) {
        if (std::strncmp(key, "rb_", 3) == 0) {
            if (rb_ty_compiled(a) && b >= 1 && b <= 256 && (c == 2 || c == 3)) ctx->rb_tuned[key] = RbShape{a, b, c};
        } else if (shape_compiled(a, b) && c >= 1 && c <= 256) {
            ctx->tuned[key] = Shape{a, b, c};
        }
    }
    std::fclose(f);
}

static int size_class(int e) { return e <= 16 ? 16 : e <= 32 ? 32 : e <= 64 ? 64 : e <= 128 ? 128 : e <= 256 ? 256 : 512; }

std::string shape_key(int nx, int ny, int nz, int precision) {
    char b[64];
    std::snprintf(b, sizeof(b), "%d_%d_%d_%d", size_class(nx), size_class(ny), size_class(nz), precision);
    return b;
}

std::string rb_key(int nx, int ny, int nz) {
    char b[64];
    std::snprintf(b, sizeof(b), "rb_%d_%d_%d", size_class(nx), size_class(ny), size_class(nz));
    return b;
}

bool rb_ty_compiled(int ty) { return ty == 2 || ty == 4 || ty == 8 || ty == 9 || ty == 10 || ty == 12 || ty == 13; }

int rb_tz(int hx, int ny, int nz, int ty) {
    const long long txy = (long long)((hx + kRbTX - 1) / kRbTX) * ((ny + ty - 1) / ty);
    return (int)std::max<long long>(4, std::min<long long>(32, (long long)nz * txy / 2048));
}

RbShape rb_shape(const fast_ctx* ctx, int nx, int ny, int nz) {
    RbShape r{4, 0, 2};
    if (ctx->rb_forced.ty) {
        r = ctx->rb_forced;
    } else {
        auto it = ctx->rb_tuned.find(rb_key(nx, ny, nz));
        if (it != ctx->rb_tuned.end()) r = it->second;
    }
    if (r.tz <= 0) r.tz = rb_tz((nx + 1) / 2, ny, nz, r.ty);
    if (const char* e = std::getenv("FAST_TAU_RB_TZ")) {
        const int v = std::atoi(e);
        if (v > 0) r.tz = v;
    }
    return r;
}

}  // namespace fth

using namespace fth;

// ------------------------------------------------------------------------------------------
// C ABI: context and volumes
// ------------------------------------------------------------------------------------------
extern "C" {

void fast_tau_params_default(fast_tau_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dir = FAST_X_POS;
    p->solver = FAST_CG;
    p->precision = FAST_FP64;
    p->tolerance = 1e-6;
    p->max_iter = 10000;
    p->d_pore = 1.0;
    p->d_solid = 0.001;
    p->porosity = -1.0;
}

int fast_ctx_create(int device, size_t hbm_budget, fast_ctx** out) {
    if (!out) return FAST_E_BADARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return FAST_E_NODEV;
    }
    if (device < 0 || device >= n) return FAST_E_BADARG;
    if (hipSetDevice(device) != hipSuccess) return FAST_E_HIP;
    auto* ctx = new fast_ctx();
    ctx->device = device;
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        delete ctx;
        return FAST_E_HIP;
    }
    ctx->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)freeb);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return FAST_E_HIP;
    }
    for (auto& e : ctx->ev) (void)hipEventCreate(&e);
    (void)hipEventCreate(&ctx->t0);
    (void)hipEventCreate(&ctx->t1);
    ctx->zero_count = (unsigned long long*)dalloc(ctx, 4096);
    if (!ctx->zero_count) {
        fast_ctx_destroy(ctx);
        return FAST_E_OOM;
    }
    if (fta::mc_scratch_address(device, &ctx->alpha_scratch) != hipSuccess) {
        (void)hipGetLastError();
        fast_ctx_destroy(ctx);
        return FAST_E_HIP;
    }
    load_tuning(ctx);
    if (const char* e = std::getenv("FAST_TAU_MIXED_RESTART")) ctx->mixed_restart_cos = std::atof(e);
    if (const char* e = std::getenv("FAST_TAU_MIXED_BLOWUP")) {
        const double g = std::atof(e);   // growth of |r| (not squared); 0 = never on growth alone
        ctx->mixed_blowup2 = g > 0 ? g * g : HUGE_VAL;
    }
    if (const char* e = std::getenv("FAST_TAU_GRAPHS")) ctx->use_graphs = std::atoi(e) != 0;
    if (const char* e = std::getenv("FAST_TAU_GROUP_MB")) ctx->group_bytes = (size_t)std::atol(e) << 20;
    if (const char* e = std::getenv("FAST_TAU_UPDATE_EVERY")) ctx->update_every = std::max(1, std::atoi(e));
    *out = ctx;
    return FAST_OK;
}

void fast_ctx_destroy(fast_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // volumes still alive die with their context (their device memory is in `live`)
    for (auto& kv : ctx->volumes) delete kv.first;
    ctx->volumes.clear();
    for (auto& kv : ctx->live) (void)hipFree(kv.first);
    for (auto& kv : ctx->freelist) (void)hipFree(kv.second);
    for (auto& pb : ctx->pin)
        if (pb.ptr) (void)hipHostFree(pb.ptr);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* fast_ctx_last_error(const fast_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int fast_ctx_set_tile_shape(fast_ctx* ctx, int tx, int ty, int tz) {
    if (!ctx) return FAST_E_BADARG;
    if (tx == 0) {
        ctx->forced = Shape{0, 0, 0};
        return FAST_OK;
    }
    if (!shape_compiled(tx, ty) || tz < 1 || tz > 256) return FAST_E_BADARG;
    ctx->forced = Shape{tx, ty, tz};
    return FAST_OK;
}

int fast_ctx_set_rb_shape(fast_ctx* ctx, int ty, int tz, int ahead) {
    if (!ctx) return FAST_E_BADARG;
    if (ty == 0) {
        ctx->rb_forced = RbShape{0, 0, 0};
        return FAST_OK;
    }
    if (!rb_ty_compiled(ty) || tz < 0 || tz > 256 || (ahead != 2 && ahead != 3)) return FAST_E_BADARG;
    ctx->rb_forced = RbShape{ty, tz, ahead};
    return FAST_OK;
}

int fast_ctx_rb_shape(const fast_ctx* ctx, int nx, int ny, int nz, int* shape) {
    if (!ctx || !shape || nx < 1 || ny < 1 || nz < 1) return FAST_E_BADARG;
    const RbShape r = rb_shape(ctx, nx, ny, nz);
    shape[0] = r.ty;
    shape[1] = r.tz;
    shape[2] = r.ahead;
    return FAST_OK;
}

int fast_ctx_set_round(fast_ctx* ctx, int iterations) {
    if (!ctx || iterations < 0 || iterations > 1024) return FAST_E_BADARG;
    ctx->chunk = iterations ? iterations : 16;
    return FAST_OK;
}

size_t fast_ctx_device_bytes(const fast_ctx* ctx) { return ctx ? ctx->held : 0; }

static int volume_init(fast_ctx* ctx, const uint8_t* src, bool src_device, int nx, int ny, int nz, fast_volume** out) {
    if (!ctx || !src || !out || nx <= 0 || ny <= 0 || nz <= 0) return FAST_E_BADARG;
    const size_t n = (size_t)nx * ny * nz;
    if (n >= (size_t)1 << 31) {
        set_error(ctx, "volume of %zu voxels exceeds the 2^31 voxel limit", n);
        return FAST_E_BADARG;
    }
    (void)hipSetDevice(ctx->device);
    auto* v = new fast_volume();
    v->ctx = ctx;
    v->nx = nx;
    v->ny = ny;
    v->nz = nz;
    v->d_mask = (uint8_t*)dalloc(ctx, n);
    v->d_code = (uint8_t*)dalloc(ctx, n);
    if (!v->d_mask || !v->d_code) {
        fast_volume_destroy(v);
        return FAST_E_OOM;
    }
    hipError_t e = hipMemcpyAsync(v->d_mask, src, n, src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->zero_count, 0, sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = ft::build_code(v->d_mask, v->d_code, nx, ny, nz, ctx->zero_count, ctx->stream);
    unsigned long long zeros = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&zeros, ctx->zero_count, sizeof(zeros), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (!hip_ok(ctx, e, "volume upload")) {
        fast_volume_destroy(v);
        return FAST_E_HIP;
    }
    v->zeros = zeros;
    ctx->volumes[v] = 1;
    *out = v;
    return FAST_OK;
}

int fast_volume_create(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz, fast_volume** out) {
    return volume_init(ctx, mask, false, nx, ny, nz, out);
}

int fast_volume_create_device(fast_ctx* ctx, const uint8_t* d_mask, int nx, int ny, int nz, fast_volume** out) {
    return volume_init(ctx, d_mask, true, nx, ny, nz, out);
}

void fast_volume_destroy(fast_volume* v) {
    if (!v) return;
    fast_ctx* ctx = v->ctx;
    dfree(ctx, v->d_mask);
    dfree(ctx, v->d_code);
    dfree(ctx, v->d_labels);
    ctx->volumes.erase(v);
    delete v;
}

int fast_volume_count_zero(fast_volume* v, uint64_t* count) {
    if (!v || !count) return FAST_E_BADARG;
    *count = v->zeros;
    return FAST_OK;
}

int fast_porosity(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz, double* porosity) {
    if (!porosity) return FAST_E_BADARG;
    fast_volume* v = nullptr;
    int rc = fast_volume_create(ctx, mask, nx, ny, nz, &v);
    if (rc) return rc;
    // getPorosity<double>: T(sumZeroElems) / T(totalElems)
    *porosity = (double)v->zeros / (double)((size_t)nx * ny * nz);
    fast_volume_destroy(v);
    return FAST_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// the batched solve driver
// ------------------------------------------------------------------------------------------
namespace fth {

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

template <class T, class A>
struct Batch {
    fast_ctx* ctx = nullptr;
    std::vector<Job*> jobs;
    Shape shape{64, 8, 16};
    bool mixed = false;
    std::vector<VolDesc> hvol;
    using C = typename CmpOf<T, A>::type;
    std::vector<int> htile, hplane, hplane_k;
    std::vector<Coef<C>> hcoef;
    std::vector<Coef<double>> hcoef64;
    std::vector<VolState<A>> hst;
    int ntiles = 0, nplanes = 0;
    size_t total = 0;
    void* meta = nullptr;
    void* vecs = nullptr;
    Launch<T, A> L{};
    Coef<double>* coef64 = nullptr;
    double* plane_out = nullptr;
    int* plane_vol = nullptr;
    int* plane_k = nullptr;
    hipGraphExec_t gexec = nullptr;   // one captured polling round of this batch

    ~Batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (ctx) {
            dfree(ctx, meta);
            dfree(ctx, vecs);
        }
    }
};

// Run `enqueue` (one polling round) from a graph captured on first use.  The graph holds
// this batch's kernel arguments, so it lives and dies with the batch.
template <class BT, class F>
static hipError_t replay(BT& B, F enqueue) {
    fast_ctx* ctx = B.ctx;
    if (!ctx->use_graphs) return enqueue();
    if (!B.gexec) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        const hipError_t ee = enqueue();
        e = hipStreamEndCapture(ctx->stream, &g);
        if (ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return ee;
        }
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&B.gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            B.gexec = nullptr;
            return e;
        }
    }
    return hipGraphLaunch(B.gexec, ctx->stream);
}

template <class T, class A>
static int prepare(Batch<T, A>& B) {
    fast_ctx* ctx = B.ctx;
    const int nv = (int)B.jobs.size();
    B.hvol.resize(nv);
    B.hcoef.resize(nv);
    B.hst.assign(nv, VolState<A>{});
    if (B.mixed) B.hcoef64.resize(nv);
    int tiles = 0, planes = 0;
    size_t total = 0;
    for (int i = 0; i < nv; ++i) {
        const Job& j = *B.jobs[i];
        const fast_volume* v = j.vol;
        if ((size_t)v->nx * v->ny * v->nz > kMaxSolveVoxels) {
            set_error(ctx, "volume of %d x %d x %d voxels exceeds the solver's limit of %zu voxels", v->nx, v->ny, v->nz,
                      kMaxSolveVoxels);
            return FAST_E_BADARG;
        }
        VolDesc& d = B.hvol[i];
        d.nx = v->nx;
        d.ny = v->ny;
        d.nz = v->nz;
        d.dir = j.prm.dir;
        d.primary = j.prm.dir / 2;
        d.sgn = (j.prm.dir % 2 == 0) ? 1 : -1;
        d.tiles_x = (v->nx + B.shape.tx - 1) / B.shape.tx;
        d.tiles_y = (v->ny + B.shape.ty - 1) / B.shape.ty;
        d.tz = j.tz;
        d.tiles_z = (v->nz + d.tz - 1) / d.tz;
        d.tile_base = tiles;
        d.ntiles = d.tiles_x * d.tiles_y * d.tiles_z;
        d.plane_base = planes;
        d.flux_nu = j.prm.flux_nonuniform ? 1 : 0;
        d.off = (long long)total;
        d.code = v->d_code;
        tiles += d.ntiles;
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        planes += np;
        total += align_up((size_t)v->nx * v->ny * v->nz, 64);
        B.hcoef[i] = make_coef<typename Batch<T, A>::C>(v->nx, v->ny, v->nz, j.prm.dir, j.prm.d_pore, j.prm.d_solid);
        if (B.mixed) B.hcoef64[i] = make_coef<double>(v->nx, v->ny, v->nz, j.prm.dir, j.prm.d_pore, j.prm.d_solid);
        VolState<A>& s = B.hst[i];
        s.status = ST_RUNNING;
        s.tol = A(j.prm.tolerance);
        s.maxiter = j.prm.max_iter;
        s.eps2 = (sizeof(T) == 8) ? A(2.220446049250313e-16) * A(2.220446049250313e-16)
                                  : A(1.1920928955078125e-07f) * A(1.1920928955078125e-07f);
        if (B.mixed) {
            s.mode = 1;
            s.status = ST_FRESH;   // the first residual step starts every volume
            s.otol = j.outer_tol;
            s.delta2 = j.update_delta * j.update_delta;
            s.restart_cos = ctx->mixed_restart_cos;
            s.blowup2 = ctx->mixed_blowup2;
            s.solver = j.prm.solver;
        }
    }
    B.ntiles = tiles;
    B.nplanes = planes;
    B.total = total;
    B.htile.resize(tiles);
    B.hplane.resize(planes);
    B.hplane_k.resize(planes);
    for (int i = 0; i < nv; ++i) {
        for (int t = 0; t < B.hvol[i].ntiles; ++t) B.htile[B.hvol[i].tile_base + t] = i;
        const int np = B.hvol[i].primary == 0 ? B.hvol[i].nx : (B.hvol[i].primary == 1 ? B.hvol[i].ny : B.hvol[i].nz);
        for (int k = 0; k < np; ++k) {
            B.hplane[B.hvol[i].plane_base + k] = i;
            B.hplane_k[B.hvol[i].plane_base + k] = k;
        }
    }
    // meta layout
    size_t o = 0;
    const size_t o_vol = o;
    o = align_up(o + sizeof(VolDesc) * nv, 256);
    const size_t o_coef = o;
    o = align_up(o + sizeof(Coef<typename Batch<T, A>::C>) * nv, 256);
    const size_t o_coef64 = o;
    o = align_up(o + sizeof(Coef<double>) * nv, 256);
    const size_t o_st = o;
    o = align_up(o + sizeof(VolState<A>) * nv, 256);
    const size_t o_tile = o;
    o = align_up(o + sizeof(int) * tiles, 256);
    const size_t o_part = o;
    o = align_up(o + sizeof(double) * 8 * (size_t)tiles, 256);
    const size_t o_pvol = o;
    o = align_up(o + sizeof(int) * planes, 256);
    const size_t o_pk = o;
    o = align_up(o + sizeof(int) * planes, 256);
    const size_t o_pout = o;
    o = align_up(o + sizeof(double) * planes, 256);
    const size_t o_tab = o;
    o = align_up(o + sizeof(typename Batch<T, A>::C) * 2 * kTabEntries * nv, 256);
    const size_t meta_bytes = o;
    B.meta = dalloc(ctx, meta_bytes);
    // x, r, r0, s in T; p[2], v[2] in the directions' type (FP64 for mixed BiCGSTAB, DirOf);
    // the mixed form's FP64 solution
    bool bicg = false;
    for (const Job* j : B.jobs) bicg = bicg || j->prm.solver == FAST_BICGSTAB;
    const size_t dsz = bicg ? sizeof(typename DirOf<T, A>::type) : sizeof(T);
    const size_t vec_bytes = 4 * total * sizeof(T) + 4 * total * dsz + (B.mixed ? total * sizeof(double) : 0);
    B.vecs = dalloc(ctx, vec_bytes);
    if (!B.meta || !B.vecs) return FAST_E_OOM;
    std::vector<char> h(meta_bytes, 0);
    std::memcpy(h.data() + o_vol, B.hvol.data(), sizeof(VolDesc) * nv);
    std::memcpy(h.data() + o_coef, B.hcoef.data(), sizeof(Coef<typename Batch<T, A>::C>) * nv);
    if (B.mixed) std::memcpy(h.data() + o_coef64, B.hcoef64.data(), sizeof(Coef<double>) * nv);
    std::memcpy(h.data() + o_st, B.hst.data(), sizeof(VolState<A>) * nv);
    std::memcpy(h.data() + o_tile, B.htile.data(), sizeof(int) * tiles);
    std::memcpy(h.data() + o_pvol, B.hplane.data(), sizeof(int) * planes);
    std::memcpy(h.data() + o_pk, B.hplane_k.data(), sizeof(int) * planes);
    hipError_t e = hipMemcpyAsync(B.meta, h.data(), meta_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // h goes out of scope
    if (!hip_ok(ctx, e, "batch upload")) return FAST_E_HIP;
    char* m = (char*)B.meta;
    B.L.tx = B.shape.tx;
    B.L.ty = B.shape.ty;
    B.L.ntiles = tiles;
    B.L.nvol = nv;
    B.L.vol = (const VolDesc*)(m + o_vol);
    B.L.tile_vol = (const int*)(m + o_tile);
    B.L.coef = (const Coef<typename Batch<T, A>::C>*)(m + o_coef);
    B.L.st = (VolState<A>*)(m + o_st);
    B.L.part = (double*)(m + o_part);
    B.coef64 = (Coef<double>*)(m + o_coef64);
    B.plane_vol = (int*)(m + o_pvol);
    B.plane_k = (int*)(m + o_pk);
    B.plane_out = (double*)(m + o_pout);
    B.L.tab = (typename Batch<T, A>::C*)(m + o_tab);
    char* vb = (char*)B.vecs;
    const size_t tb = total * sizeof(T), db = total * dsz;
    VecSet& vs = B.L.vs;
    vs.x = vb;
    vs.r = vb + tb;
    vs.r0 = vb + 2 * tb;
    vs.s = vb + 3 * tb;
    vs.p[0] = vb + 4 * tb;
    vs.p[1] = vb + 4 * tb + db;
    vs.v[0] = vb + 4 * tb + 2 * db;
    vs.v[1] = vb + 4 * tb + 3 * db;
    vs.x64 = B.mixed ? (double*)(vb + 4 * tb + 4 * db) : nullptr;
    e = ft::build_tables(B.L, ctx->stream);
    return hip_ok(ctx, e, "diagonal tables") ? FAST_OK : FAST_E_HIP;
}

// Poll the device-resident states every `chunk` iterations, keeping one chunk in flight
// (no host bubble).  `step` enqueues one chunk.  Returns when every volume reached a final
// status (`final_status`); restarts are served synchronously (they are rare).
// `poll` sees each snapshot of the states while the next chunk is in flight; what it enqueues
// runs after that chunk.
template <class BT, class Step, class Restart, class Poll>
static int drive(BT& B, int final_status, long max_chunks, Step step, Restart restart, Poll poll) {
    using St = typename decltype(B.hst)::value_type;
    fast_ctx* ctx = B.ctx;
    const int nv = (int)B.jobs.size();
    const size_t sbytes = sizeof(St) * nv;
    St* snap[2] = {(St*)pinned(ctx, 0, sbytes), (St*)pinned(ctx, 1, sbytes)};
    if (!snap[0] || !snap[1]) {
        set_error(ctx, "pinned host allocation failed");
        return FAST_E_OOM;
    }
    int slot = 0;
    bool pending = false;
    long chunks = 0;
    for (;;) {
        hipError_t e = step();
        if (e == hipSuccess) e = hipMemcpyAsync(snap[slot], B.L.st, sbytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev[slot], ctx->stream);
        if (!hip_ok(ctx, e, "enqueue")) return FAST_E_HIP;
        ++chunks;
        if (pending) {
            const int prev = slot ^ 1;
            e = hipEventSynchronize(ctx->ev[prev]);
            if (!hip_ok(ctx, e, "poll")) return FAST_E_HIP;
            bool all_final = true, any_restart = false;
            if (ctx->verbose) {
                for (int i = 0; i < nv; ++i) {
                    const St& q = snap[prev][i];
                    std::fprintf(stderr, "fasttau: round %ld vol %d status %d iter %d rsq %.3e tol2 %.3e r64sq %.3e "
                                 "updates %d restarts %d rho %.3e alpha %.3e omega %.3e\n",
                                 chunks - 1, i, q.status, q.iter, (double)q.rsq, (double)q.tol2, (double)q.r64sq,
                                 q.refinements, q.restarts, (double)q.rho, (double)q.alpha, (double)q.omega);
                }
            }
            for (int i = 0; i < nv; ++i) {
                const int st = snap[prev][i].status;
                if (st == ST_RESTART) any_restart = true;
                if (st != final_status) all_final = false;
            }
            if (all_final) break;
            if (any_restart) {
                e = hipStreamSynchronize(ctx->stream);
                if (!hip_ok(ctx, e, "restart sync")) return FAST_E_HIP;
                e = restart();
                if (!hip_ok(ctx, e, "restart")) return FAST_E_HIP;
                pending = false;
                slot ^= 1;
                continue;
            }
            e = poll(snap[prev]);
            if (!hip_ok(ctx, e, "poll step")) return FAST_E_HIP;
        }
        if (chunks > max_chunks) {
            set_error(ctx, "solve did not terminate within %ld polling rounds", max_chunks);
            return FAST_E_NOCONV;
        }
        pending = true;
        slot ^= 1;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(B.hst.data(), B.L.st, sbytes, hipMemcpyDeviceToHost);
    return hip_ok(ctx, e, "final state") ? FAST_OK : FAST_E_HIP;
}

template <class BT, class Step, class Restart>
static int drive(BT& B, int final_status, long max_chunks, Step step, Restart restart) {
    return drive(B, final_status, max_chunks, step, restart, [](const void*) { return hipSuccess; });
}

// Flux integral on the device, then tau = porosity / (dc * N_primary) in T like the
// reference (VolumeMeasures.cpp:266-275).
template <class TC, class BT>
static int finish(BT& B, const VolDesc* vol_dev, const TC* conc, const Coef<TC>* coef_dev, const std::vector<bool>& ok) {
    fast_ctx* ctx = B.ctx;
    const int nv = (int)B.jobs.size();
    hipError_t e = ft::flux_planes<TC>(B.nplanes, vol_dev, B.plane_vol, B.plane_k, coef_dev, conc, B.plane_out,
                                       ctx->stream);
    std::vector<double> planes(B.nplanes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(planes.data(), B.plane_out, sizeof(double) * B.nplanes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->t1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (!hip_ok(ctx, e, "flux")) return FAST_E_HIP;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->t0, ctx->t1);
    for (int i = 0; i < nv; ++i) {
        Job& j = *B.jobs[i];
        const VolDesc& d = B.hvol[i];
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        TC sum = TC(0);
        for (int k = 0; k < np; ++k) sum += (TC)planes[d.plane_base + k];
        const TC dc = sum / TC(np);
        const TC por = TC(j.porosity);
        j.res->porosity = j.porosity;
        j.res->dc = (double)dc;
        j.res->solve_ms = ms;
        j.res->tau = ok[i] ? (double)(por / (dc * TC(np))) : 0.0;
        if (j.conc) {
            const size_t n = (size_t)d.nx * d.ny * d.nz;
            std::vector<TC> tmp(n);
            e = hipMemcpy(tmp.data(), conc + d.off, n * sizeof(TC), hipMemcpyDeviceToHost);
            if (!hip_ok(ctx, e, "conc readback")) return FAST_E_HIP;
            for (size_t q = 0; q < n; ++q) j.conc[q] = (double)tmp[q];
        }
    }
    return FAST_OK;
}

template <class T, class A>
static int solve_plain(fast_ctx* ctx, std::vector<Job*>& jobs, Shape shape, int solver) {
    Batch<T, A> B;
    B.ctx = ctx;
    B.jobs = jobs;
    B.shape = shape;
    int rc = prepare(B);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    hipError_t e = hipEventRecord(ctx->t0, s);
    if (e == hipSuccess) e = (solver == FAST_BICGSTAB) ? Passes<T, A>::init_bicg(B.L, s) : Passes<T, A>::init_cg(B.L, s);
    if (!hip_ok(ctx, e, "init")) return FAST_E_HIP;
    long maxit = 0;
    for (auto* j : jobs) maxit = std::max(maxit, j->prm.max_iter);
    const int chunk = ctx->chunk;
    const long max_chunks = 3 * (maxit / chunk + 2) + 16;
    auto round = [&]() -> hipError_t {
        for (int i = 0; i < chunk; ++i) {
            hipError_t e2 = (solver == FAST_BICGSTAB) ? Passes<T, A>::bicg_iter(B.L, s) : Passes<T, A>::cg_iter(B.L, s);
            if (e2 != hipSuccess) return e2;
        }
        return hipSuccess;
    };
    auto step = [&]() -> hipError_t { return replay(B, round); };
    auto restart = [&]() -> hipError_t { return Passes<T, A>::restart(B.L, s); };
    rc = drive(B, ST_DONE, max_chunks, step, restart);
    if (rc) return rc;
    std::vector<bool> ok(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const VolState<A>& st = B.hst[i];
        Job& j = *jobs[i];
        double err;
        if (st.zero_rhs)
            err = (solver == FAST_BICGSTAB) ? j.prm.tolerance : 0.0;
        else
            err = (double)std::sqrt(st.rsq / st.bsq);
        const bool success = (A(err) <= A(j.prm.tolerance));
        ok[i] = success;
        j.res->rel_residual = err;
        j.res->true_rel_residual = -1;
        j.res->iterations = st.iter;
        j.res->restarts = st.restarts;
        j.res->refinements = 0;
        j.res->status = success ? FAST_SOLVED : FAST_NO_CONVERGENCE;
    }
    return finish<T>(B, B.L.vol, (const T*)B.L.vs.x, B.L.coef, ok);
}

// Mixed form: FP32 BiCGSTAB with FP64 reductions and FP64 reliable updates.  After every
// iteration an update step (no-op unless a volume asked for one) moves d into x64 and
// replaces the residual by the FP64 true residual; the FP64 test ends the solve.
static int solve_mixed(fast_ctx* ctx, std::vector<Job*>& jobs, Shape shape, int solver) {
    Batch<float, double> B;
    B.ctx = ctx;
    B.jobs = jobs;
    B.shape = shape;
    B.mixed = true;
    int rc = prepare(B);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    hipError_t e = hipEventRecord(ctx->t0, s);
    if (e == hipSuccess) e = ft::mixed_update(B.L, s);
    if (e == hipSuccess) e = ft::mixed_residual(B.L, B.coef64, s);
    if (!hip_ok(ctx, e, "mixed init")) return FAST_E_HIP;
    long maxit = 0;
    for (auto* j : jobs) maxit = std::max(maxit, j->prm.max_iter);
    const int chunk = ctx->chunk;
    const long max_chunks = 3 * (maxit / chunk + 2) + 16;
    auto round = [&]() -> hipError_t {
        hipError_t e2 = hipSuccess;
        for (int i = 0; i < chunk && e2 == hipSuccess; ++i) {
            e2 = (solver == FAST_BICGSTAB) ? Passes<float, double>::bicg_iter(B.L, s) : Passes<float, double>::cg_iter(B.L, s);
            // A volume that asked for an update skips the iterations until the update step
            // runs, so running it every few iterations changes nothing but the idle launches.
            const bool upd = (i + 1) % ctx->update_every == 0 || i + 1 == chunk;
            if (e2 == hipSuccess && upd) e2 = ft::mixed_update(B.L, s);
            if (e2 == hipSuccess && upd) e2 = ft::mixed_residual(B.L, B.coef64, s);
        }
        return e2;
    };
    auto step = [&]() -> hipError_t { return replay(B, round); };
    auto restart = [&]() -> hipError_t { return hipSuccess; };   // breakdowns restart on the device
    rc = drive(B, ST_IDLE, max_chunks, step, restart);
    if (rc) return rc;
    std::vector<bool> ok(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const VolState<double>& st = B.hst[i];
        Job& j = *jobs[i];
        const double err = st.b64sq > 0 ? std::sqrt(st.r64sq / st.b64sq) : 0.0;
        const bool success = err <= j.outer_tol;
        ok[i] = success;
        j.res->rel_residual = err;
        j.res->true_rel_residual = err;
        j.res->iterations = st.iter;
        j.res->restarts = st.restarts;
        j.res->refinements = st.refinements;
        j.res->status = success ? FAST_SOLVED : FAST_NO_CONVERGENCE;
    }
    return finish<double>(B, B.L.vol, B.L.vs.x64, B.coef64, ok);
}

// Mixed form of CG on the red-black Schur complement (see ft::RbDesc).  One batch holds
// compact per-colour vectors; the FP64 solution of both colours is scattered into a full-grid
// vector for the flux integral, which aliases the FP32 vectors (dead by then).
struct RbBatch {
    fast_ctx* ctx = nullptr;
    std::vector<Job*> jobs;
    std::vector<RbDesc> hrb;
    std::vector<VolDesc> hvol;
    std::vector<Coef<double>> hcoef;
    std::vector<VolState<double>> hst;
    std::vector<int> htile, hplane, hplane_k;
    int ntiles = 0, nplanes = 0;
    void* meta = nullptr;
    void* vecs = nullptr;
    RbLaunch L{};
    const VolDesc* dvol = nullptr;
    double* plane_out = nullptr;
    int* plane_vol = nullptr;
    int* plane_k = nullptr;
    int* tile_vol_c = nullptr;   // compacted launch: volume and batch index of each live tile
    int* tile_id_c = nullptr;
    hipGraphExec_t gexec = nullptr;

    ~RbBatch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (ctx) {
            dfree(ctx, meta);
            dfree(ctx, vecs);
        }
    }
};

int rb_mode() {
    const char* e = std::getenv("FAST_TAU_RB_STAGED");
    return e ? std::min(2, std::max(0, std::atoi(e))) : 2;
}

static int prepare_rb(RbBatch& B) {
    fast_ctx* ctx = B.ctx;
    const int nv = (int)B.jobs.size();
    B.hrb.resize(nv);
    B.hvol.resize(nv);
    B.hcoef.resize(nv);
    B.hst.assign(nv, VolState<double>{});
    // rows per tile and prefetch distance: those of the largest volume (one launch set serves
    // the batch); planes per tile: each volume's own
    const fast_volume* big = B.jobs[0]->vol;
    for (const Job* j : B.jobs)
        if ((size_t)j->vol->nx * j->vol->ny * j->vol->nz > (size_t)big->nx * big->ny * big->nz) big = j->vol;
    const RbShape bs = rb_shape(ctx, big->nx, big->ny, big->nz);
    int tiles = 0, planes = 0;
    size_t total = 0, total_c = 0;
    for (int i = 0; i < nv; ++i) {
        const Job& j = *B.jobs[i];
        const fast_volume* v = j.vol;
        if ((size_t)v->nx * v->ny * v->nz > kMaxSolveVoxels) {
            set_error(ctx, "volume of %d x %d x %d voxels exceeds the solver's limit of %zu voxels", v->nx, v->ny, v->nz,
                      kMaxSolveVoxels);
            return FAST_E_BADARG;
        }
        RbDesc& r = B.hrb[i];
        r.nx = v->nx;
        r.ny = v->ny;
        r.nz = v->nz;
        r.hx = (v->nx + 1) / 2;
        r.px = rb_row_pitch(r.hx);
        r.pp = (unsigned)r.px * (unsigned)(v->ny + 2);
        r.dir = j.prm.dir;
        r.primary = j.prm.dir / 2;
        r.sgn = (j.prm.dir % 2 == 0) ? 1 : -1;
        r.tiles_x = (r.hx + kRbTX - 1) / kRbTX;
        r.tiles_y = (v->ny + bs.ty - 1) / bs.ty;
        r.tz = rb_shape(ctx, v->nx, v->ny, v->nz).tz;
        r.tiles_z = (v->nz + r.tz - 1) / r.tz;
        r.tile_base = tiles;
        r.ntiles = r.tiles_x * r.tiles_y * r.tiles_z;
        r.off_c = (long long)total_c;
        r.off = (long long)total;
        r.code = v->d_code;
        tiles += r.ntiles;
        total_c += align_up((size_t)r.pp * (size_t)(v->nz + 2), 64);
        // full-grid descriptor for the flux integral (only the geometry and offsets are read)
        VolDesc& d = B.hvol[i];
        d.nx = v->nx;
        d.ny = v->ny;
        d.nz = v->nz;
        d.dir = r.dir;
        d.primary = r.primary;
        d.sgn = r.sgn;
        d.tiles_x = d.tiles_y = d.tiles_z = d.tz = 1;
        d.plane_base = planes;
        d.flux_nu = j.prm.flux_nonuniform ? 1 : 0;
        d.off = r.off;
        d.code = v->d_code;
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        planes += np;
        total += align_up((size_t)v->nx * v->ny * v->nz, 64);
        B.hcoef[i] = make_coef<double>(v->nx, v->ny, v->nz, j.prm.dir, j.prm.d_pore, j.prm.d_solid);
        VolState<double>& s = B.hst[i];
        s.tol = j.prm.tolerance;
        s.maxiter = j.prm.max_iter;
        s.eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
        s.mode = 1;
        s.status = ST_FRESH;
        s.otol = j.outer_tol;
        s.delta2 = j.update_delta * j.update_delta;
        s.restart_cos = ctx->mixed_restart_cos;
        s.blowup2 = ctx->mixed_blowup2;
        s.solver = FAST_CG;
    }
    B.ntiles = tiles;
    B.nplanes = planes;
    B.htile.resize(tiles);
    B.hplane.resize(planes);
    B.hplane_k.resize(planes);
    for (int i = 0; i < nv; ++i) {
        for (int t = 0; t < B.hrb[i].ntiles; ++t) B.htile[B.hrb[i].tile_base + t] = i;
        const VolDesc& d = B.hvol[i];
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        for (int k = 0; k < np; ++k) {
            B.hplane[d.plane_base + k] = i;
            B.hplane_k[d.plane_base + k] = k;
        }
    }
    size_t o = 0;
    const size_t o_rb = o;
    o = align_up(o + sizeof(RbDesc) * nv, 256);
    const size_t o_vol = o;
    o = align_up(o + sizeof(VolDesc) * nv, 256);
    const size_t o_coef = o;
    o = align_up(o + sizeof(Coef<double>) * nv, 256);
    const size_t o_st = o;
    o = align_up(o + sizeof(VolState<double>) * nv, 256);
    const size_t o_tile = o;
    o = align_up(o + sizeof(int) * tiles, 256);
    const size_t o_tvc = o;
    o = align_up(o + sizeof(int) * tiles, 256);
    const size_t o_tidc = o;
    o = align_up(o + sizeof(int) * tiles, 256);
    const size_t o_part = o;
    o = align_up(o + sizeof(double) * 8 * (size_t)tiles, 256);
    const size_t o_partb = o;
    o = align_up(o + sizeof(double) * 8 * (size_t)tiles, 256);
    const size_t o_pvol = o;
    o = align_up(o + sizeof(int) * planes, 256);
    const size_t o_pk = o;
    o = align_up(o + sizeof(int) * planes, 256);
    const size_t o_pout = o;
    o = align_up(o + sizeof(double) * planes, 256);
    const size_t o_tab = o;
    o = align_up(o + sizeof(double) * 2 * kTabEntries * nv, 256);
    const size_t meta_bytes = o;
    // FP32 red vectors x, r, p[2], q (their region later holds the full-grid solution), then
    // FP64 y (black), x64 (red), x64b (black), then the per-colour codes
    const size_t f_bytes = align_up(std::max(5 * total_c * sizeof(float), total * sizeof(double)), 256);
    const size_t d_bytes = 3 * total_c * sizeof(double);
    const size_t vec_bytes = f_bytes + d_bytes + 2 * total_c;
    B.meta = dalloc(ctx, meta_bytes);
    B.vecs = dalloc(ctx, vec_bytes);
    if (!B.meta || !B.vecs) return FAST_E_OOM;
    std::vector<char> h(meta_bytes, 0);
    std::memcpy(h.data() + o_rb, B.hrb.data(), sizeof(RbDesc) * nv);
    std::memcpy(h.data() + o_vol, B.hvol.data(), sizeof(VolDesc) * nv);
    std::memcpy(h.data() + o_coef, B.hcoef.data(), sizeof(Coef<double>) * nv);
    std::memcpy(h.data() + o_st, B.hst.data(), sizeof(VolState<double>) * nv);
    std::memcpy(h.data() + o_tile, B.htile.data(), sizeof(int) * tiles);
    std::memcpy(h.data() + o_pvol, B.hplane.data(), sizeof(int) * planes);
    std::memcpy(h.data() + o_pk, B.hplane_k.data(), sizeof(int) * planes);
    hipError_t e = hipMemcpyAsync(B.meta, h.data(), meta_bytes, hipMemcpyHostToDevice, ctx->stream);
    // ghost and padding elements must read +0 (they are never written)
    if (e == hipSuccess) e = hipMemsetAsync(B.vecs, 0, vec_bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // h goes out of scope
    if (!hip_ok(ctx, e, "batch upload")) return FAST_E_HIP;
    char* m = (char*)B.meta;
    RbLaunch& L = B.L;
    L.ntiles = tiles;
    L.nvol = nv;
    L.staged = rb_mode();
    L.fold_alpha = (nv == 1 && tiles <= kRbFoldMaxTiles) ? 1 : 0;
    L.ty = bs.ty;
    L.ahead = bs.ahead;
    L.vol = (const RbDesc*)(m + o_rb);
    L.tile_vol = (const int*)(m + o_tile);
    L.tile_id = nullptr;
    B.tile_vol_c = (int*)(m + o_tvc);
    B.tile_id_c = (int*)(m + o_tidc);
    L.coef = (const Coef<double>*)(m + o_coef);
    L.st = (VolState<double>*)(m + o_st);
    L.part = (double*)(m + o_part);
    L.part_b = (double*)(m + o_partb);
    L.tab = (double*)(m + o_tab);
    B.dvol = (const VolDesc*)(m + o_vol);
    B.plane_vol = (int*)(m + o_pvol);
    B.plane_k = (int*)(m + o_pk);
    B.plane_out = (double*)(m + o_pout);
    float* fb = (float*)B.vecs;
    L.x = fb + 0 * total_c;
    L.r = fb + 1 * total_c;
    L.p[0] = fb + 2 * total_c;
    L.p[1] = fb + 3 * total_c;
    L.q = fb + 4 * total_c;
    L.xfull = (double*)B.vecs;
    double* db = (double*)((char*)B.vecs + f_bytes);
    L.y = db + 0 * total_c;
    L.x64 = db + 1 * total_c;
    L.x64b = db + 2 * total_c;
    uint8_t* codes = (uint8_t*)B.vecs + f_bytes + d_bytes;
    for (int i = 0; i < nv; ++i) {
        B.hrb[i].code_c[0] = codes + B.hrb[i].off_c;
        B.hrb[i].code_c[1] = codes + total_c + B.hrb[i].off_c;
    }
    e = hipMemcpyAsync(m + o_rb, B.hrb.data(), sizeof(RbDesc) * nv, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = ft::rb_setup(L, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // hrb is read by the copy
    return hip_ok(ctx, e, "diagonal tables") ? FAST_OK : FAST_E_HIP;
}

// CG with FP32 vectors and FP64 reliable updates on the Schur complement of the red voxels;
// the same update rule and FP64 convergence test as solve_mixed.  The FP64 residual it tests
// is that of the full system: the black equations hold exactly (x_b is eliminated in FP64),
// so the full residual is the Schur residual on the red voxels.
static int solve_mixed_rb(fast_ctx* ctx, std::vector<Job*>& jobs) {
    RbBatch B;
    B.ctx = ctx;
    B.jobs = jobs;
    int rc = prepare_rb(B);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    hipError_t e = hipEventRecord(ctx->t0, s);
    if (e == hipSuccess) e = ft::rb_update(B.L, s);
    if (!hip_ok(ctx, e, "mixed init")) return FAST_E_HIP;
    long maxit = 0;
    for (auto* j : jobs) maxit = std::max(maxit, j->prm.max_iter);
    const int chunk = ctx->chunk;
    const long max_chunks = 3 * (maxit / chunk + 2) + 16;
    auto round = [&]() -> hipError_t {
        hipError_t e2 = hipSuccess;
        for (int i = 0; i < chunk && e2 == hipSuccess; ++i) {
            e2 = ft::rb_iter(B.L, s);
            const bool upd = (i + 1) % ctx->update_every == 0 || i + 1 == chunk;
            if (e2 == hipSuccess && upd) e2 = ft::rb_update(B.L, s);
        }
        return e2;
    };
    auto step = [&]() -> hipError_t { return replay(B, round); };
    auto restart = [&]() -> hipError_t { return hipSuccess; };
    // Batch tails: once the unfinished volumes hold at most half of the launched tiles, later
    // rounds launch only theirs (a captured round is replaced by a new capture).
    const RbLaunch full = B.L;
    auto poll = [&](const VolState<double>* snap) -> hipError_t {
        long live = 0;
        for (size_t i = 0; i < jobs.size(); ++i)
            if (snap[i].status != ST_IDLE) live += B.hrb[i].ntiles;
        if (live == 0 || 2 * live > B.L.ntiles || B.L.ntiles < kRbCompactMin) return hipSuccess;
        int* h = (int*)pinned(ctx, 2, 2 * sizeof(int) * (size_t)live);
        if (!h) return hipErrorOutOfMemory;
        long k = 0;
        for (size_t i = 0; i < jobs.size(); ++i) {
            if (snap[i].status == ST_IDLE) continue;
            for (int t = 0; t < B.hrb[i].ntiles; ++t, ++k) {
                h[k] = (int)i;
                h[live + k] = B.hrb[i].tile_base + t;
            }
        }
        hipError_t e2 = hipMemcpyAsync(B.tile_vol_c, h, sizeof(int) * live, hipMemcpyHostToDevice, s);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(B.tile_id_c, h + live, sizeof(int) * live, hipMemcpyHostToDevice, s);
        B.L.tile_vol = B.tile_vol_c;
        B.L.tile_id = B.tile_id_c;
        B.L.ntiles = (int)live;
        if (B.gexec) {
            (void)hipGraphExecDestroy(B.gexec);
            B.gexec = nullptr;
        }
        return e2;
    };
    rc = drive(B, ST_IDLE, max_chunks, step, restart, poll);
    B.L = full;   // the scatter covers every volume
    if (rc) return rc;
    e = ft::rb_finish(B.L, s);
    if (!hip_ok(ctx, e, "scatter")) return FAST_E_HIP;
    std::vector<bool> ok(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const VolState<double>& st = B.hst[i];
        Job& j = *jobs[i];
        const double err = st.b64sq > 0 ? std::sqrt(st.r64sq / st.b64sq) : 0.0;
        const bool success = err <= j.outer_tol;
        ok[i] = success;
        j.res->rel_residual = err;
        j.res->true_rel_residual = err;
        j.res->iterations = st.iter;
        j.res->restarts = st.restarts;
        j.res->refinements = st.refinements;
        j.res->status = success ? FAST_SOLVED : FAST_NO_CONVERGENCE;
    }
    return finish<double>(B, B.dvol, B.L.xfull, B.L.coef, ok);
}

// Validate, fill defaults, group by (precision, solver, tile shape) and run each group as
// one batched solve.
int solve_jobs(fast_ctx* ctx, std::vector<Job>& jobs) {
    ctx->verbose = 0;
    for (auto& j : jobs) ctx->verbose |= j.prm.verbose;
    for (auto& j : jobs) {
        const fast_tau_params& p = j.prm;
        if (!j.vol || p.dir < 0 || p.dir > 5 || (p.solver != FAST_BICGSTAB && p.solver != FAST_CG) ||
            p.precision < FAST_FP64 || p.precision > FAST_MIXED || !(p.tolerance > 0) || p.max_iter < 0) {
            set_error(ctx, "invalid tau parameters");
            return FAST_E_BADARG;
        }
        const size_t n = (size_t)j.vol->nx * j.vol->ny * j.vol->nz;
        // porosity: precomputed value, else float(#zero / N) as fast/src/TauModule.h:62
        j.porosity = (p.porosity >= 0) ? p.porosity : (double)(float)((double)j.vol->zeros / (double)n);
        j.update_delta = (p.mixed_inner_tol > 0 && p.mixed_inner_tol < 1) ? p.mixed_inner_tol : 0.1;
        j.outer_tol = p.mixed_outer_tol > 0 ? p.mixed_outer_tol : p.tolerance;
        const Shape sh = choose_shape(j.vol->nx, j.vol->ny, j.vol->nz, p.precision, ctx);
        j.tz = sh.tz;
        j.shape = sh;
        std::memset(j.res, 0, sizeof(*j.res));
        j.res->true_rel_residual = -1;
    }
    // one batched solve per (precision, solver); its block shape is the one chosen for the
    // largest volume (results do not depend on the tile shape), each volume keeps its own tz
    std::map<std::tuple<int, int, int>, std::vector<Job*>> groups;
    for (auto& j : jobs) {
        const int rb = j.prm.precision == FAST_MIXED && j.prm.solver == FAST_CG && !j.prm.mixed_full_grid;
        groups[{j.prm.precision, j.prm.solver, rb}].push_back(&j);
    }
    int worst = FAST_OK;
    for (auto& g : groups) {
        const int prec = std::get<0>(g.first), solver = std::get<1>(g.first);
        if (std::get<2>(g.first)) {
            // A launch over volumes whose red-black state together fits the last-level cache
            // (MI355X: 256 MB Infinity Cache) runs faster per voxel than one over several large
            // volumes: the red-black iteration touches ~23 B per voxel, about the cache size at
            // 256^3 (3 x 256^3 as one launch: 257 ms; one after another: 223 ms).  A batch of
            // only large volumes (each at least half of ctx->group_bytes, so each still fills
            // the GPU) therefore runs as consecutive sub-batches of at most ctx->group_bytes in
            // submission order.  Batches with smaller volumes stay one launch: consecutive
            // sub-batches of those end in long low-occupancy tails (96 uneven volumes: 1475 ms
            // as one launch, 1720 ms in 320 MB sub-batches).  Results do not depend on the
            // grouping.
            auto state_bytes = [](const Job* j) {
                const fast_volume* v = j->vol;
                return 46 * (size_t)rb_row_pitch((v->nx + 1) / 2) * (size_t)(v->ny + 2) * (size_t)(v->nz + 2);
            };
            bool all_large = ctx->group_bytes > 0;
            for (const Job* j : g.second) all_large = all_large && 2 * state_bytes(j) >= ctx->group_bytes;
            std::vector<Job*> sub;
            size_t bytes = 0;
            for (Job* j : g.second) {
                const size_t b = state_bytes(j);
                if (all_large && !sub.empty() && bytes + b > ctx->group_bytes) {
                    const int rc = solve_mixed_rb(ctx, sub);
                    if (rc) return rc;
                    sub.clear();
                    bytes = 0;
                }
                sub.push_back(j);
                bytes += b;
            }
            const int rc = solve_mixed_rb(ctx, sub);
            if (rc) return rc;
            continue;
        }
        const Job* big = g.second[0];
        for (const Job* j : g.second)
            if ((size_t)j->vol->nx * j->vol->ny * j->vol->nz > (size_t)big->vol->nx * big->vol->ny * big->vol->nz) big = j;
        const Shape sh{big->shape.tx, big->shape.ty, 16};
        int rc;
        if (prec == FAST_FP64)
            rc = solve_plain<double, double>(ctx, g.second, sh, solver);
        else if (prec == FAST_FP32)
            rc = solve_plain<float, float>(ctx, g.second, sh, solver);
        else
            rc = solve_mixed(ctx, g.second, sh, solver);
        if (rc) return rc;
    }
    for (auto& j : jobs)
        if (j.res->status != FAST_SOLVED) worst = FAST_E_NOCONV;
    return worst;
}

}  // namespace fth

extern "C" {

int fast_tau_solve_volume(fast_volume* vol, const fast_tau_params* params, fast_tau_result* result, double* conc_out) {
    if (!vol || !params || !result) return FAST_E_BADARG;
    fast_ctx* ctx = vol->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    std::vector<Job> jobs(1);
    jobs[0].vol = vol;
    jobs[0].prm = *params;
    jobs[0].res = result;
    jobs[0].conc = conc_out;
    try {
        return solve_jobs(ctx, jobs);
    } catch (const std::bad_alloc&) {   // host staging buffers; device state is released by its owners
        set_error(ctx, "host allocation failed");
        return FAST_E_OOM;
    }
}

int fast_tau_solve(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz, const fast_tau_params* params,
                   fast_tau_result* result, double* conc_out) {
    if (!ctx || !params || !result) return FAST_E_BADARG;
    fast_volume* v = nullptr;
    int rc = fast_volume_create(ctx, mask, nx, ny, nz, &v);
    if (rc) {
        if (result) std::memset(result, 0, sizeof(*result));
        return rc;
    }
    rc = fast_tau_solve_volume(v, params, result, conc_out);
    fast_volume_destroy(v);
    return rc;
}

int fast_tau_solve_batch(fast_ctx* ctx, int count, fast_volume* const* vols, const fast_tau_params* params,
                         fast_tau_result* results) {
    if (!ctx || count < 0 || (count > 0 && (!vols || !params || !results))) return FAST_E_BADARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    std::vector<Job> jobs(count);
    for (int i = 0; i < count; ++i) {
        if (!vols[i] || vols[i]->ctx != ctx) return FAST_E_BADARG;
        jobs[i].vol = vols[i];
        jobs[i].prm = params[i];
        jobs[i].res = &results[i];
        jobs[i].conc = nullptr;
    }
    try {
        return solve_jobs(ctx, jobs);
    } catch (const std::bad_alloc&) {
        set_error(ctx, "host allocation failed");
        return FAST_E_OOM;
    }
}

}  // extern "C"

// ---- diagnostics ----
template <class T>
static int debug_apply(fast_volume* vol, int dir, const T* x, T* y) {
    if (!vol || !x || !y || dir < 0 || dir > 5) return FAST_E_BADARG;
    fast_ctx* ctx = vol->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    Job j;
    j.vol = vol;
    fast_tau_params_default(&j.prm);
    j.prm.dir = dir;
    fast_tau_result res;
    j.res = &res;
    const Shape sh = choose_shape(vol->nx, vol->ny, vol->nz, sizeof(T) == 8 ? FAST_FP64 : FAST_FP32, ctx);
    j.tz = sh.tz;
    Batch<T, T> B;
    B.ctx = ctx;
    B.jobs = {&j};
    B.shape = sh;
    int rc = prepare(B);
    if (rc) return rc;
    const size_t n = (size_t)vol->nx * vol->ny * vol->nz;
    T* dx = (T*)B.L.vs.x;
    T* dy = (T*)B.L.vs.s;
    hipError_t e = hipMemcpyAsync(dx, x, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = Passes<T, T>::apply(B.L, dx, dy, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return hip_ok(ctx, e, "debug apply") ? FAST_OK : FAST_E_HIP;
}

// Time the fused passes of one solver on one volume: set up the solve, run `iters`
// iterations with events around every pass (tolerance 0: nothing converges), and return the
// mean device time per launch of each pass and of the whole iteration (ms).
template <class T, class A>
static int time_passes(fast_volume* vol, int dir, int solver, int iters, double* ms) {
    fast_ctx* ctx = vol->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    Job j;
    j.vol = vol;
    fast_tau_params_default(&j.prm);
    j.prm.dir = dir;
    j.prm.solver = solver;
    j.prm.tolerance = 0.0;
    j.prm.max_iter = 1L << 40;
    fast_tau_result res;
    j.res = &res;
    const Shape sh = choose_shape(vol->nx, vol->ny, vol->nz, sizeof(T) == 8 ? FAST_FP64 : FAST_MIXED, ctx);
    j.tz = sh.tz;
    Batch<T, A> B;
    B.ctx = ctx;
    B.jobs = {&j};
    B.shape = sh;
    int rc = prepare(B);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const bool cg = (solver == FAST_CG);
    const int nev = cg ? 4 : 6;   // events around each stencil kernel
    std::vector<hipEvent_t> ev((size_t)iters * nev);
    for (auto& e : ev) (void)hipEventCreate(&e);
    hipError_t e = cg ? Passes<T, A>::init_cg(B.L, s) : Passes<T, A>::init_bicg(B.L, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = cg ? Passes<T, A>::cg_iter(B.L, s) : Passes<T, A>::bicg_iter(B.L, s);
    for (int i = 0; i < iters && e == hipSuccess; ++i)
        e = cg ? Passes<T, A>::cg_iter_timed(B.L, s, &ev[(size_t)i * nev])
               : Passes<T, A>::bicg_iter_timed(B.L, s, &ev[(size_t)i * nev]);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    double acc[4] = {0, 0, 0, 0};
    const int npass = nev / 2;
    if (e == hipSuccess) {
        for (int i = 0; i < iters; ++i) {
            float t;
            for (int k = 0; k < npass; ++k) {
                (void)hipEventElapsedTime(&t, ev[(size_t)i * nev + 2 * k], ev[(size_t)i * nev + 2 * k + 1]);
                acc[k] += t;
            }
            if (i + 1 < iters) {
                (void)hipEventElapsedTime(&t, ev[(size_t)i * nev], ev[(size_t)(i + 1) * nev]);
                acc[3] += t;
            }
        }
    }
    // Kernel durations: each stencil pass launched `iters` times back to back between two
    // events (no finalize in between, so every launch does the full pass), which keeps the
    // event overhead out of the per-launch average.
    for (int k = 0; k < npass && e == hipSuccess; ++k) {
        e = hipEventRecord(ev[0], s);
        for (int i = 0; i < iters && e == hipSuccess; ++i) e = Passes<T, A>::pass_only(B.L, solver, k, s);
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
        acc[k] = t;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (!hip_ok(ctx, e, "pass timing")) return FAST_E_HIP;
    for (int k = 0; k < 3; ++k) ms[k] = acc[k] / iters;
    ms[3] = iters > 1 ? acc[3] / (iters - 1) : 0.0;
    return FAST_OK;
}

// Red-black passes P, Y, Q, U: each launched `iters` times back to back between two events
// (ms[0..3], average per launch), then `iters` whole iterations (ms[4], average period).  The
// solve is held in its Krylov phase (no reliable update) so that every launch does full work.
static int time_rb_passes(fast_volume* vol, int dir, int iters, double* ms) {
    fast_ctx* ctx = vol->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    Job j;
    j.vol = vol;
    fast_tau_params_default(&j.prm);
    j.prm.dir = dir;
    j.prm.solver = FAST_CG;
    j.prm.precision = FAST_MIXED;
    j.prm.max_iter = 1L << 40;
    j.outer_tol = 1e-300;
    j.update_delta = 1e-150;
    fast_tau_result res;
    j.res = &res;
    RbBatch B;
    B.ctx = ctx;
    B.jobs = {&j};
    int rc = prepare_rb(B);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    hipEvent_t ev[2];
    (void)hipEventCreate(&ev[0]);
    (void)hipEventCreate(&ev[1]);
    hipError_t e = ft::rb_update(B.L, s);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = ft::rb_iter(B.L, s);
    for (int k = 0; k < 5 && e == hipSuccess; ++k) {
        e = hipEventRecord(ev[0], s);
        for (int i = 0; i < iters && e == hipSuccess; ++i) e = k == 4 ? ft::rb_iter(B.L, s) : ft::rb_pass_only(B.L, k, s);
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
        ms[k] = t / iters;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return hip_ok(ctx, e, "pass timing") ? FAST_OK : FAST_E_HIP;
}

extern "C" {

int fast_debug_time_passes(fast_volume* vol, int dir, int precision, int solver, int iters, double* ms) {
    if (!vol || !ms || iters < 1 || dir < 0 || dir > 5 || (solver != FAST_BICGSTAB && solver != FAST_CG))
        return FAST_E_BADARG;
    if (precision == FAST_FP64) return time_passes<double, double>(vol, dir, solver, iters, ms);
    if (precision == FAST_MIXED) return time_passes<float, double>(vol, dir, solver, iters, ms);
    if (precision == FAST_FP32) return time_passes<float, float>(vol, dir, solver, iters, ms);
    return FAST_E_BADARG;
}

int fast_debug_time_rb_passes(fast_volume* vol, int dir, int iters, double* ms) {
    if (!vol || !ms || iters < 1 || dir < 0 || dir > 5) return FAST_E_BADARG;
    return time_rb_passes(vol, dir, iters, ms);
}

int fast_debug_stream(fast_ctx* ctx, size_t bytes, int iters, double* ms) {
    if (!ctx || !ms || iters < 1 || bytes < 16) return FAST_E_BADARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    const size_t n4 = bytes / 16;
    float4* a = (float4*)dalloc(ctx, n4 * 16);
    float4* b = (float4*)dalloc(ctx, n4 * 16);
    float4* c = (float4*)dalloc(ctx, n4 * 16);
    int rc = (a && b && c) ? FAST_OK : FAST_E_OOM;
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    // cases: triad one float4 per thread, triad on 8192 striding blocks, copy, fill
    const int kinds[4] = {0, 0, 1, 2};
    const unsigned grids[4] = {0, 8192, 0, 0};
    if (!rc) e = ft::stream_kernel(2, a, nullptr, nullptr, n4, 0, s);
    if (e == hipSuccess && !rc) e = ft::stream_kernel(2, b, nullptr, nullptr, n4, 0, s);
    for (int k = 0; k < 4 && e == hipSuccess && !rc; ++k) {
        e = ft::stream_kernel(kinds[k], a, b, c, n4, grids[k], s);   // untimed first launch
        if (e == hipSuccess) e = hipEventRecord(ctx->t0, s);
        for (int i = 0; i < iters && e == hipSuccess; ++i) e = ft::stream_kernel(kinds[k], a, b, c, n4, grids[k], s);
        if (e == hipSuccess) e = hipEventRecord(ctx->t1, s);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->t1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->t0, ctx->t1);
        ms[k] = (double)t / iters;
    }
    if (!rc && !hip_ok(ctx, e, "stream timing")) rc = FAST_E_HIP;
    (void)hipStreamSynchronize(s);
    dfree(ctx, c);
    dfree(ctx, b);
    dfree(ctx, a);
    return rc;
}

int fast_debug_apply_f64(fast_volume* vol, int dir, const double* x, double* y) { return debug_apply<double>(vol, dir, x, y); }
int fast_debug_apply_f32(fast_volume* vol, int dir, const float* x, float* y) { return debug_apply<float>(vol, dir, x, y); }

int fast_debug_system_f64(fast_volume* vol, int dir, double* b, double* x0, double* invdiag) {
    if (!vol || dir < 0 || dir > 5) return FAST_E_BADARG;
    fast_ctx* ctx = vol->ctx;
    const size_t n = (size_t)vol->nx * vol->ny * vol->nz;
    // b, x0 and 1/diag are what the init pass writes: run it for BiCGSTAB with one iteration
    // cap of zero and read r (= b - A x0) and x (= x0) back; b follows from A x0.
    std::vector<double> xx(n), ax(n), rr(n);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        Job j;
        j.vol = vol;
        fast_tau_params_default(&j.prm);
        j.prm.dir = dir;
        j.prm.solver = FAST_CG;
        fast_tau_result res;
        j.res = &res;
        const Shape sh = choose_shape(vol->nx, vol->ny, vol->nz, FAST_FP64, ctx);
        j.tz = sh.tz;
        Batch<double, double> B;
        B.ctx = ctx;
        B.jobs = {&j};
        B.shape = sh;
        int rc = prepare(B);
        if (rc) return rc;
        hipError_t e = Passes<double, double>::init_cg(B.L, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(xx.data(), B.L.vs.x, n * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rr.data(), B.L.vs.r, n * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (!hip_ok(ctx, e, "debug system")) return FAST_E_HIP;
    }
    int rc = fast_debug_apply_f64(vol, dir, xx.data(), ax.data());
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        if (x0) x0[i] = xx[i];
        if (b) b[i] = rr[i] + ax[i];
    }
    if (invdiag) {
        // 1/diag: apply A to unit vectors would be O(N^2); the diagonal is recovered from
        // A e over a 2x2x2 colouring (each colour class has no two voxels sharing a row).
        std::vector<double> e(n), ae(n);
        for (int c = 0; c < 8; ++c) {
            for (int z = 0; z < vol->nz; ++z)
                for (int y = 0; y < vol->ny; ++y)
                    for (int x = 0; x < vol->nx; ++x) {
                        const size_t i = (size_t)x + (size_t)vol->nx * (y + (size_t)vol->ny * z);
                        e[i] = (((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)) == c) ? 1.0 : 0.0;
                    }
            rc = fast_debug_apply_f64(vol, dir, e.data(), ae.data());
            if (rc) return rc;
            for (size_t i = 0; i < n; ++i)
                if (e[i] != 0.0) invdiag[i] = 1.0 / ae[i];
        }
    }
    return FAST_OK;
}

}  // extern "C"
