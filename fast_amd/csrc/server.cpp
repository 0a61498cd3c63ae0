// Continuous-batching request loop for tau solves (no reference counterpart: the reference
// solves one volume per direction at a time, fast/src/TauModule.h:64-114, and its only
// many-volume loop is sequential, fast/src/SliceModule.h:126-133).
//
// A worker thread owns the GPU.  Requests wait in a FIFO; at every round boundary the worker
// admits queued requests while a slot and room in the device arena remain (admission by HBM
// budget), uploads their masks and builds their stencil codes, and adds their tiles to the
// launch.  Every round runs `iters_per_round` iterations of the mixed-form
// kernels over all live volumes at once (CG on the red-black Schur complement, or BiCGSTAB on
// the full grid); each volume carries its own scalars and status, so volumes join and leave at
// round boundaries without disturbing the others.  Finished volumes get their flux integral
// and tau, release their memory, and are handed to fast_server_wait.  All dot products are
// double-double (tile- and batch-invariant) and each volume's tiling depends on its own
// dimensions only, so a request's result is bit-identical to solving it alone.
#include "fast_tau.h"
#include "tau_host.h"
#include "tau_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

using namespace ft;
using namespace fth;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kTX = 64, kTY = 8;      // one launch shape for the whole live set (full grid)
constexpr size_t kAlign = 64;         // arena granularity (elements)
constexpr size_t kPlanes = 4;         // full grid: FP32 vectors x (the correction), r, r0, s
constexpr size_t kDirPlanes = 4;      // full grid: FP64 directions p[2], v[2] (BiCGSTAB's, see DirOf)
// red-black: per padded per-colour element, FP32 x, r, p[2], q, FP64 y, x64, x64b and two code bytes
constexpr size_t kRbBytes = 5 * sizeof(float) + 3 * sizeof(double) + 2;

struct Req {
    uint64_t ticket = 0;
    std::vector<uint8_t> mask;
    int nx = 0, ny = 0, nz = 0;
    fast_tau_params prm{};
    double t_submit = 0, t_done = 0;
    // live state
    uint8_t* d_mask = nullptr;
    uint8_t* d_code = nullptr;
    int slot = -1;
    size_t off = 0, elems = 0;   // arena range (voxels, or padded per-colour elements for red-black)
    int tz = 16;
    fast_tau_result res{};
    int rc = FAST_OK;
};

// First-fit allocator of element ranges in the arena planes.
struct Arena {
    size_t cap = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    void init(size_t c) {
        cap = c;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(size_t n, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o = it->first + n;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct fast_server {
    fast_ctx* ctx = nullptr;
    fast_server_config cfg{};
    int solver = FAST_CG;
    bool rb = true;                           // CG on the red-black Schur complement
    fth::RbShape rbs{4, 0, 2};                // red-black rows per tile and prefetch (tz per request)
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_in, cv_out;
    std::deque<Req*> queue;
    std::deque<Req*> done;
    bool stop = false;
    uint64_t next_ticket = 1;
    int pending = 0;
    int fatal = FAST_OK;

    hipStream_t up = nullptr;                 // mask uploads
    hipEvent_t up_done = nullptr;
    void* arena_mem = nullptr;
    size_t S = 0;                             // elements per arena plane
    Arena arena;
    int max_active = 64;
    std::vector<Req*> slots;
    VolDesc* d_vol = nullptr;
    Coef<double>* d_coef = nullptr;
    VolState<double>* d_st = nullptr;
    unsigned long long* d_zero = nullptr;     // per-slot zero-voxel counters
    double* d_tab = nullptr;                  // per-slot diagonal tables
    int* d_tile = nullptr;
    double* d_part = nullptr;
    int* d_pvol = nullptr;
    int* d_pk = nullptr;
    double* d_pout = nullptr;
    size_t tile_cap = 0, plane_cap = 0;
    std::vector<VolDesc> h_vol;
    VolState<double>* h_st = nullptr;         // pinned snapshot
    Launch<float, double> L{};
    // red-black form
    RbDesc* d_rb = nullptr;
    std::vector<RbDesc> h_rb;
    double* d_partb = nullptr;
    int* d_tile_fin = nullptr;                // tiles of the finishing slots (scatter)
    size_t tile_fin_cap = 0;
    double* d_xfull = nullptr;                // full-grid solutions of the finishing slots
    size_t xfull_cap = 0;
    RbLaunch RL{};
    int ntiles = 0;
};

namespace {

bool grow(fast_server* s, void** p, size_t* cap, size_t need, size_t elem) {
    if (*cap >= need) return true;
    dfree(s->ctx, *p);
    size_t c = std::max(need, *cap * 2);
    *p = dalloc(s->ctx, c * elem);
    *cap = *p ? c : 0;
    return *p != nullptr;
}

float* rb_f32(fast_server* s, int k) { return (float*)s->arena_mem + (size_t)k * s->S; }
double* rb_f64(fast_server* s, int k) { return (double*)((float*)s->arena_mem + 5 * s->S) + (size_t)k * s->S; }
uint8_t* rb_u8(fast_server* s, int k) { return (uint8_t*)(rb_f64(s, 0) + 3 * s->S) + (size_t)k * s->S; }

// Upload masks and stencil codes of newly admitted requests on the upload stream.
int upload(fast_server* s, const std::vector<Req*>& adm) {
    fast_ctx* ctx = s->ctx;
    for (Req* r : adm) {
        const size_t n = (size_t)r->nx * r->ny * r->nz;
        r->d_mask = (uint8_t*)dalloc(ctx, n);
        r->d_code = (uint8_t*)dalloc(ctx, n);
        if (!r->d_mask || !r->d_code) return FAST_E_OOM;
        hipError_t e = hipMemcpyAsync(r->d_mask, r->mask.data(), n, hipMemcpyHostToDevice, s->up);
        if (e == hipSuccess) e = hipMemsetAsync(s->d_zero + r->slot, 0, sizeof(unsigned long long), s->up);
        if (e == hipSuccess) e = ft::build_code(r->d_mask, r->d_code, r->nx, r->ny, r->nz, s->d_zero + r->slot, s->up);
        // red-black: ghost and padding elements of the request's range must read +0
        for (int k = 0; s->rb && k < 5 && e == hipSuccess; ++k)
            e = hipMemsetAsync(rb_f32(s, k) + r->off, 0, r->elems * sizeof(float), s->up);
        for (int k = 0; s->rb && k < 3 && e == hipSuccess; ++k)
            e = hipMemsetAsync(rb_f64(s, k) + r->off, 0, r->elems * sizeof(double), s->up);
        for (int k = 0; s->rb && k < 2 && e == hipSuccess; ++k) e = hipMemsetAsync(rb_u8(s, k) + r->off, 0, r->elems, s->up);
        if (!hip_ok(ctx, e, "server upload")) return FAST_E_HIP;
    }
    hipError_t e = hipEventRecord(s->up_done, s->up);
    if (e == hipSuccess) e = hipStreamSynchronize(s->up);   // host mask copies may be released after this
    if (!hip_ok(ctx, e, "server upload")) return FAST_E_HIP;
    for (Req* r : adm) {
        r->mask.clear();
        r->mask.shrink_to_fit();
    }
    return FAST_OK;
}

// Initial state of a newly admitted slot (every request starts with the FP64 residual step).
VolState<double> fresh_state(const fast_server* s, const Req* r) {
    VolState<double> v{};
    v.status = ST_FRESH;
    v.mode = 1;
    v.solver = s->solver;
    v.tol = r->prm.tolerance;
    v.maxiter = r->prm.max_iter;
    v.otol = r->prm.mixed_outer_tol > 0 ? r->prm.mixed_outer_tol : r->prm.tolerance;
    const double delta = (r->prm.mixed_inner_tol > 0 && r->prm.mixed_inner_tol < 1) ? r->prm.mixed_inner_tol : 0.1;
    v.delta2 = delta * delta;
    v.restart_cos = s->ctx->mixed_restart_cos;
    v.blowup2 = s->ctx->mixed_blowup2;
    v.eps2 = s->rb ? 2.220446049250313e-16 * 2.220446049250313e-16 : 1.1920928955078125e-07 * 1.1920928955078125e-07;
    return v;
}

// Red-black form: descriptors and tile list of the live set (tile_base as the full launch
// sees it), the per-slot states and coefficients of newly admitted requests, tables and codes.
int rebuild_rb(fast_server* s, const std::vector<Req*>& adm) {
    fast_ctx* ctx = s->ctx;
    int tiles = 0;
    std::vector<int> tv;
    for (int i = 0; i < s->max_active; ++i) {
        RbDesc& d = s->h_rb[i];
        Req* r = s->slots[i];
        d = RbDesc{};
        d.tile_base = tiles;
        if (!r) continue;
        d.nx = r->nx;
        d.ny = r->ny;
        d.nz = r->nz;
        d.hx = (r->nx + 1) / 2;
        d.px = rb_row_pitch(d.hx);
        d.pp = (unsigned)d.px * (unsigned)(r->ny + 2);
        d.dir = r->prm.dir;
        d.primary = r->prm.dir / 2;
        d.sgn = (r->prm.dir % 2 == 0) ? 1 : -1;
        d.tiles_x = (d.hx + kRbTX - 1) / kRbTX;
        d.tiles_y = (r->ny + s->rbs.ty - 1) / s->rbs.ty;
        d.tz = r->tz;
        d.tiles_z = (r->nz + d.tz - 1) / d.tz;
        d.ntiles = d.tiles_x * d.tiles_y * d.tiles_z;
        d.off_c = (long long)r->off;
        d.code = r->d_code;
        d.code_c[0] = rb_u8(s, 0) + r->off;
        d.code_c[1] = rb_u8(s, 1) + r->off;
        for (int t = 0; t < d.ntiles; ++t) tv.push_back(i);
        tiles += d.ntiles;
    }
    if (!grow(s, (void**)&s->d_tile, &s->tile_cap, std::max(tiles, 1), sizeof(int))) return FAST_E_OOM;
    dfree(ctx, s->d_part);   // partials: 8 doubles per tile of the tile capacity, red and black
    dfree(ctx, s->d_partb);
    s->d_part = (double*)dalloc(ctx, s->tile_cap * 8 * sizeof(double));
    s->d_partb = (double*)dalloc(ctx, s->tile_cap * 8 * sizeof(double));
    if (!s->d_part || !s->d_partb) return FAST_E_OOM;
    hipStream_t st = ctx->stream;
    hipError_t e = hipStreamWaitEvent(st, s->up_done, 0);
    if (e == hipSuccess && tiles)
        e = hipMemcpyAsync(s->d_tile, tv.data(), sizeof(int) * tiles, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_rb, s->h_rb.data(), sizeof(RbDesc) * s->max_active, hipMemcpyHostToDevice, st);
    for (Req* r : adm) {
        if (e != hipSuccess) break;
        const Coef<double> c = make_coef<double>(r->nx, r->ny, r->nz, r->prm.dir, r->prm.d_pore, r->prm.d_solid);
        const VolState<double> v = fresh_state(s, r);
        e = hipMemcpyAsync(s->d_coef + r->slot, &c, sizeof(c), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s->d_st + r->slot, &v, sizeof(v), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // c and v live on this stack frame
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);       // tv and h_rb are read by the copies
    if (!hip_ok(ctx, e, "server rebuild")) return FAST_E_HIP;
    RbLaunch& L = s->RL;
    L.ntiles = tiles;
    L.nvol = s->max_active;
    L.staged = rb_mode();
    L.fold_alpha = 0;
    L.ty = s->rbs.ty;
    L.ahead = s->rbs.ahead;
    L.vol = s->d_rb;
    L.tile_vol = s->d_tile;
    L.tile_id = nullptr;   // tiles are listed in batch order
    L.coef = s->d_coef;
    L.st = s->d_st;
    L.part = s->d_part;
    L.part_b = s->d_partb;
    L.tab = s->d_tab;
    L.x = rb_f32(s, 0);
    L.r = rb_f32(s, 1);
    L.p[0] = rb_f32(s, 2);
    L.p[1] = rb_f32(s, 3);
    L.q = rb_f32(s, 4);
    L.y = rb_f64(s, 0);
    L.x64 = rb_f64(s, 1);
    L.x64b = rb_f64(s, 2);
    L.xfull = s->d_xfull;
    s->ntiles = tiles;
    if (!adm.empty()) {
        const hipError_t et = ft::rb_setup(L, ctx->stream);   // tables of every slot, codes of the live set
        if (!hip_ok(ctx, et, "server tables")) return FAST_E_HIP;
    }
    return FAST_OK;
}

// Descriptors and tile list of the live set; states of newly admitted slots.
int rebuild(fast_server* s, const std::vector<Req*>& adm) {
    if (s->rb) return rebuild_rb(s, adm);
    fast_ctx* ctx = s->ctx;
    int tiles = 0;
    std::vector<int> tv;
    for (int i = 0; i < s->max_active; ++i) {
        VolDesc& d = s->h_vol[i];
        Req* r = s->slots[i];
        if (!r) {
            d.ntiles = 0;
            d.tile_base = tiles;
            continue;
        }
        d.nx = r->nx;
        d.ny = r->ny;
        d.nz = r->nz;
        d.dir = r->prm.dir;
        d.primary = r->prm.dir / 2;
        d.sgn = (r->prm.dir % 2 == 0) ? 1 : -1;
        d.tiles_x = (r->nx + kTX - 1) / kTX;
        d.tiles_y = (r->ny + kTY - 1) / kTY;
        d.tz = r->tz;
        d.tiles_z = (r->nz + d.tz - 1) / d.tz;
        d.tile_base = tiles;
        d.ntiles = d.tiles_x * d.tiles_y * d.tiles_z;
        d.plane_base = 0;
        d.flux_nu = r->prm.flux_nonuniform ? 1 : 0;
        d.off = (long long)r->off;
        d.code = r->d_code;
        for (int t = 0; t < d.ntiles; ++t) tv.push_back(i);
        tiles += d.ntiles;
    }
    if (!grow(s, (void**)&s->d_tile, &s->tile_cap, std::max(tiles, 1), sizeof(int))) return FAST_E_OOM;
    dfree(ctx, s->d_part);   // partials: 8 doubles per tile of the tile capacity
    s->d_part = (double*)dalloc(ctx, s->tile_cap * 8 * sizeof(double));
    if (!s->d_part) return FAST_E_OOM;
    hipStream_t st = ctx->stream;
    hipError_t e = hipStreamWaitEvent(st, s->up_done, 0);
    if (e == hipSuccess && tiles)
        e = hipMemcpyAsync(s->d_tile, tv.data(), sizeof(int) * tiles, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_vol, s->h_vol.data(), sizeof(VolDesc) * s->max_active, hipMemcpyHostToDevice, st);
    for (Req* r : adm) {
        if (e != hipSuccess) break;
        const Coef<double> c = make_coef<double>(r->nx, r->ny, r->nz, r->prm.dir, r->prm.d_pore, r->prm.d_solid);
        const VolState<double> v = fresh_state(s, r);
        e = hipMemcpyAsync(s->d_coef + r->slot, &c, sizeof(c), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s->d_st + r->slot, &v, sizeof(v), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // c and v live on this stack frame
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // tv and h_vol are read by the copies
    if (!hip_ok(ctx, e, "server rebuild")) return FAST_E_HIP;
    s->ntiles = tiles;
    Launch<float, double>& L = s->L;
    L.tx = kTX;
    L.ty = kTY;
    L.ntiles = tiles;
    L.nvol = s->max_active;
    L.vol = s->d_vol;
    L.tile_vol = s->d_tile;
    L.coef = s->d_coef;
    L.st = s->d_st;
    L.part = s->d_part;
    L.tab = s->d_tab;
    float* a = (float*)s->arena_mem;
    const size_t S = s->S;
    L.vs.x = a + 0 * S;
    L.vs.r = a + 1 * S;
    L.vs.r0 = a + 2 * S;
    L.vs.s = a + 3 * S;
    double* d = (double*)(a + kPlanes * S);
    L.vs.p[0] = d + 0 * S;
    L.vs.p[1] = d + 1 * S;
    L.vs.v[0] = d + 2 * S;
    L.vs.v[1] = d + 3 * S;
    L.vs.x64 = d + kDirPlanes * S;
    if (!adm.empty()) {
        const hipError_t et = ft::build_tables(L, ctx->stream);   // descriptors and coefficients are on the device
        if (!hip_ok(ctx, et, "server tables")) return FAST_E_HIP;
    }
    return FAST_OK;
}

// Flux and tau of the finished slots; release them.
int retire(fast_server* s, const std::vector<int>& fin) {
    fast_ctx* ctx = s->ctx;
    // red-black: both colours of each finishing slot go into a full-grid vector first; the
    // scatter runs over the finishing slots' tiles only (tile_base and the full-grid offset
    // as that launch sees them; the next rebuild restores the descriptors)
    std::vector<int> ft_tiles;
    size_t xn = 0;
    if (s->rb) {
        for (int i : fin) {
            RbDesc& d = s->h_rb[i];
            VolDesc& v = s->h_vol[i];
            v = VolDesc{};
            v.nx = d.nx;
            v.ny = d.ny;
            v.nz = d.nz;
            v.dir = d.dir;
            v.primary = d.primary;
            v.sgn = d.sgn;
            v.flux_nu = s->slots[i]->prm.flux_nonuniform ? 1 : 0;
            v.off = (long long)xn;
            v.code = s->slots[i]->d_code;
            d.off = (long long)xn;
            d.tile_base = (int)ft_tiles.size();
            for (int t = 0; t < d.ntiles; ++t) ft_tiles.push_back(i);
            xn += align_up((size_t)d.nx * d.ny * d.nz, kAlign);
        }
    }
    std::vector<int> pv, pk;
    for (int i : fin) {
        const VolDesc& d = s->h_vol[i];
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        for (int k = 0; k < np; ++k) {
            pv.push_back(i);
            pk.push_back(k);
        }
    }
    const size_t np_all = pv.size();
    size_t cap = s->plane_cap;
    if (cap < np_all) {
        dfree(ctx, s->d_pvol);
        dfree(ctx, s->d_pk);
        dfree(ctx, s->d_pout);
        cap = std::max(np_all, 2 * cap);
        s->d_pvol = (int*)dalloc(ctx, cap * sizeof(int));
        s->d_pk = (int*)dalloc(ctx, cap * sizeof(int));
        s->d_pout = (double*)dalloc(ctx, cap * sizeof(double));
        if (!s->d_pvol || !s->d_pk || !s->d_pout) return FAST_E_OOM;
        s->plane_cap = cap;
    }
    std::vector<double> planes(np_all);
    std::vector<unsigned long long> zeros(s->max_active);
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    const double* conc = s->L.vs.x64;
    if (s->rb) {
        if (!grow(s, (void**)&s->d_xfull, &s->xfull_cap, std::max<size_t>(xn, 1), sizeof(double)) ||
            !grow(s, (void**)&s->d_tile_fin, &s->tile_fin_cap, std::max<size_t>(ft_tiles.size(), 1), sizeof(int)))
            return FAST_E_OOM;
        RbLaunch F = s->RL;
        F.ntiles = (int)ft_tiles.size();
        F.tile_vol = s->d_tile_fin;
        F.xfull = s->d_xfull;
        e = hipMemcpyAsync(s->d_rb, s->h_rb.data(), sizeof(RbDesc) * s->max_active, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && F.ntiles)
            e = hipMemcpyAsync(s->d_tile_fin, ft_tiles.data(), sizeof(int) * ft_tiles.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->d_vol, s->h_vol.data(), sizeof(VolDesc) * s->max_active, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && F.ntiles) e = ft::rb_finish(F, st);
        conc = s->d_xfull;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_pvol, pv.data(), np_all * sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_pk, pk.data(), np_all * sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = ft::flux_planes<double>((int)np_all, s->d_vol, s->d_pvol, s->d_pk, s->d_coef, conc, s->d_pout, st);
    if (e == hipSuccess) e = hipMemcpyAsync(planes.data(), s->d_pout, np_all * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(zeros.data(), s->d_zero, s->max_active * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (!hip_ok(ctx, e, "server flux")) return FAST_E_HIP;
    size_t pos = 0;
    const double t = now_s();
    std::vector<Req*> out;
    out.reserve(fin.size());   // no allocation between taking a request out of its slot and listing it
    for (int i : fin) {
        Req* r = s->slots[i];
        const VolState<double>& q = s->h_st[i];
        const VolDesc& d = s->h_vol[i];
        const int np = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
        double sum = 0;
        for (int k = 0; k < np; ++k) sum += planes[pos + k];
        pos += np;
        const double dc = sum / np;
        const size_t n = (size_t)r->nx * r->ny * r->nz;
        const double por = r->prm.porosity >= 0 ? r->prm.porosity : (double)(float)((double)zeros[i] / (double)n);
        const double err = q.b64sq > 0 ? std::sqrt(q.r64sq / q.b64sq) : 0.0;
        const double otol = r->prm.mixed_outer_tol > 0 ? r->prm.mixed_outer_tol : r->prm.tolerance;
        const bool ok = err <= otol;
        fast_tau_result& res = r->res;
        res.porosity = por;
        res.dc = dc;
        res.tau = ok ? por / (dc * np) : 0.0;
        res.rel_residual = err;
        res.true_rel_residual = err;
        res.iterations = q.iter;
        res.restarts = q.restarts;
        res.refinements = q.refinements;
        res.status = ok ? FAST_SOLVED : FAST_NO_CONVERGENCE;
        r->rc = ok ? FAST_OK : FAST_E_NOCONV;
        r->t_done = t;
        res.solve_ms = 1e3 * (t - r->t_submit);
        dfree(ctx, r->d_mask);
        dfree(ctx, r->d_code);
        r->d_mask = r->d_code = nullptr;
        out.push_back(r);
        s->arena.release(r->off, r->elems);
        s->slots[i] = nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Req* r : out) s->done.push_back(r);
    }
    s->cv_out.notify_all();
    return FAST_OK;
}

// Fails every admitted and queued request with `rc`.  Admitted requests give back their device
// mask and code and their arena range first (after the stream has drained, so no kernel still
// reads them); the worker stops after this.
void fail_all(fast_server* s, int rc) {
    fast_ctx* ctx = s->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(s->mu);
    s->fatal = rc;
    auto finish = [&](Req* r) {
        r->rc = rc;
        r->res.status = FAST_FAILED;
        r->t_done = now_s();
        s->done.push_back(r);
    };
    for (Req*& r : s->slots)
        if (r) {
            dfree(ctx, r->d_mask);
            dfree(ctx, r->d_code);
            r->d_mask = r->d_code = nullptr;
            try {
                s->arena.release(r->off, r->elems);
            } catch (const std::bad_alloc&) {   // the arena is not used again once the server has failed
            }
            finish(r);
            r = nullptr;
        }
    while (!s->queue.empty()) {
        finish(s->queue.front());
        s->queue.pop_front();
    }
    s->cv_out.notify_all();
}

void worker_main(fast_server* s) {
    fast_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    const int m = s->cfg.iters_per_round;
    for (;;) {
        std::vector<Req*> adm;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            const bool live = std::any_of(s->slots.begin(), s->slots.end(), [](Req* r) { return r != nullptr; });
            s->cv_in.wait(lk, [&] { return s->stop || !s->queue.empty() || live; });
            if (s->stop && s->queue.empty() && !live) return;
            // admission: FIFO, while a slot and arena room remain
            while (!s->queue.empty()) {
                Req* r = s->queue.front();
                auto it = std::find(s->slots.begin(), s->slots.end(), nullptr);
                if (it == s->slots.end()) break;
                size_t off;
                if (!s->arena.alloc(r->elems, &off)) break;
                r->off = off;
                r->slot = (int)(it - s->slots.begin());
                *it = r;
                s->queue.pop_front();
                adm.push_back(r);
            }
        }
        if (!adm.empty()) {
            int rc = upload(s, adm);
            if (rc == FAST_OK) rc = rebuild(s, adm);
            if (rc != FAST_OK) {
                fail_all(s, rc);
                return;
            }
        }
        if (s->ntiles == 0) continue;
        // one round over the live set
        hipStream_t st = ctx->stream;
        hipError_t e = hipSuccess;
        for (int i = 0; i < m && e == hipSuccess; ++i) {
            const bool upd = (i + 1) % ctx->update_every == 0 || i + 1 == m;
            if (s->rb) {
                e = ft::rb_iter(s->RL, st);
                if (e == hipSuccess && upd) e = ft::rb_update(s->RL, st);
                continue;
            }
            e = (s->solver == FAST_BICGSTAB) ? Passes<float, double>::bicg_iter(s->L, st)
                                             : Passes<float, double>::cg_iter(s->L, st);
            if (e == hipSuccess && upd) e = ft::mixed_update(s->L, st);
            if (e == hipSuccess && upd) e = ft::mixed_residual(s->L, s->d_coef, st);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->h_st, s->d_st, sizeof(VolState<double>) * s->max_active, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (!hip_ok(ctx, e, "server round")) {
            fail_all(s, FAST_E_HIP);
            return;
        }
        std::vector<int> fin;
        for (int i = 0; i < s->max_active; ++i)
            if (s->slots[i] && s->h_st[i].status == ST_IDLE) fin.push_back(i);
        if (!fin.empty()) {
            int rc = retire(s, fin);
            if (rc == FAST_OK) rc = rebuild(s, {});
            if (rc != FAST_OK) {
                fail_all(s, rc);
                return;
            }
        }
    }
}

}  // namespace

extern "C" {

int fast_server_create(fast_ctx* ctx, const fast_server_config* cfg, fast_server** out) {
    if (!ctx || !out) return FAST_E_BADARG;
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    auto* s = new fast_server();
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_active <= 0) s->cfg.max_active = 64;
    if (s->cfg.iters_per_round <= 0) s->cfg.iters_per_round = 8;
    s->solver = s->cfg.use_bicgstab ? FAST_BICGSTAB : FAST_CG;
    s->rb = (s->solver == FAST_CG);
    // one launch set serves every live request: the rows per tile and prefetch distance of the
    // 256^3 size class (the middle of the request sizes), tz per request
    s->rbs = rb_shape(ctx, 256, 256, 256);
    s->max_active = s->cfg.max_active;
    const size_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : std::min(ctx->budget / 2, (size_t)16 << 30);
    // bytes per arena element: red-black, kRbBytes per padded per-colour element; full grid,
    // 8 FP32 planes + the FP64 solution per voxel
    const size_t per = s->rb ? kRbBytes : kPlanes * sizeof(float) + (kDirPlanes + 1) * sizeof(double);
    s->S = align_up(budget / per, kAlign);
    if (s->S * per > budget && s->S > kAlign) s->S -= kAlign;
    s->arena_mem = dalloc(ctx, s->S * per);
    s->d_vol = (VolDesc*)dalloc(ctx, sizeof(VolDesc) * s->max_active);
    s->d_coef = (Coef<double>*)dalloc(ctx, sizeof(Coef<double>) * s->max_active);
    s->d_st = (VolState<double>*)dalloc(ctx, sizeof(VolState<double>) * s->max_active);
    s->d_zero = (unsigned long long*)dalloc(ctx, sizeof(unsigned long long) * s->max_active);
    s->d_tab = (double*)dalloc(ctx, sizeof(double) * 2 * kTabEntries * s->max_active);
    s->d_rb = (RbDesc*)dalloc(ctx, sizeof(RbDesc) * s->max_active);
    if (!s->arena_mem || !s->d_vol || !s->d_coef || !s->d_st || !s->d_zero || !s->d_tab || !s->d_rb ||
        hipHostMalloc((void**)&s->h_st, sizeof(VolState<double>) * s->max_active, hipHostMallocDefault) != hipSuccess ||
        hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&s->up_done) != hipSuccess) {
        (void)hipGetLastError();
        fast_server_destroy(s);
        return FAST_E_OOM;
    }
    s->arena.init(s->S);
    s->slots.assign(s->max_active, nullptr);
    s->h_vol.assign(s->max_active, VolDesc{});
    s->h_rb.assign(s->max_active, RbDesc{});
    // every slot starts idle
    std::vector<VolState<double>> idle(s->max_active);
    for (auto& v : idle) v.status = ST_IDLE;
    hipError_t e = hipMemcpy(s->d_st, idle.data(), sizeof(VolState<double>) * s->max_active, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventRecord(s->up_done, s->up);
    if (!hip_ok(ctx, e, "server init")) {
        fast_server_destroy(s);
        return FAST_E_HIP;
    }
    s->worker = std::thread([s] {
        try {
            worker_main(s);
        } catch (const std::bad_alloc&) {   // host bookkeeping: fail the requests, keep the process
            fail_all(s, FAST_E_OOM);
        }
    });
    *out = s;
    return FAST_OK;
}

int fast_server_submit(fast_server* s, const uint8_t* mask, int nx, int ny, int nz, const fast_tau_params* params,
                       uint64_t* ticket) {
    if (!s || !mask || !params || !ticket || nx <= 0 || ny <= 0 || nz <= 0) return FAST_E_BADARG;
    if (params->dir < 0 || params->dir > 5 || !(params->tolerance > 0) || params->max_iter < 0) return FAST_E_BADARG;
    const size_t n = (size_t)nx * ny * nz;
    if (n > kMaxSolveVoxels) return FAST_E_BADARG;
    Req* r = new (std::nothrow) Req();
    if (!r) return FAST_E_OOM;
    try {
        r->mask.assign(mask, mask + n);   // the request keeps its own copy until upload
    } catch (const std::bad_alloc&) {
        delete r;
        return FAST_E_OOM;
    }
    r->nx = nx;
    r->ny = ny;
    r->nz = nz;
    r->prm = *params;
    if (s->rb) {   // padded per-colour elements
        const int hx = (nx + 1) / 2;
        r->elems = align_up((size_t)rb_row_pitch(hx) * (size_t)(ny + 2) * (size_t)(nz + 2), kAlign);
        r->tz = rb_shape(s->ctx, nx, ny, nz).tz;
    } else {
        r->elems = align_up(n, kAlign);
        r->tz = choose_shape(nx, ny, nz, FAST_MIXED, s->ctx).tz;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal != FAST_OK) {
            delete r;
            return s->fatal;
        }
        if (r->elems > s->S) {
            delete r;
            return FAST_E_OOM;   // larger than the whole arena: can never be admitted
        }
        r->ticket = s->next_ticket++;
        r->t_submit = now_s();
        s->queue.push_back(r);
        ++s->pending;
        *ticket = r->ticket;
    }
    s->cv_in.notify_one();
    return FAST_OK;
}

int fast_server_wait(fast_server* s, int timeout_ms, uint64_t* ticket, fast_tau_result* result, double* latency_ms) {
    if (!s) return FAST_E_BADARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done.empty(); };
    if (timeout_ms < 0)
        s->cv_out.wait(lk, ready);
    else if (!s->cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return FAST_E_BUSY;
    Req* r = s->done.front();
    s->done.pop_front();
    --s->pending;
    lk.unlock();
    if (ticket) *ticket = r->ticket;
    if (result) *result = r->res;
    if (latency_ms) *latency_ms = 1e3 * (r->t_done - r->t_submit);
    const int rc = r->rc;
    delete r;
    return rc;
}

int fast_server_pending(fast_server* s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    return s->pending;
}

void fast_server_destroy(fast_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_in.notify_all();
    if (s->worker.joinable()) s->worker.join();
    fast_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (Req* r : s->slots)
        if (r) {
            dfree(ctx, r->d_mask);
            dfree(ctx, r->d_code);
            delete r;
        }
    for (Req* r : s->queue) delete r;
    for (Req* r : s->done) delete r;
    dfree(ctx, s->arena_mem);
    dfree(ctx, s->d_vol);
    dfree(ctx, s->d_coef);
    dfree(ctx, s->d_st);
    dfree(ctx, s->d_zero);
    dfree(ctx, s->d_tab);
    dfree(ctx, s->d_tile);
    dfree(ctx, s->d_part);
    dfree(ctx, s->d_pvol);
    dfree(ctx, s->d_pk);
    dfree(ctx, s->d_pout);
    dfree(ctx, s->d_rb);
    dfree(ctx, s->d_partb);
    dfree(ctx, s->d_tile_fin);
    dfree(ctx, s->d_xfull);
    if (s->h_st) (void)hipHostFree(s->h_st);
    if (s->up_done) (void)hipEventDestroy(s->up_done);
    if (s->up) (void)hipStreamDestroy(s->up);
    delete s;
}

}  // extern "C"
