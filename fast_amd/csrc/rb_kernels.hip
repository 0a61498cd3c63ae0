// gfx950 kernels of the mixed-precision CG on the red-black Schur complement (see RbDesc in
// tau_internal.h for the padded per-colour layout), and the streaming kernels the HBM
// ceiling is measured with.
#include "device_common.h"

#include <algorithm>

namespace ft {

// ---- red-black reduced system (mixed form, CG) --------------------------------------------
namespace {

// Element of slot k of row (y, z) in the padded per-colour layout (a ghost layer around the
// box; ghost slots and the slots past the end of odd rows hold 0 and are never written).
__device__ __forceinline__ unsigned rb_index(const RbDesc& d, int k, int y, int z) {
    return (unsigned)(k + 1) + (unsigned)d.px * (unsigned)(y + 1) + d.pp * (unsigned)(z + 1);
}

// Position of a thread of an RB tile: slot k of row (y, z) of one colour, z marching.
struct RbPos {
    int k, y, z0, z1;
    bool row_ok;
};

template <int TY>
__device__ __forceinline__ RbPos rb_pos(const RbDesc& d, int lt) {
    const int tix = lt % d.tiles_x;
    const int tiy = (lt / d.tiles_x) % d.tiles_y;
    const int tiz = lt / (d.tiles_x * d.tiles_y);
    RbPos p;
    p.k = tix * kRbTX + (int)(threadIdx.x % kRbTX);
    p.y = tiy * TY + (int)(threadIdx.x / kRbTX);
    p.z0 = tiz * d.tz;
    p.z1 = min(p.z0 + d.tz, d.nz);
    p.row_ok = (p.k < d.hx) && (p.y < d.ny);
    return p;
}

// Per-volume LDS tables shared by the RB passes.
struct RbTabs {
    double tab[2 * kTabEntries];
    FacePair<double> ftab[8];
};

// PART: 0 diagonals only, 1 inverses only, 2 both (a pass copies only what it reads; the
// other half of the LDS table stays unset and unread).
template <int PART, int NT>
__device__ __forceinline__ void rb_load_tables(const RbLaunch& L, int v, const Coef<double>& c, RbTabs& t) {
    const double* src = L.tab + (long long)v * 2 * kTabEntries;
    constexpr int lo = PART == 1 ? kTabEntries : 0, hi = PART == 0 ? kTabEntries : 2 * kTabEntries;
    for (int i = lo + (int)threadIdx.x; i < hi; i += NT) t.tab[i] = src[i];
    if (threadIdx.x < 8) t.ftab[threadIdx.x] = face_pair(c, threadIdx.x);
    __syncthreads();
}

// Loads of the six neighbours of the other colour, z-, y-, x-, x+, y+, z+, of the voxel
// x = 2k + s at element ci.  Rows of the other colour hold voxel x at slot k, x + 1 at slot
// k + s and x - 1 at slot k - 1 + s; a neighbour outside the box is a ghost (+0).
template <class U>
__device__ __forceinline__ void rb_nbrs(const U* u, const RbDesc& d, unsigned ci, int s, U (&o)[6]) {
    o[0] = ld(u, ci - d.pp);
    o[1] = ld(u, ci - (unsigned)d.px);
    o[2] = ld(u, ci - 1u + (unsigned)s);
    o[3] = ld(u, ci + (unsigned)s);
    o[4] = ld(u, ci + (unsigned)d.px);
    o[5] = ld(u, ci + d.pp);
}

// Sum over the faces of c_face * (neighbour value), in the order of the row product
// (Z-, Y-, X-, X+, Y+, Z+), each term fused into the sum (one rounding per face); ghost terms
// are fma(c, +0, acc) = acc and leave the sum unchanged.  The red-black form is held to
// accuracy bars, not to the reference's rounding, so it takes the fused form throughout.
template <class U>
__device__ __forceinline__ double rb_sum(const FacePair<double>* ftab, int code, const U (&o)[6]) {
    const int o4 = (code >> 4) & 4;
    const FacePair<double> px = ftab[(code & 3) | o4];
    const FacePair<double> py = ftab[((code >> 2) & 3) | o4];
    const FacePair<double> pz = ftab[((code >> 4) & 3) | o4];
    double acc = 0.0;
    acc = fma(pz.m, (double)o[0], acc);
    acc = fma(py.m, (double)o[1], acc);
    acc = fma(px.m, (double)o[2], acc);
    acc = fma(px.p, (double)o[3], acc);
    acc = fma(py.p, (double)o[4], acc);
    acc = fma(pz.p, (double)o[5], acc);
    return acc;
}

enum RbKind : int { RB_P = 0, RB_Y, RB_Q, RB_U, RB_UPD, RB_Y64, RB_R64, RB_SCATTER, RB_CODE };

// Loaded inputs of one voxel of an iteration pass.
struct RbIn {
    int code;
    float a, b, c, e;   // centre values
    float u[6];         // Y: red neighbours (p)
    double w[6];        // Q: black neighbours (Y)
};

// The iteration passes, one thread per slot of a row, marching tz planes:
//   P (red):   p = D^-1 r (first iteration) or D^-1 r + beta p_old; and, when x lacks the
//              last two iterations' steps (xpend == 2), x += a0 p_prev + a1 p_old, p_prev being
//              the older p that this pass overwrites
//   Y (black): Y = D_b^-1 (A_br p)
//   Q (red):   q = D_r p - A_rb Y = S p; partial p.q
//   U (red):   r -= alpha q; partials r.r and r.(D^-1 r).  alpha = (r.z) / (p.q) comes from a
//              finalize launch between Q and U, or (RbLaunch::fold_alpha, one-volume launch
//              sets) every U block reduces the Q pass's partials itself, which saves that launch
// x += alpha p is deferred from U: every second P pass adds the last two iterations' steps
// (it loads p_old anyway, and p_prev is still in the buffer it overwrites), so U reads r, q
// and the code (13 B per red voxel instead of 25) and P reads and writes x every other
// iteration (13 or 25 B, 19 on average).  x rounds once per two steps; the update pass adds
// whatever is pending before it folds x into x64.
// Loads are issued AH planes ahead of use into AH + 1 register sets used in rotation (the
// loop is unrolled by AH + 1), all unconditional: an invalid point reads the first interior
// slot, and ghosts make neighbour reads safe, so each plane waits only for its own loads.
template <int KIND, int TY, int AH>
__global__ __launch_bounds__(kRbTX* TY) void k_rb_it(RbLaunch L) {
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    if (slot >= L.ntiles) return;
    const int v = L.tile_vol[slot];
    const int tile = L.tile_id ? L.tile_id[slot] : slot;   // the tile's index in the batch
    const RbDesc d = L.vol[v];   // by value: kept in registers across the stores
    VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    const Coef<double> c = L.coef[v];
    rb_load_tables<KIND == RB_Q ? 0 : 1, kRbTX * TY>(L, v, c, tabs);   // Q reads D; P, Y and U read D^-1
    const RbPos ps = rb_pos<TY>(d, tile - d.tile_base);
    constexpr int kColour = (KIND == RB_Y) ? 1 : 0;
    double alpha = vs->alpha;   // U: replaced by this iteration's alpha below
    const double beta = vs->beta;
    const bool first = vs->first != 0;
    const bool xpend = KIND == RB_P && vs->xpend == 2;   // uniform over the volume
    const double xa0 = vs->xalpha[0], xa1 = vs->xalpha[1];
    const bool odd = (vs->phase & 1) != 0;
    const float* p_old = L.p[odd ? 1 : 0] + d.off_c;
    float* p_new = L.p[odd ? 0 : 1] + d.off_c;
    float* X = L.x + d.off_c;
    float* Rr = L.r + d.off_c;
    float* Q = L.q + d.off_c;
    double* Y = L.y + d.off_c;
    const uint8_t* code_c = d.code_c[kColour];
    const unsigned c00 = rb_index(d, 0, 0, 0);
    auto valid = [&](int z, int& s, int& x) {
        s = (kColour + ps.y + z) & 1;
        x = 2 * ps.k + s;
        return ps.row_ok && z < ps.z1 && x < d.nx;
    };
    auto load = [&](int z, RbIn& in) {
        int s, x;
        const unsigned ci = valid(z, s, x) ? rb_index(d, ps.k, ps.y, z) : c00;
        in.code = ld(code_c, ci);
        if (KIND == RB_P) {
            in.a = ld(Rr, ci);
            in.b = ld(p_old, ci);   // unused on the first iteration
            if (xpend) {
                in.c = ld(X, ci);
                in.e = ld(p_new, ci);   // p two iterations back, about to be overwritten
            }
        } else if (KIND == RB_Y) {
            rb_nbrs(p_new, d, ci, s, in.u);
        } else if (KIND == RB_Q) {
            in.a = ld(p_new, ci);
            rb_nbrs(Y, d, ci, s, in.w);
        } else {
            in.c = ld(Rr, ci);
            in.e = ld(Q, ci);
        }
    };
    AccDD acc[2];
    acc[0].init();
    acc[1].init();
    auto proc = [&](int z, const RbIn& in) {
        int s, x;
        if (!valid(z, s, x)) return;
        const unsigned ci = rb_index(d, ps.k, ps.y, z);
        const int code = in.code;
        const int bx = box_flags(x, ps.y, z, d.nx, d.ny, d.nz);
        const DiagInv<double> di = diag_at<double>(tabs.tab, code, bx, d.primary, c);
        if (KIND == RB_P) {
            const double z_ = di.iv * (double)in.a;
            const double pn = first ? z_ : z_ + beta * (double)in.b;
            st(p_new, ci, (float)pn);
            if (xpend) st(X, ci, (float)(((double)in.c + xa0 * (double)in.e) + xa1 * (double)in.b));
        } else if (KIND == RB_Y) {
            st(Y, ci, rb_sum(tabs.ftab, code, in.u) * di.iv);
        } else if (KIND == RB_Q) {
            const double pv = (double)in.a;
            const double qv = fma(di.d, pv, -rb_sum(tabs.ftab, code, in.w));
            st(Q, ci, (float)qv);
            acc[0].add(pv, qv);
        } else {
            const float rs = (float)((double)in.c - alpha * (double)in.e);
            st(Rr, ci, rs);
            acc[0].add(rs, rs);
            acc[1].add((double)rs, di.iv * (double)rs);
        }
    };
    // loads AH planes ahead into a ring of AH + 1 register sets; no early exit: a step past
    // the last plane loads the first slot and does nothing, which keeps the loop body straight
    // (the compiler then waits for each set's own loads only)
    constexpr int R = AH + 1;
    RbIn in[R];
#pragma unroll
    for (int i = 0; i < AH; ++i) load(ps.z0 + i, in[i]);
    if (KIND == RB_U && L.fold_alpha) {
        // alpha = (r.z) / (p.q) from the Q pass's tile partials of the volume, reduced by every
        // block alike while its first planes load (a double-double sum rounds to the same value
        // for any reduction tree), as the finalize would; the block of the volume's first tile
        // stores it (the finalize after U records it with x's pending step), and a non-finite
        // alpha starts the volume afresh
        __shared__ double pq;
        double tot[1];
        reduce_volume<1, kRbTX * TY>(L.part, d.tile_base, d.ntiles, tot);
        if (threadIdx.x == 0) pq = tot[0];
        __syncthreads();
        alpha = vs->absnew / pq;
        if (tile == d.tile_base && threadIdx.x == 0) {
            vs->alpha = alpha;
            if (vs->xpend == 2) vs->xpend = 0;   // the P pass of this iteration added both steps
            if (!isfinite(alpha)) vs->status = ST_FRESH;
        }
        if (!isfinite(alpha)) return;
    }
    for (int z = ps.z0; z < ps.z1; z += R) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            load(z + j + AH, in[(j + AH) % R]);
            proc(z + j, in[j]);
        }
    }
    if (KIND == RB_Q) block_reduce_store<AccDD, 1, kRbTX * TY>(reinterpret_cast<AccDD(&)[1]>(acc), L.part + (long long)tile * kPart);
    // U's partials go after the Q pass's p.q in the tile's slots: blocks still reducing p.q
    // read the first two
    if (KIND == RB_U) block_reduce_store<AccDD, 2, kRbTX * TY>(acc, L.part + (long long)tile * kPart + 2);
}

// Q with the black Y planes staged in LDS: a block owns kRbTX slots x TY rows of red
// voxels and marches z, keeping black planes z - 1, z, z + 1 over its tile plus a one-slot,
// one-row halo in a four-plane ring.  Per plane a thread loads at most two Y values (the halo
// plane in block-wide order) instead of six neighbours.  With four ring planes one barrier per
// plane suffices: the plane written at step z + 1 is not among those read at step z, and the
// one written at step z + 2 can only be written once every thread has passed step z + 1's
// barrier.  Same values, roundings and summation order as k_rb_it<RB_Q>.  (Staging p for the
// Y pass the same way measured slower than k_rb_it<RB_Y>: its loads are 4-byte.)
template <int TY>
__global__ __launch_bounds__(kRbTX* TY) void k_rb_q(RbLaunch L) {
    using U = double;
    constexpr int W = kRbTX + 2, PL = W * (TY + 2);
    constexpr int NT = kRbTX * TY;
    constexpr int RND = (PL + NT - 1) / NT;
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    if (slot >= L.ntiles) return;
    const int v = L.tile_vol[slot];
    const int tile = L.tile_id ? L.tile_id[slot] : slot;   // the tile's index in the batch
    const RbDesc d = L.vol[v];
    const VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    __shared__ U ring[4 * PL];
    const Coef<double> c = L.coef[v];
    rb_load_tables<0, NT>(L, v, c, tabs);   // Q reads D only
    const int lt = tile - d.tile_base;
    const int k0 = (lt % d.tiles_x) * kRbTX;
    const int y0 = ((lt / d.tiles_x) % d.tiles_y) * TY;
    const int z0 = (lt / (d.tiles_x * d.tiles_y)) * d.tz;
    const int z1 = min(z0 + d.tz, d.nz);
    const bool odd = (vs->phase & 1) != 0;
    const float* P = L.p[odd ? 0 : 1] + d.off_c;
    const U* src = L.y + d.off_c;
    const uint8_t* code_c = d.code_c[0];
    // element of halo point j of this thread in black plane zc (-1 <= zc <= nz: the ghost layer
    // covers the box's neighbours); points that no red voxel reads load element 0
    unsigned hofs[RND];
#pragma unroll
    for (int j = 0; j < RND; ++j) {
        const int i = (int)threadIdx.x + j * NT;
        const int kk = k0 - 1 + i % W, yy = y0 - 1 + i / W;
        hofs[j] = (i < PL && kk <= d.hx && yy <= d.ny) ? (unsigned)(kk + 1) + (unsigned)d.px * (unsigned)(yy + 1) : ~0u;
    }
    auto h_load = [&](int zc, U (&h)[RND]) {
#pragma unroll
        for (int j = 0; j < RND; ++j) {
            const bool ok = hofs[j] != ~0u && zc <= d.nz;
            h[j] = ld(src, ok ? hofs[j] + d.pp * (unsigned)(zc + 1) : 0u);
        }
    };
    auto h_store = [&](int zc, const U (&h)[RND]) {
        U* dst = ring + (zc & 3) * PL;
#pragma unroll
        for (int j = 0; j < RND; ++j) {
            const int i = (int)threadIdx.x + j * NT;
            if (i < PL) dst[i] = h[j];
        }
    };
    const int k = k0 + (int)(threadIdx.x % kRbTX), y = y0 + (int)(threadIdx.x / kRbTX);
    const bool row_ok = k < d.hx && y < d.ny;
    const int col = (k - k0 + 1) + (y - y0 + 1) * W;   // this thread's slot in a ring plane
    const unsigned c00 = rb_index(d, 0, 0, 0);
    auto c_load = [&](int z, int& code, float& pv) {
        const int x = 2 * k + ((y + z) & 1);
        const unsigned ci = (row_ok && z < z1 && x < d.nx) ? rb_index(d, k, y, z) : c00;
        code = ld(code_c, ci);
        pv = ld(P, ci);
    };
    AccDD acc[1];
    acc[0].init();
    auto plane = [&](int z, int code, float pv) {
        const int s = (y + z) & 1;
        const int x = 2 * k + s;
        if (!(row_ok && z < z1 && x < d.nx)) return;
        const int bx = box_flags(x, y, z, d.nx, d.ny, d.nz);
        const DiagInv<double> di = diag_at<double>(tabs.tab, code, bx, d.primary, c);
        const U* pl = ring + (z & 3) * PL + col;
        U o[6];
        o[0] = ring[((z - 1) & 3) * PL + col];
        o[1] = pl[-W];
        o[2] = pl[s - 1];
        o[3] = pl[s];
        o[4] = pl[W];
        o[5] = ring[((z + 1) & 3) * PL + col];
        const double p64 = (double)pv;
        const double qv = fma(di.d, p64, -rb_sum(tabs.ftab, code, o));
        st(L.q + d.off_c, rb_index(d, k, y, z), (float)qv);
        acc[0].add(p64, qv);
    };
    U ha[RND], hb[RND];
    h_load(z0 - 1, ha);
    h_store(z0 - 1, ha);
    h_load(z0, ha);
    h_store(z0, ha);
    h_load(z0 + 1, ha);
    int code_a, code_b;
    float pv_a, pv_b;
    c_load(z0, code_a, pv_a);
    // unrolled by 2 over alternating register sets, no early exit (a step past the last plane
    // stores an unused ring plane and computes nothing)
    for (int z = z0; z < z1; z += 2) {
        h_store(z + 1, ha);
        h_load(z + 2, hb);
        c_load(z + 1, code_b, pv_b);
        __syncthreads();
        plane(z, code_a, pv_a);
        h_store(z + 2, hb);
        h_load(z + 3, ha);
        c_load(z + 2, code_a, pv_a);
        __syncthreads();
        plane(z + 1, code_b, pv_b);
    }
    block_reduce_store<AccDD, 1, NT>(acc, L.part + (long long)tile * kPart);
}

// Compile-time integer, used to unroll the fused pass's march with constant register indices.
template <int N>
struct IC {
    static constexpr int value = N;
};

// Steps J .. N-1 of an unrolled stretch of a march over a register ring of length R:
// f(IC<j mod R>, t + j) for each j, leaving the stretch (true) after the step of plane `last`.
template <int J, int N, int R, class F>
__device__ __forceinline__ bool ring_turn(F& f, int t, int last) {
    if constexpr (J < N) {
        f(IC<J % R>{}, t + J);
        if (t + J >= last) return true;
        return ring_turn<J + 1, N, R>(f, t, last);
    } else {
        return false;
    }
}

// Y and Q fused: q = S p = D_r p - A_rb (D_b^-1 A_br p) for a block's red tile without
// storing Y.  A block owns kRbTX slots x TY rows of red voxels and marches z; its threads are
// columns (slot k, row y) that march with it:
//   waves 0 .. TY-1:  the block's own rows y0 .. y0+TY-1 (lane = slot k0 + lane);
//   waves TY, TY+1:   the halo rows y0-1 and y0+TY;
//   wave TY+2:        lanes 0 .. 2TY-1 are the side columns k0-1 and k0+64 of the own rows.
// Each column keeps its red p of planes t-1 .. t+AHEAD in registers (loads issued AHEAD planes
// ahead, the march unrolled by the ring length so no register is copied while its load is in
// flight) and computes the black Y of its column; the own rows also form q.  The x and y
// neighbours come from LDS: a p plane over slots k0-2 .. k0+65 and rows y0-2 .. y0+TY+1
// (the points no column owns are loaded by the lanes of waves TY .. TY+2, one each),
// double-buffered, and a Y plane over slots k0-1 .. k0+64 and rows y0-1 .. y0+TY in a ring of
// R buffers indexed like the registers.  Step t writes p(t), passes one barrier, computes Y(t)
// and q(t-1), whose z neighbours Y(t-2) and Y(t) are the column's own (Y(t-2) read back from
// the ring rather than kept in registers); a buffer is rewritten two steps or more after it
// was last read, with a barrier in between.  Y and q take the same values, roundings and
// summation order as k_rb_it<RB_Y> followed by k_rb_it<RB_Q>, so results are bit-identical.
// Per red voxel the pass reads p and the two codes and writes q (10 bytes, against 30 for the
// separate passes, whose Y is FP64); the halo's Y is computed by both blocks that read it.
// With a prefetch of 2 planes the pass fits 80 VGPRs, 6 waves per SIMD: two blocks of up to
// 12 waves (9 rows) per CU.  At 84 VGPRs (5 waves per SIMD) a CU held one block of 8 to 13 rows.
template <int TY, int AHEAD>
__global__ __launch_bounds__(kRbTX*(TY + 3)) __attribute__((amdgpu_waves_per_eu(AHEAD == 2 ? 6 : 1)))
void k_rb_yq(RbLaunch L) {
    static_assert(AHEAD >= 2, "p of plane t+1 is needed at step t");
    constexpr int R = AHEAD + 2;   // ring length: planes t-1 .. t+AHEAD
    constexpr int NT = kRbTX * (TY + 3);
    constexpr int PW = kRbTX + 4, PH = TY + 4;   // p plane: slots k0-2 .. k0+65, rows y0-2 .. y0+TY+1
    constexpr int YW = kRbTX + 2, YH = TY + 2;   // Y plane: slots k0-1 .. k0+64, rows y0-1 .. y0+TY
    constexpr int NPO = 2 * PW + 2 * (TY + 2) + 4;   // p points that no column owns
    static_assert(kRbTX == 64 && NPO <= 3 * 64, "one p-only point per lane of waves TY .. TY+2");
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    if (slot >= L.ntiles) return;
    const int v = L.tile_vol[slot];
    const int tile = L.tile_id ? L.tile_id[slot] : slot;   // the tile's index in the batch
    const RbDesc d = L.vol[v];
    const VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    __shared__ float pbuf[2][PW * PH + 1];    // + a spare slot for lanes with nothing to store
    __shared__ double ybuf[R][YW * YH + 1];
    const Coef<double> c = L.coef[v];
    rb_load_tables<2, NT>(L, v, c, tabs);   // Y reads D_b^-1, Q reads D_r
    const int lt = tile - d.tile_base;
    const int k0 = (lt % d.tiles_x) * kRbTX;
    const int y0 = ((lt / d.tiles_x) % d.tiles_y) * TY;
    const int z0 = (lt / (d.tiles_x * d.tiles_y)) * d.tz;
    const int z1 = min(z0 + d.tz, d.nz);
    const bool odd = (vs->phase & 1) != 0;
    const float* P = L.p[odd ? 0 : 1] + d.off_c;
    float* Q = L.q + d.off_c;
    const uint8_t* code_r = d.code_c[0];
    const uint8_t* code_b = d.code_c[1];
    const int tid = (int)threadIdx.x, w = tid >> 6, lane = tid & 63;
    // this thread's column (slot kc, row yc)
    const bool own = w < TY;
    const bool has_col = w < TY + 2 || lane < 2 * TY;
    const int kc = w <= TY + 1 ? k0 + lane : (lane < TY ? k0 - 1 : k0 + kRbTX);
    const int yc = own ? y0 + w : (w == TY ? y0 - 1 : (w == TY + 1 ? y0 + TY : y0 + (lane % TY)));
    const bool col_pad = has_col && kc >= -1 && kc <= d.hx && yc >= -1 && yc <= d.ny;   // inside the padded layout
    const bool col_box = has_col && kc >= 0 && kc < d.hx && yc >= 0 && yc < d.ny;
    // Element of the column in plane -1.  Outside the padded layout it is element 0, and a
    // plane offset outside -1 .. nz is 0: either way the load reads a ghost (+0, code 0).
    const unsigned cel = col_pad ? (unsigned)(kc + 1) + (unsigned)d.px * (unsigned)(yc + 1) : 0u;
    // The codes are read as the 4-byte word holding the column's byte: plane pitches are
    // multiples of 4 (rb_row_pitch), so the byte sits at the same shift in every plane, and a
    // 32-bit register ring needs no re-extension of a byte between the steps of the march.
    const unsigned cel4 = cel >> 2, csh = (cel & 3u) * 8u;
    const unsigned* code_r4 = reinterpret_cast<const unsigned*>(code_r);
    const unsigned* code_b4 = reinterpret_cast<const unsigned*>(code_b);
    const int pdst = has_col ? (kc - k0 + 2) + (yc - y0 + 2) * PW : PW * PH;
    const int ypos = (kc - k0 + 1) + (yc - y0 + 1) * YW;
    const int ydst = has_col ? ypos : YW * YH;
    // per colour parity s of the voxel x = 2 kc + s of the column: inside the box, and the
    // x-face bits of the diagonal-table index; rows on a y face take the formula
    const bool ok0 = col_box && 2 * kc < d.nx, ok1 = col_box && 2 * kc + 1 < d.nx;
    const int xo0 = (2 * kc == d.nx - 1 ? 1 << 7 : 0) | (kc == 0 ? 1 << 8 : 0);
    const int xo1 = 2 * kc + 1 == d.nx - 1 ? 1 << 7 : 0;
    const bool yface = yc == 0 || yc == d.ny - 1;
    const int sb0 = (1 + yc) & 1;   // parity of the column's black voxel in even planes (red: 1 - sb0)
    // this thread's p-only point (lanes of waves TY .. TY+2)
    int podst = PW * PH;
    unsigned pel = 0u;
    {
        const int i = (w - TY) * 64 + lane;
        if (w >= TY && i < NPO) {
            int kk, yy;
            if (i < 2 * PW) {
                kk = k0 - 2 + i % PW;
                yy = i < PW ? y0 - 2 : y0 + TY + 1;
            } else if (i < 2 * PW + 2 * (TY + 2)) {
                const int j = i - 2 * PW;
                kk = j < TY + 2 ? k0 - 2 : k0 + kRbTX + 1;
                yy = y0 - 1 + j % (TY + 2);
            } else {
                const int j = i - 2 * PW - 2 * (TY + 2);
                kk = (j & 1) ? k0 + kRbTX : k0 - 1;
                yy = (j & 2) ? y0 + TY : y0 - 1;
            }
            podst = (kk - k0 + 2) + (yy - y0 + 2) * PW;
            if (kk >= -1 && kk <= d.hx && yy >= -1 && yy <= d.ny) pel = (unsigned)(kk + 1) + (unsigned)d.px * (unsigned)(yy + 1);
        }
    }
    // plane zz's offset in elements (uniform); 0 outside -1 .. nz
    auto plane = [&](int zz) { return (unsigned)(zz + 1) <= (unsigned)(d.nz + 1) ? d.pp * (unsigned)(zz + 1) : 0u; };
    // diagonal entry (and inverse) of a row: tables, or the formula on a y or z box face
    auto diag = [&](int code, int s, int yy, int zz, bool zface) {
        if (yface || zface) return diag_at<double>(tabs.tab, code, box_flags(2 * kc + s, yy, zz, d.nx, d.ny, d.nz), d.primary, c);
        const int i = code | (s ? xo1 : xo0);
        return DiagInv<double>{tabs.tab[i], tabs.tab[kTabEntries + i]};
    };
    AccDD acc[1];
    acc[0].init();
    // register rings indexed by (plane - t0) mod R, t0 = z0 - 1: p of the column (planes t-1 ..
    // t+AHEAD), the p-only point and the black code (t .. t+AHEAD), the red code (t-1 ..
    // t+AHEAD-1)
    const int t0 = z0 - 1;
    float pr[R], po[R];
    unsigned cb[R], cr[R];   // code words (the column's byte at shift csh)
    pr[R - 1] = ld(P + plane(t0 - 1), cel);
    cr[R - 1] = 0;   // q(t0 - 1) is never formed
#pragma unroll
    for (int i = 0; i < AHEAD; ++i) {
        const unsigned o = plane(t0 + i);
        pr[i] = ld(P + o, cel);
        po[i] = ld(P + o, pel);
        cb[i] = ld(code_b4 + (o >> 2), cel4);
        if (i < AHEAD - 1) cr[i] = ld(code_r4 + (o >> 2), cel4);
    }
    auto step = [&](auto J, int t) {
        constexpr int j = decltype(J)::value;
        constexpr int jm = (j + R - 1) % R, jn = (j + 1) % R;
        {
            const unsigned o = plane(t + AHEAD), oq = plane(t + AHEAD - 1);
            pr[(j + AHEAD) % R] = ld(P + o, cel);
            po[(j + AHEAD) % R] = ld(P + o, pel);
            cb[(j + AHEAD) % R] = ld(code_b4 + (o >> 2), cel4);
            cr[(j + AHEAD - 1) % R] = ld(code_r4 + (oq >> 2), cel4);
        }
        float* pb = pbuf[t & 1];
        pb[pdst] = pr[j];
        pb[podst] = po[j];
        __syncthreads();
        // Y(t) of the column's black voxel (every column; ghosts +0)
        {
            const int s = sb0 ^ (t & 1);
            const float* pl = pb + pdst;
            const float u[6] = {pr[jm], pl[-PW], pl[s - 1], pl[s], pl[PW], pr[jn]};
            const int code = (int)((cb[j] >> csh) & 0xffu);
            const DiagInv<double> di = diag(code, s, yc, t, t == 0 || t == d.nz - 1);
            const double yv = rb_sum(tabs.ftab, code, u) * di.iv;
            const bool ok = (s ? ok1 : ok0) && t >= 0 && t < d.nz;
            const double yt = ok ? yv : 0.0;
            ybuf[j][ydst] = yt;
            // q(t - 1) of the column's red voxel (own rows)
            const int tq = t - 1;
            const int sq = s;   // (yc + t - 1) & 1: the red voxel of plane t-1 has the black's parity of plane t
            if (own && (sq ? ok1 : ok0) && tq >= z0 && tq < z1) {
                const double* yl = ybuf[jm] + ypos;
                const double o[6] = {ybuf[(j + R - 2) % R][ypos], yl[-YW], yl[sq - 1], yl[sq], yl[YW], yt};
                const int code_q = (int)((cr[jm] >> csh) & 0xffu);
                const DiagInv<double> dq = diag(code_q, sq, yc, tq, tq == 0 || tq == d.nz - 1);
                const double p64 = (double)pr[jm];
                const double qv = fma(dq.d, p64, -rb_sum(tabs.ftab, code_q, o));
                st(Q + d.pp * (unsigned)(tq + 1), cel, (float)qv);
                acc[0].add(p64, qv);
            }
        }
    };
    // z0 - 1 .. z1: Y of the planes around the tile, q of its planes (the exits are uniform
    // over the block).  Unrolling over several ring turns (fewer back edges, where the
    // compiler waits for every load in flight) measured no faster.
    for (int t = t0;; t += R)
        if (ring_turn<0, R, R>(step, t, z1)) break;
    block_reduce_store<AccDD, 1, NT>(acc, L.part + (long long)tile * kPart);
}

// The passes outside the Krylov loop, one thread per slot of a row, marching tz planes:
//   CODE:    per-colour stencil codes from the full-grid codes (set-up)
//   UPD:     x64 += x (kept unless discarded; x's pending steps alpha p added first), x = 0
//                                                                          [update / fresh start]
//   Y64:     black x64b = D_b^-1 (b_b - A_br x64), partial |b_b|^2       [update / fresh start]
//   R64:     red r = b_S - S x64 = (b_r - D_r x64) - A_rb x64b in FP64, stored in FP32;
//            partials |r64|^2, |b_r|^2, |r32|^2, r32.(D^-1 r32)           [update / fresh start]
//   SCATTER: both colours of x64 / x64b into the full-grid solution        [end]
// On the first update x64 is the reference's initial guess x0, which depends only on the
// primary coordinate: Y64 evaluates it for the red neighbours, R64 stores it.
template <int KIND, int TY>
__global__ __launch_bounds__(kRbTX* TY) void k_rb(RbLaunch L) {
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    if (slot >= L.ntiles) return;
    const int v = L.tile_vol[slot];
    const int tile = L.tile_id ? L.tile_id[slot] : slot;   // the tile's index in the batch
    const RbDesc d = L.vol[v];
    VolState<double>* vs = L.st + v;
    if (KIND == RB_UPD || KIND == RB_Y64 || KIND == RB_R64) {
        const int status = vs->status;
        if (status != ST_UPDATE && status != ST_FRESH) return;
    }
    constexpr bool kTables = (KIND == RB_Y64 || KIND == RB_R64);
    __shared__ RbTabs tabs;
    const Coef<double> c = L.coef[v];
    if (kTables) rb_load_tables<2, kRbTX * TY>(L, v, c, tabs);
    const RbPos ps = rb_pos<TY>(d, tile - d.tile_base);
    constexpr int kColour = (KIND == RB_Y64) ? 1 : 0;
    const bool first_res = vs->refinements == 0;   // x64 not yet set: use x0
    const bool keep = vs->refinements > 0 && !vs->discard;
    // x lacks the steps of the last xpend iterations: the newest in the buffer the next P pass
    // reads as p_old, an older one in the other buffer; added in the P pass's order
    const int xpend = KIND == RB_UPD ? vs->xpend : 0;
    const double xa0 = vs->xalpha[0], xa1 = vs->xalpha[1];
    const float* p_last = L.p[(vs->phase & 1) ? 1 : 0] + d.off_c;
    const float* p_prev = L.p[(vs->phase & 1) ? 0 : 1] + d.off_c;
    const int nprim = d.primary == 0 ? d.nx : (d.primary == 1 ? d.ny : d.nz);
    float* X = L.x + d.off_c;
    float* Rr = L.r + d.off_c;
    double* X64 = L.x64 + d.off_c;
    double* X64B = L.x64b + d.off_c;
    AccDD acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i].init();
    if (ps.row_ok) {
        for (int z = ps.z0; z < ps.z1; ++z) {
            const int s = (kColour + ps.y + z) & 1;
            const int x = 2 * ps.k + s;
            const unsigned ci = rb_index(d, ps.k, ps.y, z);
            if (KIND == RB_SCATTER) {   // slot k of both colours: red x = 2k + s, black x = 2k + 1 - s
                double* row = L.xfull + d.off + (long long)d.nx * ((long long)ps.y + (long long)d.ny * z);
                if (x < d.nx) row[x] = X64[ci];
                if (2 * ps.k + 1 - s < d.nx) row[2 * ps.k + 1 - s] = X64B[ci];
                continue;
            }
            if (KIND == RB_CODE) {   // both colours
                const unsigned lin = (unsigned)d.nx * ((unsigned)ps.y + (unsigned)d.ny * (unsigned)z);
                uint8_t* cr = const_cast<uint8_t*>(d.code_c[0]);
                uint8_t* cb = const_cast<uint8_t*>(d.code_c[1]);
                if (x < d.nx) cr[ci] = d.code[lin + (unsigned)x];
                if (2 * ps.k + 1 - s < d.nx) cb[ci] = d.code[lin + (unsigned)(2 * ps.k + 1 - s)];
                continue;
            }
            if (x >= d.nx) continue;
            if (KIND == RB_UPD) {
                if (keep) {
                    const float xv = xpend == 2   ? (float)(((double)X[ci] + xa0 * (double)p_prev[ci]) + xa1 * (double)p_last[ci])
                                     : xpend == 1 ? (float)((double)X[ci] + xa0 * (double)p_last[ci])
                                                  : X[ci];
                    X64[ci] = X64[ci] + (double)xv;
                }
                X[ci] = 0.0f;
                continue;
            }
            const int code = ld(d.code_c[kColour], ci);
            const int bx = box_flags(x, ps.y, z, d.nx, d.ny, d.nz);
            const DiagInv<double> di = diag_at<double>(tabs.tab, code, bx, d.primary, c);
            const int kp = d.primary == 0 ? x : (d.primary == 1 ? ps.y : z);
            if (KIND == RB_Y64) {
                double g;
                if (first_res) {
                    const int o4 = (code >> 4) & 4;
                    const FacePair<double> px = tabs.ftab[(code & 3) | o4];
                    const FacePair<double> py = tabs.ftab[((code >> 2) & 3) | o4];
                    const FacePair<double> pz = tabs.ftab[((code >> 4) & 3) | o4];
                    auto x0n = [&](int dxp, int dyp, int dzp, bool ok) {
                        if (!ok) return 0.0;
                        const int kk = d.primary == 0 ? x + dxp : (d.primary == 1 ? ps.y + dyp : z + dzp);
                        return x0_of<double>(kk, nprim, d.sgn);
                    };
                    g = 0.0;
                    g += pz.m * x0n(0, 0, -1, z > 0);
                    g += py.m * x0n(0, -1, 0, ps.y > 0);
                    g += px.m * x0n(-1, 0, 0, x > 0);
                    g += px.p * x0n(1, 0, 0, x + 1 < d.nx);
                    g += py.p * x0n(0, 1, 0, ps.y + 1 < d.ny);
                    g += pz.p * x0n(0, 0, 1, z + 1 < d.nz);
                } else {
                    double o[6];
                    rb_nbrs(X64, d, ci, s, o);
                    g = rb_sum(tabs.ftab, code, o);
                }
                const double b = rhs_of<double>(code, kp, nprim, c);
                st(X64B, ci, (b - g) / di.d);
                acc[0].add(b, b);
            } else if (KIND == RB_R64) {
                const double xr = first_res ? x0_of<double>(kp, nprim, d.sgn) : ld(X64, ci);
                if (first_res) st(X64, ci, xr);
                const double b = rhs_of<double>(code, kp, nprim, c);
                double o[6];
                rb_nbrs(X64B, d, ci, s, o);
                const double r64 = (b - di.d * xr) - rb_sum(tabs.ftab, code, o);
                const float r32 = (float)r64;
                st(Rr, ci, r32);
                acc[0].add(r64, r64);
                acc[1].add(b, b);
                acc[2].add(r32, r32);
                acc[3].add((double)r32, di.iv * (double)r32);
            }
        }
    }
    if (KIND == RB_Y64) block_reduce_store<AccDD, 1, kRbTX * TY>(reinterpret_cast<AccDD(&)[1]>(acc), L.part_b + (long long)tile * kPart);
    if (KIND == RB_R64) block_reduce_store<AccDD, 4, kRbTX * TY>(acc, L.part + (long long)tile * kPart);
}

// Per-volume finalize of an RB pass (one block per volume, fixed-order reduction).
template <int FIN>
__global__ __launch_bounds__(kFinThreads) void k_rb_fin(RbLaunch L) {
    const int v = blockIdx.x;
    if (v >= L.nvol) return;
    VolState<double>* s = L.st + v;
    const RbDesc& d = L.vol[v];
    if (FIN == F_MIXRES) {
        if (s->status != ST_UPDATE && s->status != ST_FRESH) return;
        double tot[4], totb[1];
        reduce_volume<4, kFinThreads>(L.part, d.tile_base, d.ntiles, tot);
        reduce_volume<1, kFinThreads>(L.part_b, d.tile_base, d.ntiles, totb);
        if (threadIdx.x == 0) {
            tot[1] += totb[0];   // |b|^2 over both colours
            s->xpend = 0;        // the update pass added it
            mixres_update(s, tot);
        }
        return;
    }
    if (s->status != ST_RUNNING) return;
    constexpr int NP = (FIN == F_CG2) ? 2 : 1;
    double tot[NP];
    // the U pass's partials sit after the Q pass's p.q in each tile's slots
    reduce_volume<NP, kFinThreads>(L.part + (FIN == F_CG2 ? 2 : 0), d.tile_base, d.ntiles, tot);
    if (threadIdx.x == 0) {
        double ta[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NP; ++k) ta[k] = tot[k];
        if (FIN == F_CG1 && s->xpend == 2) s->xpend = 0;   // the P pass added both steps
        if (FIN == F_CG2) s->xalpha[s->xpend++] = s->alpha;   // x lacks this iteration's step
        fin_update<double, FIN>(s, ta);
    }
}

__global__ __launch_bounds__(256) void k_rb_tables(const RbDesc* vol, const Coef<double>* coef, double* tab, int nvol) {
    const int v = blockIdx.x;
    if (v >= nvol) return;
    const Coef<double> c = coef[v];
    const int primary = vol[v].primary;
    double* t = tab + (long long)v * 2 * kTabEntries;
    for (int i = threadIdx.x; i < kTabEntries; i += 256) {
        const double dg = diag_of<double>(i & 127, i >> 7, primary, c);
        t[i] = dg;
        t[kTabEntries + i] = 1.0 / dg;
    }
}

// Rows per tile (RbLaunch::ty) as a template argument: every kernel of a batch uses the same
// tiles, so the partial sums of a pass line up with the tile ranges the finalizes reduce.
#define FT_RB_TY(F)   \
    switch (L.ty) {    \
        case 2: F(2);  \
        case 8: F(8);  \
        case 9: F(9);  \
        case 10: F(10); \
        case 12: F(12); \
        case 13: F(13); \
        default: F(4); \
    }

template <int KIND>
void rb_launch(const RbLaunch& L, hipStream_t s) {
#define FT_GO(TYV)                                                                              \
    hipLaunchKernelGGL((k_rb<KIND, TYV>), dim3(L.ntiles), dim3(kRbTX * TYV), 0, s, L); \
    return;
    FT_RB_TY(FT_GO)
#undef FT_GO
}

// Planes the streaming passes P and U load ahead.  3, 4 and 6 measured no faster than 2 on
// MI355X (256^3 and 512^3), so every plane-marching pass loads two ahead.
constexpr int kRbStreamAhead = 2;

template <int KIND>
void rb_launch_it(const RbLaunch& L, hipStream_t s) {
    constexpr int AH = (KIND == RB_P || KIND == RB_U) ? kRbStreamAhead : 2;
#define FT_GO(TYV)                                                                                     \
    hipLaunchKernelGGL((k_rb_it<KIND, TYV, AH>), dim3(L.ntiles), dim3(kRbTX * TYV), 0, s, L); \
    return;
    FT_RB_TY(FT_GO)
#undef FT_GO
}

}  // namespace

namespace {
// Streaming kernels of known traffic, the measured ceiling the stencil passes are compared
// with: c = a + s b (two reads, one write), b = a (one read, one write), a = s (one write);
// 16-byte accesses, one per thread per sweep, the grid striding when it is capped.
__global__ __launch_bounds__(256) void k_stream_triad(const float4* __restrict__ a, const float4* __restrict__ b,
                                                      float4* __restrict__ c, size_t n4, float sc) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        const float4 x = a[i], y = b[i];
        c[i] = make_float4(x.x + sc * y.x, x.y + sc * y.y, x.z + sc * y.z, x.w + sc * y.w);
    }
}
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) b[i] = a[i];
}
__global__ __launch_bounds__(256) void k_stream_fill(float4* __restrict__ a, size_t n4, float sc) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
        a[i] = make_float4(sc, sc, sc, sc);
}
}  // namespace

// kind 0 triad, 1 copy, 2 fill over buffers of n4 float4; blocks = 0: one float4 per thread.
hipError_t stream_kernel(int kind, float4* a, float4* b, float4* c, size_t n4, unsigned blocks, hipStream_t s) {
    const size_t need = (n4 + 255) / 256;
    const unsigned g = blocks ? blocks : (unsigned)std::min<size_t>(need, 0x7FFFFFFFu);
    if (kind == 0) hipLaunchKernelGGL(k_stream_triad, dim3(g), dim3(256), 0, s, a, b, c, n4, 0.5f);
    else if (kind == 1) hipLaunchKernelGGL(k_stream_copy, dim3(g), dim3(256), 0, s, a, b, n4);
    else hipLaunchKernelGGL(k_stream_fill, dim3(g), dim3(256), 0, s, a, n4, 0.5f);
    return hipGetLastError();
}

hipError_t rb_setup(const RbLaunch& L, hipStream_t s) {
    if (L.nvol <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rb_tables, dim3(L.nvol), dim3(256), 0, s, L.vol, L.coef, L.tab, L.nvol);
    rb_launch<RB_CODE>(L, s);
    return hipGetLastError();
}

static void rb_launch_q(const RbLaunch& L, hipStream_t s) {
    if (L.staged == 2) {
#define FT_GO(TYV)                                                                                                  \
    if (L.ahead == 2) hipLaunchKernelGGL((k_rb_yq<TYV, 2>), dim3(L.ntiles), dim3(kRbTX * (TYV + 3)), 0, s, L);     \
    else hipLaunchKernelGGL((k_rb_yq<TYV, 3>), dim3(L.ntiles), dim3(kRbTX * (TYV + 3)), 0, s, L);                  \
    return;
        FT_RB_TY(FT_GO)
#undef FT_GO
    }
    if (L.staged) {
#define FT_GO(TYV)                                                                          \
    hipLaunchKernelGGL((k_rb_q<TYV>), dim3(L.ntiles), dim3(kRbTX * TYV), 0, s, L); \
    return;
        FT_RB_TY(FT_GO)
#undef FT_GO
    }
    rb_launch_it<RB_Q>(L, s);
}

hipError_t rb_iter(const RbLaunch& L, hipStream_t s) {
    rb_launch_it<RB_P>(L, s);
    if (L.staged < 2) rb_launch_it<RB_Y>(L, s);   // the fused form computes Y itself
    rb_launch_q(L, s);
    if (!L.fold_alpha) hipLaunchKernelGGL((k_rb_fin<F_CG1>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);
    rb_launch_it<RB_U>(L, s);
    hipLaunchKernelGGL((k_rb_fin<F_CG2>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);
    return hipGetLastError();
}

hipError_t rb_update(const RbLaunch& L, hipStream_t s) {
    rb_launch<RB_UPD>(L, s);
    rb_launch<RB_Y64>(L, s);
    rb_launch<RB_R64>(L, s);
    hipLaunchKernelGGL((k_rb_fin<F_MIXRES>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);
    return hipGetLastError();
}

hipError_t rb_pass_only(const RbLaunch& L, int k, hipStream_t s) {
    switch (k) {
        case 0: rb_launch_it<RB_P>(L, s); break;
        case 1:
            if (L.staged < 2) rb_launch_it<RB_Y>(L, s);
            break;
        case 2: rb_launch_q(L, s); break;
        case 3: rb_launch_it<RB_U>(L, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rb_finish(const RbLaunch& L, hipStream_t s) {
    rb_launch<RB_SCATTER>(L, s);
    return hipGetLastError();
}

}  // namespace ft
