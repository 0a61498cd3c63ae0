// gfx950 kernels of the matrix-free tortuosity solve.
//
// Operator.  Every voxel is an unknown; its row of the 7-point system is fully determined by
// a one-byte stencil code (bit j = solid flag of the neighbour across face j, Dir order
// X+,X-,Y+,Y-,Z+,Z-; bit 6 = own solid flag) and by which faces lie on the box boundary.
// This restates the reference assembly (fastlib/src/cuda/Diffusion.cuh:54-148,
// fastlib/src/solver/Diffusion.cpp:64-153) without storing the seven coefficient arrays:
//   off-diagonal(face j) = ci[own + nbr_j] for an interior face, absent on a box face,
//   diagonal            = 0 - c_0 - c_1 - ... - c_5 in Dir order over interior faces and
//                         primary-axis box faces (half-cell coefficient cb[own]),
//   rhs                 = -cb[own] * c_begin / c_end on the first / last primary plane,
//   x0                  = 1 - k/(Np+1) (positive direction) or k/(Np+1).
// A row product sums in the reference's CSR column order Z-,Y-,X-,diag,X+,Y+,Z+ starting
// from zero (Eigen's row-major sparse*dense), and the build uses -ffp-contract=off, so
// y = A x is bit-identical to the reference CPU operator.
//
// Tiling.  A block of TX x TY threads owns a TX x TY column brick of one volume and marches
// tz planes along z.  The stencil input of the current plane (plus a one-voxel halo) is
// staged in a double-buffered LDS plane; the planes below and above live in registers.
// Blocks are mapped to tiles XCD-aware: the hardware deals workgroups round-robin over the
// 8 XCDs, so consecutive tiles (which share halo rows) are given to workgroups on one XCD.
//
// Reductions.  Each tile writes its partial sums (fixed in-block tree); one wave per volume
// then reduces that volume's tiles in a fixed order and updates the device-resident scalars.
// Results therefore depend only on the volume and its tile shape, never on what else is in
// the batch, and are identical run to run.  Every dot product is accumulated as exact
// products (fma error term) in double-double and rounded once, so a dot rounds to the same
// value whatever the reduction tree: the Krylov trajectory is independent of the tile shape
// and of the batch, and matches a sequential CPU evaluation of the same recurrences bit for
// bit.
#include "device_common.h"

#include <algorithm>
#include <cfloat>

namespace ft {

namespace {

#ifndef FT_PREFETCH
#define FT_PREFETCH 2
#endif
constexpr int kPrefetch = FT_PREFETCH;   // planes the stencil loads run ahead of their use

enum PassKind : int {
    K_INIT_BICG = 0,
    K_INIT_CG,
    K_RESTART,
    K_BICG1,
    K_BICG2,
    K_BICG3,
    K_CG1,
    K_CG2,
    K_APPLY,
    K_MIXRES,
};

template <class T>
struct Raw {
    T a, b, c, e, f, g;
    int code;
};

template <class T>
struct Pt {
    T u, aux, e, f, g, dg, iv;   // iv = 1 / dg where a pass needs it
    int code;
};

// Everything a pass needs, held in registers: coefficients and scalars are copied at kernel
// entry (a pointer dereference inside the plane loop would be re-read from memory after every
// store, and its wait would also drain the prefetched loads), and the ping-pong buffers are
// resolved up front (a dynamically indexed array would live in scratch memory).
template <class T, class A>
struct PassCtx {
    using C = typename CmpOf<T, A>::type;
    Coef<C> c;
    const C* tab;                // diagonal tables of the volume (LDS)
    const FacePair<C>* ftab;     // face coefficient pairs (LDS)
    int primary, sgn, np, first;
    C alpha, beta, omega;
    const uint8_t* code;
    T* x;
    T* r;
    T* r0;
    T* p_old;   // p of the previous iteration (read)
    T* p_new;   // p of this iteration (written)
    T* v_old;
    T* v_new;
    T* s;       // BiCGSTAB's s = r - alpha v (pass 2 -> pass 3); the apply diagnostic's output
    // BiCGSTAB's p and v in their own storage type (the same buffers as p_old .. v_new)
    using D = typename DirOf<T, A>::type;
    D* bp_old;
    D* bp_new;
    D* bv_old;
    D* bv_new;
};

// Raw loads of the BiCGSTAB passes: residual-side vectors in T, directions in D.
template <class T, class D>
struct RawB {
    T a, e, f;
    D b, c;
    int code;
};

template <class T, class A, int KIND>
struct PassOps;

// INIT_BICG: u = x0; r = b - A x0, x = x0, r0 = r, p = v = 0
template <class T, class A>
struct PassOps<T, A, K_INIT_BICG> {
    using C = typename CmpOf<T, A>::type;
    using D = typename DirOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 2;   // |r|^2, |b|^2
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> p;
        p.code = w.code;
        p.u = x0_of<C>(k, pc.np, pc.sgn);
        p.dg = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c).d;
        return p;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        const C b = rhs_of<C>(pt.code, k, pc.np, pc.c);
        const C r = b - Au;
        st(pc.x, i, T(pt.u));
        st(pc.r, i, T(r));
        st(pc.r0, i, T(r));
        st(pc.bp_old, i, D(0));
        st(pc.bv_old, i, D(0));
        acc[0].add(r, r);
        acc[1].add(b, b);
    }
};

// INIT_CG: u = x0; r = b - A x0, x = x0
template <class T, class A>
struct PassOps<T, A, K_INIT_CG> {
    using C = typename CmpOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 3;   // |r|^2, |b|^2, r.(M^-1 r)
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> p;
        p.code = w.code;
        p.u = x0_of<C>(k, pc.np, pc.sgn);
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            p.dg = di.d;
            p.iv = di.iv;
        }
        return p;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        const C b = rhs_of<C>(pt.code, k, pc.np, pc.c);
        const C r = b - Au;
        const C inv = pt.iv;
        st(pc.x, i, T(pt.u));
        st(pc.r, i, T(r));
        acc[0].add(r, r);
        acc[1].add(b, b);
        acc[2].add(r, inv * r);
    }
};

// RESTART: u = x; r = b - A x, r0 = r
template <class T, class A>
struct PassOps<T, A, K_RESTART> {
    using C = typename CmpOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 1;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.x, i);
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> p;
        p.code = w.code;
        p.u = C(w.a);
        p.dg = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c).d;
        return p;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        const C b = rhs_of<C>(pt.code, k, pc.np, pc.c);
        const C r = b - Au;
        st(pc.r, i, T(r));
        st(pc.r0, i, T(r));
        acc[0].add(r, r);
    }
};

// APPLY (diagnostic): y = A x
template <class T, class A>
struct PassOps<T, A, K_APPLY> {
    using C = typename CmpOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 0;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.x, i);
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> p;
        p.code = w.code;
        p.u = C(w.a);
        p.dg = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c).d;
        return p;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&)[1]) {
        st(pc.s, i, T(Au));
    }
};

// BICG1: p = r + beta (p_old - omega v_old); u = M^-1 p; v = A u; partial r0.v
template <class T, class A>
struct PassOps<T, A, K_BICG1> {
    using C = typename CmpOf<T, A>::type;
    using D = typename DirOf<T, A>::type;
    using R = RawB<T, D>;   // raw loads in the storage types; widened in compute()
    static constexpr int NP = 1;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.r, i);
        w.b = ld(pc.bp_old, i);
        w.c = ld(pc.bv_old, i);
        if (centre) w.e = ld(pc.r0, i);
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> q;
        q.code = w.code;
        // p as it will be stored: the stencil and the x update must see the same direction
        const C p = C(D(C(w.a) + pc.beta * (C(w.b) - pc.omega * C(w.c))));
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            q.dg = di.d;
            q.iv = di.iv;
        }
        q.u = q.iv * p;
        q.aux = p;
        q.e = C(w.e);
        return q;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        st(pc.bp_new, i, D(pt.aux));
        st(pc.bv_new, i, D(Au));
        acc[0].add(pt.e, Au);
    }
};

// BICG2: s = r - alpha v (stored); u = M^-1 s; t = A u; partials t.s, t.t.  t is not stored:
// BICG3 recomputes it from s.  s is rounded to the storage type first, so that both passes
// see the same s.
template <class T, class A>
struct PassOps<T, A, K_BICG2> {
    using C = typename CmpOf<T, A>::type;
    using D = typename DirOf<T, A>::type;
    using R = RawB<T, D>;   // raw loads in the storage types; widened in compute()
    static constexpr int NP = 2;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.r, i);
        w.b = ld(pc.bv_new, i);
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> q;
        q.code = w.code;
        const C s = C(T(C(w.a) - pc.alpha * C(w.b)));
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            q.dg = di.d;
            q.iv = di.iv;
        }
        q.u = q.iv * s;
        q.aux = s;
        return q;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        st(pc.s, i, T(pt.aux));
        acc[0].add(Au, pt.aux);
        acc[1].add(Au, Au);
    }
};

// BICG3: t = A M^-1 s (recomputed from the stored s); x += alpha M^-1 p + omega M^-1 s;
// r = s - omega t; partials r.r, r0.r of the stored residual.  The stencil input is s, never
// r, so writing r in place cannot race with a neighbouring tile's halo reads.
template <class T, class A>
struct PassOps<T, A, K_BICG3> {
    using C = typename CmpOf<T, A>::type;
    using D = typename DirOf<T, A>::type;
    using R = RawB<T, D>;   // raw loads in the storage types; widened in compute()
    static constexpr int NP = 2;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.s, i);
        if (centre) {
            w.c = ld(pc.bp_new, i);
            w.e = ld(pc.r0, i);
            w.f = ld(pc.x, i);
        }
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> q;
        q.code = w.code;
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            q.dg = di.d;
            q.iv = di.iv;
        }
        q.u = q.iv * C(w.a);
        q.aux = C(w.a);
        q.e = C(w.e);
        q.f = C(w.f);
        q.g = C(w.c);
        return q;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        const C yv = pt.iv * pt.g;
        st(pc.x, i, T(pt.f + (pc.alpha * yv + pc.omega * pt.u)));
        const T rs = T(pt.aux - pc.omega * Au);   // the stored residual is what the next iteration sees
        st(pc.r, i, rs);
        acc[0].add(C(rs), C(rs));
        acc[1].add(pt.e, C(rs));
    }
};

// CG1: z = M^-1 r; p = z (first iteration) or z + beta p_old; partial p.(A p).  A p is not
// stored: CG2 recomputes it from the stored p, which costs no extra HBM traffic (CG2 reads p
// anyway) and saves writing and re-reading q.
template <class T, class A>
struct PassOps<T, A, K_CG1> {
    using C = typename CmpOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 1;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.r, i);
        w.b = ld(pc.p_old, i);   // unused on the first iteration (an unconditional load stays in flight)
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> q;
        q.code = w.code;
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            q.dg = di.d;
            q.iv = di.iv;
        }
        const C z = q.iv * C(w.a);
        // p as it will be stored: the stencil and the x update must see the same direction
        q.u = C(T(pc.first ? z : z + pc.beta * C(w.b)));
        return q;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        st(pc.p_new, i, T(pt.u));
        acc[0].add(pt.u, Au);
    }
};

// CG2: q = A p (recomputed from the stored p); x += alpha p; r -= alpha q; partials r.r and
// r.(M^-1 r) of the stored residual
template <class T, class A>
struct PassOps<T, A, K_CG2> {
    using C = typename CmpOf<T, A>::type;
    using R = Raw<T>;   // raw loads in the storage type; widened in compute()
    static constexpr int NP = 2;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        w.a = ld(pc.p_new, i);
        if (centre) {
            w.b = ld(pc.x, i);
            w.c = ld(pc.r, i);
        }
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<C> compute(const X& pc, const R& w, int bx, int k) {
        Pt<C> q;
        q.code = w.code;
        {
            const DiagInv<C> di = diag_at<C>(pc.tab, w.code, bx, pc.primary, pc.c);
            q.dg = di.d;
            q.iv = di.iv;
        }
        q.u = C(w.a);
        q.aux = C(w.b);
        q.e = C(w.c);
        return q;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, C Au, const Pt<C>& pt, Acc (&acc)[NP]) {
        st(pc.x, i, T(pt.aux + pc.alpha * pt.u));
        const T rs = T(pt.e - pc.alpha * Au);   // the stored residual is what the next iteration sees
        st(pc.r, i, rs);
        acc[0].add(C(rs), C(rs));
        acc[1].add(C(rs), pt.iv * C(rs));
    }
};

// ---------------------------------------------------------------------------------------
// The plane march.  A block owns one TX x TY column brick of a volume and walks its z
// planes.  Loads are issued two planes ahead (fetch), turned into stencil inputs one plane
// ahead (compute), and the current plane plus a one-voxel halo is staged in a
// double-buffered LDS plane; the planes below and above stay in registers.
// ---------------------------------------------------------------------------------------
template <class U, class Ops, class Ctx, class Acc, int NPA, int TX, int TY>
__device__ __forceinline__ void march(const Ctx& pc, const VolDesc& vd, int lt, U* lds, Acc (&acc)[NPA]) {
    using R = typename Ops::R;
    constexpr int W = TX + 2;
    constexpr int PL = (TY + 2) * W;
    constexpr int HWAVES = (2 * TX + 2 * TY + 63) / 64;   // leading waves that stage halo points
    const int nx = vd.nx, ny = vd.ny, nz = vd.nz, primary = vd.primary;
    const int tix = lt % vd.tiles_x;
    const int tiy = (lt / vd.tiles_x) % vd.tiles_y;
    const int tiz = lt / (vd.tiles_x * vd.tiles_y);
    const int x0 = tix * TX, y0 = tiy * TY, z0 = tiz * vd.tz;
    const int z1 = min(z0 + vd.tz, nz);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool inside = (x < nx) && (y < ny);

    // halo assignment: rows y0-1 and y0+TY, columns x0-1 and x0+TX
    const int h = threadIdx.x;
    int hx = 0, hy = 0, hl = -1;
    if (h < TX) {
        hx = x0 + h;
        hy = y0 - 1;
        hl = h + 1;
    } else if (h < 2 * TX) {
        hx = x0 + (h - TX);
        hy = y0 + TY;
        hl = (TY + 1) * W + (h - TX + 1);
    } else if (h < 2 * TX + TY) {
        hx = x0 - 1;
        hy = y0 + (h - 2 * TX);
        hl = (h - 2 * TX + 1) * W;
    } else if (h < 2 * TX + 2 * TY) {
        hx = x0 + TX;
        hy = y0 + (h - 2 * TX - TY);
        hl = (h - 2 * TX - TY + 1) * W + (TX + 1);
    }
    const bool hvalid = (hl >= 0) && hx >= 0 && hy >= 0 && hx < nx && hy < ny;
    const bool hwave = (threadIdx.x >> 6) < HWAVES;   // uniform over the wave

    // Every lane loads a voxel inside the volume (clamped coordinates, clamped plane), so the
    // loads carry no per-lane condition and stay in flight across the plane loop; what a lane
    // outside the volume (or past the last plane) loads is never used.
    const int xc = min(x, nx - 1), yc = min(y, ny - 1);
    const int hxc = min(max(hx, 0), nx - 1), hyc = min(max(hy, 0), ny - 1);
    const unsigned plane = (unsigned)nx * (unsigned)ny;
    const unsigned clin = (unsigned)xc + (unsigned)nx * (unsigned)yc;    // centre column
    const unsigned hlin = (unsigned)hxc + (unsigned)nx * (unsigned)hyc;  // halo column
    auto bxz = [&](int bxy, int pz) { return bxy | (pz == nz - 1 ? 16 : 0) | (pz == 0 ? 32 : 0); };
    const int cbxy = (x == nx - 1 ? 1 : 0) | (x == 0 ? 2 : 0) | (y == ny - 1 ? 4 : 0) | (y == 0 ? 8 : 0);
    const int hbxy = (hx == nx - 1 ? 1 : 0) | (hx == 0 ? 2 : 0) | (hy == ny - 1 ? 4 : 0) | (hy == 0 ? 8 : 0);
    auto kc = [&](int px, int py, int pz) { return primary == 0 ? px : (primary == 1 ? py : pz); };
    auto clampz = [&](int pz) { return pz < nz ? pz : nz - 1; };

    U u_prev = U(0), h_cur = U(0);
    Pt<U> c_cur{};
    // Loads run kPrefetch planes ahead of their use: rc[j] / rh[j] hold the centre / halo
    // loads of one plane, rotating through kPrefetch + 1 register sets (the loop is unrolled
    // by that many planes, so a load in flight is never copied; a copy would wait for it).
    constexpr int NB = kPrefetch + 1;
    R rc[NB], rh[NB];
    if (z0 > 0)
        u_prev = Ops::compute(pc, Ops::fetch(pc, clin + plane * (z0 - 1), false), bxz(cbxy, z0 - 1), kc(x, y, z0 - 1)).u;
    c_cur = Ops::compute(pc, Ops::fetch(pc, clin + plane * z0, true), bxz(cbxy, z0), kc(x, y, z0));
    if (hwave) h_cur = Ops::compute(pc, Ops::fetch(pc, hlin + plane * z0, false), bxz(hbxy, z0), kc(hx, hy, z0)).u;
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j) {
        rc[j] = Ops::fetch(pc, clin + plane * clampz(z0 + 1 + j), true);
        if (hwave) rh[j] = Ops::fetch(pc, hlin + plane * clampz(z0 + 1 + j), false);
    }

    // One plane z: `rc_in` / `rh_in` hold the loads of plane z+1; the loads of plane
    // z+1+kPrefetch go to `rc_out` / `rh_out`.
    auto step = [&](int z, const R& rc_in, R& rc_out, const R& rh_in, R& rh_out) {
        const bool has_next = (z + 1 < nz);
        const Pt<U> c_next = Ops::compute(pc, rc_in, bxz(cbxy, z + 1), kc(x, y, z + 1));
        U h_next = U(0);
        if (hwave) h_next = Ops::compute(pc, rh_in, bxz(hbxy, z + 1), kc(hx, hy, z + 1)).u;
        rc_out = Ops::fetch(pc, clin + plane * clampz(z + 1 + kPrefetch), true);
        if (hwave) rh_out = Ops::fetch(pc, hlin + plane * clampz(z + 1 + kPrefetch), false);
        U* buf = lds + (z & 1) * PL;
        buf[(ty + 1) * W + (tx + 1)] = inside ? c_cur.u : U(0);
        if (hl >= 0) buf[hl] = hvalid ? h_cur : U(0);
        __syncthreads();
        if (inside) {
            const U uzp = has_next ? c_next.u : U(0);
            const U Au = apply_row<U>(c_cur.code, c_cur.dg, pc.ftab, u_prev, buf[ty * W + (tx + 1)],
                                      buf[(ty + 1) * W + tx], c_cur.u, buf[(ty + 1) * W + (tx + 2)],
                                      buf[(ty + 2) * W + (tx + 1)], uzp);
            Ops::epi(pc, clin + plane * z, kc(x, y, z), Au, c_cur, acc);
        }
        u_prev = c_cur.u;
        c_cur = c_next;
        h_cur = h_next;
    };
    for (int z = z0; z < z1; z += NB) {
#pragma unroll
        for (int u = 0; u < NB; ++u)
            if (z + u < z1) step(z + u, rc[u], rc[(u + kPrefetch) % NB], rh[u], rh[(u + kPrefetch) % NB]);
    }
}

// ---------------------------------------------------------------------------------------
// End of a pass.  Every tile stores its partial sums; then one block per volume (k_fin)
// reduces that volume's partials in a fixed order and runs its scalar update.  (Finishing in
// the last tile of each volume instead would need a device-scope release fence in every
// tile, which writes back the XCD's L2 each time and costs more than the extra launch.)
// ---------------------------------------------------------------------------------------
// Copy the volume's diagonal tables into LDS and build its face-pair table there.
template <class C, int NT>
__device__ __forceinline__ void load_tables(const C* src, C* dst, FacePair<C>* fdst, const Coef<C>& c) {
    for (int i = threadIdx.x; i < 2 * kTabEntries; i += NT) dst[i] = src[i];
    if (threadIdx.x < 8) fdst[threadIdx.x] = face_pair(c, threadIdx.x);
    __syncthreads();
}

template <class T, class A, int TX, int TY, int KIND>
__global__ __launch_bounds__(TX* TY) void k_stencil(Launch<T, A> L, const T* xin, T* yout) {
    using Ops = PassOps<T, A, KIND>;
    constexpr int NP = Ops::NP;
    constexpr int NPA = NP > 0 ? NP : 1;
    using C = typename CmpOf<T, A>::type;
    __shared__ C lds[2 * (TY + 2) * (TX + 2)];
    __shared__ C tab[2 * kTabEntries];
    __shared__ FacePair<C> ftab[8];

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const VolDesc& vd = L.vol[v];
    const VolState<A>* st = L.st + v;
    if (KIND != K_APPLY && KIND != K_RESTART && st->status != ST_RUNNING) return;
    if (KIND == K_RESTART && st->status != ST_RESTART) return;

    PassCtx<T, A> pc;
    pc.c = L.coef[v];
    pc.tab = tab;
    pc.ftab = ftab;
    pc.primary = vd.primary;
    pc.sgn = vd.sgn;
    pc.np = dim_of(vd, vd.primary);
    pc.first = st->first;
    pc.alpha = C(st->alpha);
    pc.beta = C(st->beta);
    pc.omega = C(st->omega);
    pc.code = vd.code;
    {
        // per-volume base pointers: element offsets below are within the volume
        const bool odd = (st->phase & 1) != 0;
        const long long o = vd.off;
        pc.x = (T*)L.vs.x + o;
        pc.r = (T*)L.vs.r + o;
        pc.r0 = (T*)L.vs.r0 + o;
        pc.p_old = (T*)(odd ? L.vs.p[1] : L.vs.p[0]) + o;
        pc.p_new = (T*)(odd ? L.vs.p[0] : L.vs.p[1]) + o;
        pc.v_old = (T*)(odd ? L.vs.v[1] : L.vs.v[0]) + o;
        pc.v_new = (T*)(odd ? L.vs.v[0] : L.vs.v[1]) + o;
        using D = typename DirOf<T, A>::type;
        pc.bp_old = (D*)(odd ? L.vs.p[1] : L.vs.p[0]) + o;
        pc.bp_new = (D*)(odd ? L.vs.p[0] : L.vs.p[1]) + o;
        pc.bv_old = (D*)(odd ? L.vs.v[1] : L.vs.v[0]) + o;
        pc.bv_new = (D*)(odd ? L.vs.v[0] : L.vs.v[1]) + o;
        pc.s = (T*)L.vs.s + o;
        if (KIND == K_APPLY) {
            pc.x = (T*)xin + o;
            pc.s = yout + o;
        }
    }
    load_tables<C, TX * TY>(L.tab + (long long)v * 2 * kTabEntries, tab, ftab, pc.c);
    using Acc = typename AccOf<T, A>::type;
    Acc acc[NPA];
#pragma unroll
    for (int k = 0; k < NPA; ++k) acc[k].init();
    march<C, Ops, PassCtx<T, A>, Acc, NPA, TX, TY>(pc, vd, tile - vd.tile_base, lds, acc);
    if constexpr (NP > 0) block_reduce_store<Acc, NP, TX * TY>(acc, L.part + (long long)tile * kPart);
}

// ---------------------------------------------------------------------------------------
// Elementwise passes over the same tiles (no halo).
// ---------------------------------------------------------------------------------------
// Mixed form, reliable update of the volumes in ST_UPDATE / ST_FRESH: x64 += d (unless the
// correction is discarded), d = 0.  Elementwise over the same tiles (no halo).
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_mixupd(Launch<float, double> L) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const VolDesc& vd = L.vol[v];
    const VolState<double>* st = L.st + v;
    if (st->status != ST_UPDATE && st->status != ST_FRESH) return;
    const bool keep = st->refinements > 0 && !st->discard;
    const int nx = vd.nx, ny = vd.ny, nz = vd.nz;
    const int lt = tile - vd.tile_base;
    const int tix = lt % vd.tiles_x;
    const int tiy = (lt / vd.tiles_x) % vd.tiles_y;
    const int tiz = lt / (vd.tiles_x * vd.tiles_y);
    const int x = tix * TX + (int)(threadIdx.x % TX), y = tiy * TY + (int)(threadIdx.x / TX);
    const int z0 = tiz * vd.tz, z1 = min(z0 + vd.tz, nz);
    if (x >= nx || y >= ny) return;
    const long long plane = (long long)nx * ny;
    float* D = (float*)L.vs.x;
    for (int z = z0; z < z1; ++z) {
        const long long i = vd.off + (long long)x + (long long)nx * y + plane * z;
        if (keep) L.vs.x64[i] = L.vs.x64[i] + (double)D[i];
        D[i] = 0.0f;
    }
}

// partial sums each finalize reads
__host__ __device__ constexpr int np_of(int fin) {
    return fin == F_INIT_CG ? 3 : (fin == F_INIT_BICG || fin == F_BICG2 || fin == F_BICG3 || fin == F_CG2) ? 2 : 1;
}

// One block per volume: reduce the volume's tile partials of the pass just run (fixed order)
// and update its scalars.
template <class T, class A, int KIND>
__global__ __launch_bounds__(kFinThreads) void k_fin(Launch<T, A> L) {
    constexpr int NP = np_of(KIND);
    const int v = blockIdx.x;
    if (v >= L.nvol) return;
    VolState<A>* s = L.st + v;
    const int status = s->status;
    if (KIND == F_RESTART) {
        if (status != ST_RESTART) return;
    } else if (status != ST_RUNNING) {
        return;
    }
    const VolDesc& vd = L.vol[v];
    double tot[NP];
    reduce_volume<NP, kFinThreads>(L.part, vd.tile_base, vd.ntiles, tot);
    if (threadIdx.x == 0) {
        A ta[3] = {A(0), A(0), A(0)};
#pragma unroll
        for (int k = 0; k < NP; ++k) ta[k] = A(tot[k]);
        fin_update<A, KIND>(s, ta);
    }
}

template <class T, class A, int TX, int TY>
hipError_t launch_iter_bicg(const Launch<T, A>& L, hipStream_t s, hipEvent_t* ev = nullptr) {
    const dim3 g(L.ntiles), b(TX * TY), gf(L.nvol), bf(kFinThreads);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL((k_stencil<T, A, TX, TY, K_BICG1>), g, b, 0, s, L, (const T*)nullptr, (T*)nullptr);
    if (ev) (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL((k_fin<T, A, F_BICG1>), gf, bf, 0, s, L);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL((k_stencil<T, A, TX, TY, K_BICG2>), g, b, 0, s, L, (const T*)nullptr, (T*)nullptr);
    if (ev) (void)hipEventRecord(ev[3], s);
    hipLaunchKernelGGL((k_fin<T, A, F_BICG2>), gf, bf, 0, s, L);
    if (ev) (void)hipEventRecord(ev[4], s);
    hipLaunchKernelGGL((k_stencil<T, A, TX, TY, K_BICG3>), g, b, 0, s, L, (const T*)nullptr, (T*)nullptr);
    if (ev) (void)hipEventRecord(ev[5], s);
    hipLaunchKernelGGL((k_fin<T, A, F_BICG3>), gf, bf, 0, s, L);
    return hipGetLastError();
}

template <class T, class A, int TX, int TY>
hipError_t launch_iter_cg(const Launch<T, A>& L, hipStream_t s, hipEvent_t* ev = nullptr) {
    const dim3 g(L.ntiles), b(TX * TY), gf(L.nvol), bf(kFinThreads);
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL((k_stencil<T, A, TX, TY, K_CG1>), g, b, 0, s, L, (const T*)nullptr, (T*)nullptr);
    if (ev) (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL((k_fin<T, A, F_CG1>), gf, bf, 0, s, L);
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL((k_stencil<T, A, TX, TY, K_CG2>), g, b, 0, s, L, (const T*)nullptr, (T*)nullptr);
    if (ev) (void)hipEventRecord(ev[3], s);
    hipLaunchKernelGGL((k_fin<T, A, F_CG2>), gf, bf, 0, s, L);
    return hipGetLastError();
}

// Shape dispatch: the compiled block shapes.
#define FT_SHAPES(X) X(64, 8) X(64, 4) X(64, 16) X(32, 8) X(32, 16) X(32, 32) X(16, 16)

}  // namespace

// ---- mask preprocessing ----------------------------------------------------------------
// Work item w = (x chunk of 256, group of 4 rows), row = y + ny z: a thread builds the codes of
// one x in 4 rows.  Its 28 byte loads are unconditional and independent (a neighbour outside
// the box is loaded from the voxel itself, which is the reference's clamp), so they are all in
// flight at once; index arithmetic is 32-bit per work item (a 64-bit division per voxel made
// this kernel ~100x slower than its bandwidth).  Zeros are counted per thread and added with
// one atomic per block (atomics on one address serialise in L2).
__global__ __launch_bounds__(256) void k_build_code(const uint8_t* __restrict__ mask, uint8_t* __restrict__ code,
                                                    int nx, int ny, int nz, unsigned long long* zero_count) {
    constexpr int R = 4;
    const unsigned chunks = ((unsigned)nx + 255u) / 256u;
    const unsigned rows = (unsigned)ny * (unsigned)nz;
    const unsigned items = chunks * ((rows + R - 1) / R);
    const long long plane = (long long)nx * ny;
    unsigned long long zeros = 0;
    for (unsigned w = blockIdx.x; w < items; w += gridDim.x) {
        const unsigned rg = w / chunks;
        const int x = (int)((w - rg * chunks) * 256u + threadIdx.x);
        const int xc = min(x, nx - 1);
        uint8_t m[R][7];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const unsigned row = min(rg * R + j, rows - 1);
            const int y = (int)(row % (unsigned)ny), z = (int)(row / (unsigned)ny);
            const long long i = (long long)row * nx + xc;
            m[j][0] = mask[i];
            m[j][1] = mask[xc + 1 < nx ? i + 1 : i];
            m[j][2] = mask[xc > 0 ? i - 1 : i];
            m[j][3] = mask[y + 1 < ny ? i + nx : i];
            m[j][4] = mask[y > 0 ? i - nx : i];
            m[j][5] = mask[z + 1 < nz ? i + plane : i];
            m[j][6] = mask[z > 0 ? i - plane : i];
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const unsigned row = rg * R + j;
            if (x >= nx || row >= rows) continue;
            const int own = m[j][0] != 0;
            zeros += own ? 0ull : 1ull;
            code[(long long)row * nx + x] = (uint8_t)((m[j][1] != 0) | ((m[j][2] != 0) << 1) | ((m[j][3] != 0) << 2) |
                                                      ((m[j][4] != 0) << 3) | ((m[j][5] != 0) << 4) |
                                                      ((m[j][6] != 0) << 5) | (own << 6));
        }
    }
    zeros = wave_sum(zeros);
    __shared__ unsigned long long wz[4];
    if ((threadIdx.x & 63) == 0) wz[threadIdx.x >> 6] = zeros;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = wz[0] + wz[1] + wz[2] + wz[3];
        if (t) atomicAdd(zero_count, t);
    }
}

hipError_t build_code(const uint8_t* mask, uint8_t* code, int nx, int ny, int nz, unsigned long long* zero_count,
                      hipStream_t s) {
    const unsigned long long rows = (unsigned long long)ny * nz;
    const unsigned long long items = (((unsigned long long)nx + 255) / 256) * ((rows + 3) / 4);
    if (rows > 0xFFFFFFFFull || items > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)std::min<unsigned long long>(items, 2048);
    hipLaunchKernelGGL(k_build_code, dim3(blocks), dim3(256), 0, s, mask, code, nx, ny, nz, zero_count);
    return hipGetLastError();
}

// ---- diagonal tables ---------------------------------------------------------------------
namespace {
template <class C>
__global__ __launch_bounds__(256) void k_tables(const VolDesc* vol, const Coef<C>* coef, C* tab, int nvol) {
    const int v = blockIdx.x;
    if (v >= nvol) return;
    const Coef<C> c = coef[v];
    const int primary = vol[v].primary;
    C* t = tab + (long long)v * 2 * kTabEntries;
    for (int i = threadIdx.x; i < kTabEntries; i += 256) {
        const C d = diag_of<C>(i & 127, i >> 7, primary, c);   // x box faces from bits 7, 8
        t[i] = d;
        t[kTabEntries + i] = C(1) / d;
    }
}
}  // namespace

template <class T, class A>
hipError_t build_tables(const Launch<T, A>& L, hipStream_t s) {
    using C = typename CmpOf<T, A>::type;
    if (L.nvol <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_tables<C>), dim3(L.nvol), dim3(256), 0, s, L.vol, L.coef, L.tab, L.nvol);
    return hipGetLastError();
}
template hipError_t build_tables(const Launch<double, double>&, hipStream_t);
template hipError_t build_tables(const Launch<float, float>&, hipStream_t);
template hipError_t build_tables(const Launch<float, double>&, hipStream_t);

// ---- pass launchers ----------------------------------------------------------------------
#define FT_SHAPE_CASE(TXV, TYV, CALL) \
    if (L.tx == TXV && L.ty == TYV) { CALL(TXV, TYV); }

template <class T, class A>
hipError_t Passes<T, A>::init_bicg(const Launch<T, A>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                        \
    hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_INIT_BICG>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L,       \
                       (const T*)nullptr, (T*)nullptr);                                                          \
    hipLaunchKernelGGL((k_fin<T, A, F_INIT_BICG>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);                    \
    return hipGetLastError();
#define FT_TRY(TXV, TYV) FT_SHAPE_CASE(TXV, TYV, FT_CALL)
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::init_cg(const Launch<T, A>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                        \
    hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_INIT_CG>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L,         \
                       (const T*)nullptr, (T*)nullptr);                                                          \
    hipLaunchKernelGGL((k_fin<T, A, F_INIT_CG>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);                    \
    return hipGetLastError();
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::restart(const Launch<T, A>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                        \
    hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_RESTART>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L,         \
                       (const T*)nullptr, (T*)nullptr);                                                          \
    hipLaunchKernelGGL((k_fin<T, A, F_RESTART>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);                    \
    return hipGetLastError();
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::bicg_iter(const Launch<T, A>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV) return launch_iter_bicg<T, A, TXV, TYV>(L, s);
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::bicg_iter_timed(const Launch<T, A>& L, hipStream_t s, hipEvent_t* ev) {
#define FT_CALL(TXV, TYV) return launch_iter_bicg<T, A, TXV, TYV>(L, s, ev);
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::cg_iter(const Launch<T, A>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV) return launch_iter_cg<T, A, TXV, TYV>(L, s);
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::cg_iter_timed(const Launch<T, A>& L, hipStream_t s, hipEvent_t* ev) {
#define FT_CALL(TXV, TYV) return launch_iter_cg<T, A, TXV, TYV>(L, s, ev);
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::pass_only(const Launch<T, A>& L, int solver, int k, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                  \
    {                                                                                                      \
        const dim3 g(L.ntiles), b(TXV * TYV);                                                              \
        const T* n0 = nullptr;                                                                             \
        T* n1 = nullptr;                                                                                   \
        if (solver == 1 && k == 0) hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_CG1>), g, b, 0, s, L, n0, n1);   \
        if (solver == 1 && k == 1) hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_CG2>), g, b, 0, s, L, n0, n1);   \
        if (solver == 0 && k == 0) hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_BICG1>), g, b, 0, s, L, n0, n1); \
        if (solver == 0 && k == 1) hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_BICG2>), g, b, 0, s, L, n0, n1); \
        if (solver == 0 && k == 2) hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_BICG3>), g, b, 0, s, L, n0, n1); \
        return hipGetLastError();                                                                          \
    }
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

template <class T, class A>
hipError_t Passes<T, A>::apply(const Launch<T, A>& L, const T* x, T* y, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                        \
    hipLaunchKernelGGL((k_stencil<T, A, TXV, TYV, K_APPLY>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L, x, y);    \
    return hipGetLastError();
    FT_SHAPES(FT_TRY)
#undef FT_CALL
#undef FT_TRY
    return hipErrorInvalidValue;
}

template struct Passes<double, double>;
template struct Passes<float, float>;
template struct Passes<float, double>;

// ---- mixed precision: FP64 true residual feeding the FP32 inner system -----------------
namespace {

struct MixCtx {
    Coef<double> c;
    const double* tab;
    const FacePair<double>* ftab;
    int primary, sgn, np, first, fresh, reshadow, solver;
    const uint8_t* code;
    double* x64;
    float* r;
    float* r0;
    double* p_cur;   // BiCGSTAB's directions (FP64 in the mixed form)
    double* v_cur;
};

// FP64 true residual of the mixed form: u = x64 (x0 on the first step); r64 = b - A x64
// replaces the FP32 residual; a fresh start also resets r0 = r and p = v = 0.
struct MixRaw {
    double a;   // x64
    float e;    // r0
    int code;
};

struct MixOps {
    static constexpr int NP = 4;   // |r64|^2, |b|^2, |float(r64)|^2, r0.float(r64)
    using R = MixRaw;
    template <class X>
    __device__ static R fetch(const X& pc, unsigned i, bool centre) {
        R w;
        if (!pc.first) w.a = ld(pc.x64, i);
        if (centre && !pc.reshadow) w.e = ld(pc.r0, i);
        w.code = ld(pc.code, i);
        return w;
    }
    template <class X>
    __device__ static Pt<double> compute(const X& pc, const R& w, int bx, int k) {
        Pt<double> p;
        p.code = w.code;
        p.u = pc.first ? x0_of<double>(k, pc.np, pc.sgn) : w.a;
        {
            const DiagInv<double> di = diag_at<double>(pc.tab, w.code, bx, pc.primary, pc.c);
            p.dg = di.d;
            p.iv = di.iv;
        }
        p.e = (double)w.e;
        return p;
    }
    template <class X, class Acc>
    __device__ static void epi(const X& pc, unsigned i, int k, double Au, const Pt<double>& pt, Acc (&acc)[NP]) {
        const double b = rhs_of<double>(pt.code, k, pc.np, pc.c);
        const double r64 = b - Au;
        const float r32 = (float)r64;
        if (pc.first) st(pc.x64, i, pt.u);
        st(pc.r, i, r32);
        if (pc.reshadow) st(pc.r0, i, r32);
        if (pc.fresh && pc.solver == 0) {   // CG starts afresh with p = z (first), reading no p
            st(pc.p_cur, i, 0.0);
            st(pc.v_cur, i, 0.0);
        }
        acc[0].add(r64, r64);
        acc[1].add(b, b);
        acc[2].add(r32, r32);
        if (pc.solver == 1)
            acc[3].add((double)r32, pt.iv * (double)r32);   // CG: r.(M^-1 r)
        else
            acc[3].add(pc.reshadow ? (double)r32 : pt.e, (double)r32);   // BiCGSTAB: r0.r
    }
};


template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_mixres(Launch<float, double> L, const Coef<double>* coef64) {
    __shared__ double lds[2 * (TY + 2) * (TX + 2)];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const VolDesc& vd = L.vol[v];
    const VolState<double>* st = L.st + v;
    if (st->status != ST_UPDATE && st->status != ST_FRESH) return;
    __shared__ double tab[2 * kTabEntries];
    __shared__ FacePair<double> ftab[8];
    MixCtx pc;
    pc.c = coef64[v];
    pc.tab = tab;
    pc.ftab = ftab;
    pc.primary = vd.primary;
    pc.sgn = vd.sgn;
    pc.np = dim_of(vd, vd.primary);
    pc.first = (st->refinements == 0);
    pc.fresh = (st->status == ST_FRESH);
    pc.solver = st->solver;
    pc.reshadow = (pc.fresh || st->reshadow) && pc.solver == 0;
    pc.code = vd.code;
    const bool odd = (st->phase & 1) != 0;
    const long long o = vd.off;
    pc.x64 = L.vs.x64 + o;
    pc.r = (float*)L.vs.r + o;
    pc.r0 = (float*)L.vs.r0 + o;
    pc.p_cur = (double*)(odd ? L.vs.p[1] : L.vs.p[0]) + o;
    pc.v_cur = (double*)(odd ? L.vs.v[1] : L.vs.v[0]) + o;
    load_tables<double, TX * TY>(L.tab + (long long)v * 2 * kTabEntries, tab, ftab, pc.c);
    AccDD acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k].init();
    march<double, MixOps, MixCtx, AccDD, 4, TX, TY>(pc, vd, tile - vd.tile_base, lds, acc);
    block_reduce_store<AccDD, 4, TX * TY>(acc, L.part + (long long)tile * kPart);
}

// One block per volume: reduce the partials of the reliable-update pass and decide.
__global__ __launch_bounds__(kFinThreads) void k_fin_mixres(Launch<float, double> L) {
    const int v = blockIdx.x;
    if (v >= L.nvol) return;
    VolState<double>* s = L.st + v;
    if (s->status != ST_UPDATE && s->status != ST_FRESH) return;
    const VolDesc& vd = L.vol[v];
    double tot[4];
    reduce_volume<4, kFinThreads>(L.part, vd.tile_base, vd.ntiles, tot);
    if (threadIdx.x == 0) mixres_update(s, tot);
}


}  // namespace

hipError_t mixed_residual(const Launch<float, double>& L, const Coef<double>* coef64, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                    \
    hipLaunchKernelGGL((k_mixres<TXV, TYV>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L, coef64);             \
    hipLaunchKernelGGL(k_fin_mixres, dim3(L.nvol), dim3(kFinThreads), 0, s, L);                               \
    return hipGetLastError();
#define FT_TRY(TXV, TYV) FT_SHAPE_CASE(TXV, TYV, FT_CALL)
    FT_SHAPES(FT_TRY)
#undef FT_CALL
    return hipErrorInvalidValue;
}

hipError_t mixed_update(const Launch<float, double>& L, hipStream_t s) {
#define FT_CALL(TXV, TYV)                                                                                    \
    hipLaunchKernelGGL((k_mixupd<TXV, TYV>), dim3(L.ntiles), dim3(TXV * TYV), 0, s, L); \
    return hipGetLastError();
    FT_SHAPES(FT_TRY)
#undef FT_CALL
#undef FT_TRY
    return hipErrorInvalidValue;
}

// ---- flux integral -------------------------------------------------------------------------
// VolumeMeasures.cpp:188-272: per plane k along the primary axis, the sum over the plane of
// D_own * (c - c_next) / dx with c_next = 0, dx = 0.5 past the last plane in the flow
// direction, divided by the plane size.  One block per plane; per-thread partial sums in
// memory order, then the fixed block tree.
namespace {
template <class T>
__global__ __launch_bounds__(256) void k_flux_planes(const VolDesc* vol, const int* plane_vol, const int* plane_k,
                                                     const Coef<T>* coef, const T* conc, double* plane_out) {
    const int gp = blockIdx.x;
    const int v = plane_vol[gp];
    const VolDesc& vd = vol[v];
    const Coef<T> cf = coef[v];
    const int k = plane_k[gp];
    const int nx = vd.nx, ny = vd.ny, nz = vd.nz, p = vd.primary;
    const int dims[3] = {nx, ny, nz};
    const long long n = (long long)dims[(p + 1) % 3] * dims[(p + 2) % 3];
    const long long plane = (long long)nx * ny;
    const long long stride_p = p == 0 ? 1 : (p == 1 ? nx : plane);
    const int np = dims[p];
    const bool has_next = vd.sgn > 0 ? (k + 1 < np) : (k > 0);
    const long long step = vd.sgn > 0 ? stride_p : -stride_p;
    // enumerate the plane in memory order: two in-plane axes a (inner) and b (outer)
    const int ia = p == 0 ? 1 : 0;     // inner axis
    const int ib = p == 2 ? 1 : 2;     // outer axis
    const int na = dims[ia];
    const long long sa = ia == 0 ? 1 : nx;
    const long long sb = ib == 1 ? nx : plane;
    const bool nu = vd.flux_nu != 0;
    AccDD acc[1];
    acc[0].init();
    unsigned pores = 0;
    for (unsigned q = threadIdx.x; q < (unsigned)n; q += 256) {   // a plane has < 2^32 voxels
        const unsigned a = q % (unsigned)na, b = q / (unsigned)na;
        const long long lin = (long long)k * stride_p + (long long)a * sa + (long long)b * sb;
        const int code = vd.code[lin];
        const int own = (code >> 6) & 1;
        pores += own ? 0u : 1u;
        if (own && nu) continue;   // non-uniform mode: pore voxels only
        const T D = own ? cf.dflux1 : cf.dflux0;
        const T c = conc[vd.off + lin];
        T cn = T(0), dx = T(0.5);
        if (has_next) {
            cn = conc[vd.off + lin + step];
            dx = T(1.0);
            // non-uniform mode: half spacing towards a solid neighbour (the code's solid flag of
            // the neighbour in the flux direction: bit 0 for +x, 1 for -x, ... in Dir order)
            if (nu && ((code >> vd.dir) & 1)) dx = T(0.5);
        }
        const T term = D * (c - cn) / dx;
        acc[0].add2((double)term, 0.0);
    }
    double sum[2] = {0.0, 0.0};
    block_reduce_store<AccDD, 1, 256>(acc, sum);   // thread 0 receives the plane sum (hi, lo)
    pores = wave_sum(pores);
    __shared__ unsigned wp[4];
    if ((threadIdx.x & 63) == 0) wp[threadIdx.x >> 6] = pores;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long pc = (unsigned long long)wp[0] + wp[1] + wp[2] + wp[3];
        plane_out[gp] = (double)(T(sum[0] + sum[1]) / (nu ? T(pc) : T(n)));
    }
}
}  // namespace

template <class T>
hipError_t flux_planes(int count, const VolDesc* vol, const int* plane_vol, const int* plane_k, const Coef<T>* coef,
                       const T* conc, double* plane_out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_flux_planes<T>), dim3(count), dim3(256), 0, s, vol, plane_vol, plane_k, coef, conc, plane_out);
    return hipGetLastError();
}
template hipError_t flux_planes<double>(int, const VolDesc*, const int*, const int*, const Coef<double>*, const double*,
                                        double*, hipStream_t);
template hipError_t flux_planes<float>(int, const VolDesc*, const int*, const int*, const Coef<float>*, const float*,
                                       double*, hipStream_t);

}  // namespace ft
