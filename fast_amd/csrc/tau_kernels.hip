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
#include "tau_internal.h"

#include <algorithm>
#include <cfloat>

namespace ft {

namespace {

constexpr int kPart = 8;         // partial-sum doubles per tile: up to 4 sums as (hi, lo) pairs
constexpr int kFinThreads = 256; // threads of a per-volume finalize block
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
enum FinKind : int {
    F_INIT_BICG = 0,
    F_RESTART,
    F_BICG1,
    F_BICG2,
    F_BICG3,
    F_INIT_CG,
    F_CG1,
    F_CG2,
    F_MIXRES,
};

__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int per = G >> 3, rem = G & 7;
    const int xcd = b & 7, idx = b >> 3;
    return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

__device__ __forceinline__ int dim_of(const VolDesc& vd, int a) { return a == 0 ? vd.nx : (a == 1 ? vd.ny : vd.nz); }

// Box-boundary flags of a voxel, bit j = face j (Dir order) lies on the volume boundary.
__device__ __forceinline__ int box_flags(int x, int y, int z, int nx, int ny, int nz) {
    return (x == nx - 1 ? 1 : 0) | (x == 0 ? 2 : 0) | (y == ny - 1 ? 4 : 0) | (y == 0 ? 8 : 0) |
           (z == nz - 1 ? 16 : 0) | (z == 0 ? 32 : 0);
}

// ---------------------------------------------------------------------------------------
// Branch-free row pieces.  A face without a coefficient contributes 0 to the diagonal and
// 0 * u to the row product.  The diagonal and the product accumulator start at +0 and can
// never become -0, so d - 0 == d and acc + (+/-0) == acc: the results are bit-identical to
// skipping the face, as the reference's CSR does.
// ---------------------------------------------------------------------------------------
// Select among coefficients by bit masks on their representations.  A plain ?: between two
// fields of one struct is folded by the compiler into a load with a computed address, which
// forces the whole context into scratch memory; the masked form stays in registers.
template <class T>
struct Bits;
template <>
struct Bits<double> {
    using type = unsigned long long;
};
template <>
struct Bits<float> {
    using type = unsigned int;
};

template <class T>
__device__ __forceinline__ T pick2(bool second, T a, T b) {
    using B = typename Bits<T>::type;
    const B m = (B)0 - (B)second;
    return __builtin_bit_cast(T, (B)((__builtin_bit_cast(B, a) & ~m) | (__builtin_bit_cast(B, b) & m)));
}

template <class T>
__device__ __forceinline__ T face_val(const Coef<T>& c, int own, int nbr) {
    const int s = own + nbr;
    return pick2(s == 2, pick2(s == 1, c.ci0, c.ci1), c.ci2);
}

// coefficient of face j as it enters the diagonal: interior face, half cell on a primary
// box face, nothing on a lateral box face (Diffusion.cuh:93-105)
template <class T>
__device__ __forceinline__ T diag_coef(const Coef<T>& c, int code, int bx, int primary, int j) {
    const int own = (code >> 6) & 1;
    const T inner = face_val(c, own, (code >> j) & 1);
    const T cb = pick2(own != 0, c.cb0, c.cb1);
    const T boxv = pick2((j >> 1) == primary, T(0), cb);
    return pick2(((bx >> j) & 1) != 0, inner, boxv);
}

// Diagonal of the row, subtraction order X+,X-,Y+,Y-,Z+,Z- (Diffusion.cuh:85-105).
template <class T>
__device__ __forceinline__ T diag_of(int code, int bx, int primary, const Coef<T>& c) {
    T d = T(0);
#pragma unroll
    for (int j = 0; j < 6; ++j) d -= diag_coef(c, code, bx, primary, j);
    return d;
}

template <class T>
__device__ __forceinline__ T rhs_of(int code, int k, int np, const Coef<T>& c) {
    const T cb = pick2(((code >> 6) & 1) != 0, c.cb0, c.cb1);
    T rhs = T(0);
    if (k == 0)
        rhs -= cb * c.cbeg;
    else if (k == np - 1)
        rhs -= cb * c.cend;
    return rhs;
}

template <class T>
__device__ __forceinline__ T x0_of(int k, int np, int sgn) {
    if (sgn == 1) return T(1.0f) - (T(k) / T(np + 1));
    return T(k) / T(np + 1);
}

// Diagonal and inverse diagonal of a row from the volume's tables (in LDS).  Rows on a y or z
// box face evaluate the formula: they fill whole rows or planes of the march, so for 64-wide
// tiles the branch is uniform over a wave.  The tables were filled by diag_of itself, so both
// ways give the same bits.
template <class C>
struct DiagInv {
    C d, iv;   // diagonal and its inverse
};

template <class C>
__device__ __forceinline__ DiagInv<C> diag_at(const C* tab, int code, int bx, int primary, const Coef<C>& c) {
    if (bx & 0x3C) {
        const C d = diag_of<C>(code, bx, primary, c);
        return DiagInv<C>{d, C(1) / d};
    }
    const int i = code | ((bx & 3) << 7);
    return DiagInv<C>{tab[i], tab[kTabEntries + i]};
}

// Coefficients of the + and - face of one axis for (own, nbr+, nbr-) = index bits (2, 0, 1):
// ci[own + nbr], the interior-face value (Diffusion.cuh:85-105).  Eight entries per volume,
// kept in LDS so a row reads its three pairs instead of selecting six values.
template <class T>
struct FacePair {
    T p, m;
};

template <class T>
__device__ __forceinline__ FacePair<T> face_pair(const Coef<T>& c, int e) {
    const int own = (e >> 2) & 1;
    return FacePair<T>{face_val(c, own, e & 1), face_val(c, own, (e >> 1) & 1)};
}

// Row product in CSR column order Z-,Y-,X-,diag,X+,Y+,Z+.  A box face has no off-diagonal
// coefficient in the reference; here its term is c * (+0): the plane march supplies +0 for
// every neighbour outside the box, c is finite and positive, and the accumulator (started
// at +0) is never -0, so the term leaves the sum unchanged bit for bit.
template <class T>
__device__ __forceinline__ T apply_row(int code, T dg, const FacePair<T>* ftab, T uzm, T uym, T uxm, T uc, T uxp,
                                       T uyp, T uzp) {
    const int o4 = (code >> 4) & 4;   // own solid flag (bit 6) as bit 2 of the pair index
    const FacePair<T> px = ftab[(code & 3) | o4];
    const FacePair<T> py = ftab[((code >> 2) & 3) | o4];
    const FacePair<T> pz = ftab[((code >> 4) & 3) | o4];
    T acc = T(0);
    acc += pz.m * uzm;
    acc += py.m * uym;
    acc += px.m * uxm;
    acc += dg * uc;
    acc += px.p * uxp;
    acc += py.p * uyp;
    acc += pz.p * uzp;
    return acc;
}

template <class A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Double-double accumulator of exact products (Dekker's add; fma gives the product error).
struct AccDD {
    double hi, lo;
    __device__ __forceinline__ void init() { hi = 0.0; lo = 0.0; }
    __device__ __forceinline__ void add2(double p, double e) {
        const double s = hi + p;
        const double v = s - hi;
        double t = (hi - (s - v)) + (p - v);
        t += lo + e;
        hi = s + t;
        lo = t - (hi - s);
    }
    __device__ __forceinline__ void add(double a, double b) {
        const double p = a * b;
        add2(p, fma(a, b, -p));
    }
    __device__ __forceinline__ void add(float a, float b) { add2((double)a * (double)b, 0.0); }   // exact
};

// Every solve accumulates its dot products in double-double: a dot then rounds to the same
// value for any reduction tree, so results do not depend on the tile shape or on the batch.
template <class T, class A>
struct AccOf {
    using type = AccDD;
};

template <class Acc>
__device__ __forceinline__ void wave_combine(Acc& a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double h = __shfl_xor(a.hi, off, 64);
        const double l = __shfl_xor(a.lo, off, 64);
        a.add2(h, l);
    }
}

// Deterministic block reduction of NP accumulators; thread 0 writes out[2k], out[2k+1].
template <class Acc, int NP, int NT>
__device__ __forceinline__ void block_reduce_store(Acc (&acc)[NP], double* out) {
    constexpr int NW = NT / 64;
    __shared__ double red[NW][NP > 0 ? NP : 1][2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        Acc a = acc[k];
        wave_combine(a);
        if (lane == 0) {
            red[w][k][0] = a.hi;
            red[w][k][1] = a.lo;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            Acc a;
            a.init();
            for (int i = 0; i < NW; ++i) a.add2(red[i][k][0], red[i][k][1]);
            out[2 * k] = a.hi;
            out[2 * k + 1] = a.lo;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Per-pass policies.  fetch(): raw loads of one voxel (vectors and stencil code; the centre
// also loads what only the epilogue needs).  compute(): the stencil input u at that voxel
// (and the centre values the epilogue needs).  epi(): outputs and partial sums at the centre.
// fetch and compute are split so the plane march can issue loads two planes ahead.
// ---------------------------------------------------------------------------------------
// Element i of one volume's vector: a uniform base pointer plus a 32-bit byte offset in the
// global address space, so loads and stores take the SGPR-base form with no 64-bit address
// arithmetic per element (a flat access would also be counted against the LDS counter and
// waited for at every barrier).  Every vector of a volume stays below 4 GiB (checked on the
// host).
template <class T>
__device__ __forceinline__ T ld(const T* base, unsigned i) {
    using GT = const __attribute__((address_space(1))) T;
    using GC = const __attribute__((address_space(1))) char;
    return *(GT*)((GC*)base + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st(T* base, unsigned i, T v) {
    using GT = __attribute__((address_space(1))) T;
    using GC = __attribute__((address_space(1))) char;
    *(GT*)((GC*)base + i * (unsigned)sizeof(T)) = v;
}

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
// Sum of the partials of tiles [base, base + n) in a fixed order; thread 0 gets the totals.
template <int NP, int NT>
__device__ __forceinline__ void reduce_volume(const double* part, int base, int n, double (&tot)[NP]) {
    __shared__ double red[2 * (NP > 0 ? NP : 1)];
    AccDD sums[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) sums[k].init();
    constexpr int U = 4;   // tiles in flight per thread
    for (int t0 = threadIdx.x; t0 < n; t0 += U * NT) {
        double v[U][2 * NP];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * NT;
            const double* p = part + (long long)(base + (t < n ? t : 0)) * kPart;
#pragma unroll
            for (int k = 0; k < 2 * NP; ++k) v[u][k] = (t < n) ? p[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < NP; ++k) sums[k].add2(v[u][2 * k], v[u][2 * k + 1]);
    }
    block_reduce_store<AccDD, NP, NT>(sums, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) tot[k] = red[2 * k] + red[2 * k + 1];
    }
}

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

// ---------------------------------------------------------------------------------------
// Per-volume scalar updates: one wave per volume reduces its tiles in a fixed order.
// These restate Eigen's BiCGSTAB / ConjugateGradient scalar recurrences (Eigen 3.3
// IterativeLinearSolvers, called from fastlib/src/solver/BICGSTAB.cpp:80-86 and
// CG.cpp:85-92; GPU mirrors fastlib/src/cuda/BICGSTAB.cu:93-213, CG.cu:68-155).
// ---------------------------------------------------------------------------------------
template <class A>
__device__ __forceinline__ void bicg_loop_head(VolState<A>* s, A rr0) {
    // while (|r|^2 > tol2 && i < maxIters) { rho_old = rho; rho = r0.r; restart?; beta = ... }
    if (!(s->rsq > s->tol2) || s->iter >= s->maxiter) {
        s->status = ST_DONE;
        return;
    }
    s->rho_old = s->rho;
    s->rho = rr0;
    if (fabs(s->rho) < s->eps2 * s->r0sq) {
        // plain solve: Eigen's restart with a fresh r0; inner solve of a refinement cycle:
        // end the cycle, the outer step recomputes the true residual and starts afresh
        s->status = (s->mode == 1) ? ST_DONE : ST_RESTART;
        return;
    }
    s->beta = (s->rho / s->rho_old) * (s->alpha / s->omega);
    s->status = ST_RUNNING;
}

// Mixed form, end of a BiCGSTAB iteration.  Reliable update (x64 += d, residual replaced by
// the FP64 true residual) when the recursive residual has dropped by delta since the last one,
// when it meets the target (the true residual decides) or at the iteration cap; the process
// then continues with p, v and the recursive r0.r (rho_rec), which are consistent with each
// other, rather than with r0 dotted with the replaced residual.  Eigen's rule (|r0.r| below
// eps^2 |r0|^2) and, when set, restart_cos also reset the shadow residual r0 = r at the
// update.  A non-finite residual, or one grown by blowup2 since the last update, drops the
// correction and starts afresh from x64.  BiCGSTAB's residual peaks by orders of magnitude
// on ill-conditioned volumes and recovers, so growth alone drops nothing by default.
template <class A>
__device__ __forceinline__ void mixed_after_iteration(VolState<A>* s, A rr0) {
    if (!isfinite(s->rsq) || s->rsq > s->blowup2 * s->rsq_ref) {
        s->status = ST_FRESH;
        s->discard = 1;
        return;
    }
    if (!(s->rsq > s->tol2) || s->rsq < s->delta2 * s->rsq_ref || s->iter >= s->maxiter) {
        s->status = ST_UPDATE;
        s->rho_rec = rr0;
        return;
    }
    if (!isfinite(rr0) || fabs(rr0) < s->eps2 * s->r0sq || fabs(rr0) < s->restart_cos * sqrt(s->r0sq * s->rsq)) {
        s->status = ST_UPDATE;
        s->reshadow = 1;
        return;
    }
    if (s->omega == A(0)) {
        s->status = ST_FRESH;
        return;
    }
    s->rho_old = s->rho;
    s->rho = rr0;
    s->beta = (s->rho / s->rho_old) * (s->alpha / s->omega);
    s->status = ST_RUNNING;
}

// Mixed form, CG: the same update / fresh-start rules; between updates the usual PCG
// recurrence (beta = r.z / previous r.z).
template <class A>
__device__ __forceinline__ void mixed_cg_after_iteration(VolState<A>* s, A rz) {
    s->iter += 1;
    if (!isfinite(s->rsq) || s->rsq > 1e4 * s->rsq_ref) {
        s->status = ST_FRESH;
        s->discard = 1;
        return;
    }
    if (!(s->rsq > s->tol2) || s->rsq < s->delta2 * s->rsq_ref || s->iter >= s->maxiter) {
        s->status = ST_UPDATE;
        return;
    }
    const A absold = s->absnew;
    s->absnew = rz;
    s->beta = s->absnew / absold;
    s->first = 0;
    s->status = ST_RUNNING;
}

// Scalar update of one volume after a pass, from the pass's dot products `tot` (run by one
// thread of the volume's last tile).
template <class A, int KIND>
__device__ void fin_update(VolState<A>* s, const A* tot) {
    switch (KIND) {
        case F_INIT_BICG: {
            s->rsq = tot[0];
            s->r0sq = tot[0];
            s->bsq = tot[1];
            s->rho = A(1);
            s->alpha = A(1);
            s->omega = A(1);
            s->iter = 0;
            s->restarts = 0;
            s->tol2 = s->tol * s->tol * s->bsq;
            if (s->bsq == A(0)) {
                s->zero_rhs = 1;
                s->status = ST_DONE;
                return;
            }
            bicg_loop_head(s, tot[0]);
            break;
        }
        case F_RESTART: {
            // r = b - A x; r0 = r; rho = r0_sqnorm = |r|^2; if (restarts++ == 0) i = 0;
            s->rho = tot[0];
            s->r0sq = tot[0];
            if (s->restarts++ == 0) s->iter = 0;
            s->beta = (s->rho / s->rho_old) * (s->alpha / s->omega);
            s->status = ST_RUNNING;
            break;
        }
        case F_BICG1:
            s->alpha = s->rho / tot[0];
            if (s->mode == 1 && !isfinite(s->alpha)) s->status = ST_FRESH;   // breakdown: start afresh
            break;
        case F_BICG2: {
            const A tmp = tot[1];
            s->omega = (tmp > A(0)) ? tot[0] / tmp : A(0);
            if (s->mode == 1 && !isfinite(s->omega)) s->status = ST_FRESH;
            break;
        }
        case F_BICG3: {
            s->rsq = tot[0];
            s->iter += 1;
            s->phase += 1;
            if (s->mode == 1)
                mixed_after_iteration(s, tot[1]);
            else
                bicg_loop_head(s, tot[1]);
            break;
        }
        case F_INIT_CG: {
            s->rsq = tot[0];
            s->bsq = tot[1];
            s->iter = 0;
            s->first = 1;
            if (s->bsq == A(0)) {
                s->zero_rhs = 1;
                s->status = ST_DONE;
                return;
            }
            const A tiny = (sizeof(A) == 8) ? A(DBL_MIN) : A(FLT_MIN);
            const A th = s->tol * s->tol * s->bsq;
            s->thresh = th > tiny ? th : tiny;
            if (s->rsq < s->thresh || s->maxiter <= 0) {
                s->status = ST_DONE;
                return;
            }
            s->absnew = tot[2];
            break;
        }
        case F_CG1:
            s->alpha = s->absnew / tot[0];
            if (s->mode == 1 && !isfinite(s->alpha)) s->status = ST_FRESH;
            break;
        case F_CG2: {
            s->rsq = tot[0];
            s->phase += 1;
            if (s->mode == 1) {
                mixed_cg_after_iteration(s, tot[1]);
                return;
            }
            if (s->rsq < s->thresh) {
                s->status = ST_DONE;
                return;
            }
            const A absold = s->absnew;
            s->absnew = tot[1];
            s->beta = s->absnew / absold;
            s->first = 0;
            s->iter += 1;
            if (s->iter >= s->maxiter) s->status = ST_DONE;
            break;
        }
        default:
            break;
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

__device__ void mixres_update(VolState<double>* s, const double* tot);

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

// Finalize of a reliable update / fresh start: the FP64 test decides whether the volume is
// finished; otherwise BiCGSTAB continues on the replaced residual (a fresh start takes the
// Eigen initial state with r0 = r).
__device__ void mixres_update(VolState<double>* s, const double* tot) {
    const int status = s->status;
    if (s->refinements == 0) {
        s->b64sq = tot[1];
        s->tol2 = s->otol * s->otol * s->b64sq;
    }
    s->refinements += 1;
    s->discard = 0;
    s->r64sq = tot[0];
    if (s->b64sq == 0.0 || !(s->r64sq > s->otol * s->otol * s->b64sq) || s->iter >= s->maxiter) {
        s->status = ST_IDLE;
        return;
    }
    const int reshadow = s->reshadow;
    s->reshadow = 0;
    s->rsq = tot[2];
    s->rsq_ref = tot[2];
    if (s->solver == 1) {   // CG: continue with beta from the replaced residual, or p = z afresh
        if (status == ST_FRESH) {
            if (s->refinements > 1) s->restarts += 1;
            s->first = 1;
            s->absnew = tot[3];
        } else {
            const double absold = s->absnew;
            s->absnew = tot[3];
            s->beta = s->absnew / absold;
            s->first = 0;
        }
        s->status = ST_RUNNING;
        return;
    }
    if (status == ST_FRESH) {
        if (s->refinements > 1) s->restarts += 1;
        s->rho = 1.0;
        s->alpha = 1.0;
        s->omega = 1.0;
    }
    if (status == ST_FRESH || reshadow) s->r0sq = tot[2];
    if (reshadow && status != ST_FRESH) s->restarts += 1;
    const double rho_new = (status != ST_FRESH && !reshadow && isfinite(s->rho_rec)) ? s->rho_rec : tot[3];
    if (status != ST_FRESH && !reshadow &&
        (!isfinite(rho_new) || fabs(rho_new) < s->restart_cos * sqrt(s->r0sq * s->rsq))) {
        s->status = ST_UPDATE;   // the replaced residual is too orthogonal to r0: reset r0 next
        s->reshadow = 1;
        return;
    }
    if (s->omega == 0.0) {
        s->status = ST_FRESH;
        return;
    }
    s->rho_old = s->rho;
    s->rho = rho_new;
    s->beta = (s->rho / s->rho_old) * (s->alpha / s->omega);
    s->status = ST_RUNNING;
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

__device__ __forceinline__ RbPos rb_pos(const RbDesc& d, int lt) {
    const int tix = lt % d.tiles_x;
    const int tiy = (lt / d.tiles_x) % d.tiles_y;
    const int tiz = lt / (d.tiles_x * d.tiles_y);
    RbPos p;
    p.k = tix * kRbTX + (int)(threadIdx.x % kRbTX);
    p.y = tiy * kRbTY + (int)(threadIdx.x / kRbTX);
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
template <int PART = 2>
__device__ __forceinline__ void rb_load_tables(const RbLaunch& L, int v, const Coef<double>& c, RbTabs& t) {
    const double* src = L.tab + (long long)v * 2 * kTabEntries;
    constexpr int lo = PART == 1 ? kTabEntries : 0, hi = PART == 0 ? kTabEntries : 2 * kTabEntries;
    for (int i = lo + (int)threadIdx.x; i < hi; i += kRbTX * kRbTY) t.tab[i] = src[i];
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
// (Z-, Y-, X-, X+, Y+, Z+); ghost terms are c * (+0) and leave the sum unchanged.
template <class U>
__device__ __forceinline__ double rb_sum(const FacePair<double>* ftab, int code, const U (&o)[6]) {
    const int o4 = (code >> 4) & 4;
    const FacePair<double> px = ftab[(code & 3) | o4];
    const FacePair<double> py = ftab[((code >> 2) & 3) | o4];
    const FacePair<double> pz = ftab[((code >> 4) & 3) | o4];
    double acc = 0.0;
    acc += pz.m * (double)o[0];
    acc += py.m * (double)o[1];
    acc += px.m * (double)o[2];
    acc += px.p * (double)o[3];
    acc += py.p * (double)o[4];
    acc += pz.p * (double)o[5];
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
//   P (red):   p = D^-1 r (first iteration) or D^-1 r + beta p_old
//   Y (black): Y = D_b^-1 (A_br p)
//   Q (red):   q = D_r p - A_rb Y = S p; partial p.q
//   U (red):   x += alpha p; r -= alpha q; partials r.r and r.(D^-1 r)
// Loads are issued two planes ahead of use into three register sets used in rotation (the loop
// is unrolled by 3), all unconditional: an invalid point reads the first interior slot, and
// ghosts make neighbour reads safe, so each plane waits only for its own loads.
template <int KIND>
__global__ __launch_bounds__(kRbTX* kRbTY) void k_rb_it(RbLaunch L) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const RbDesc d = L.vol[v];   // by value: kept in registers across the stores
    const VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    const Coef<double> c = L.coef[v];
    rb_load_tables<KIND == RB_Q ? 0 : 1>(L, v, c, tabs);   // Q reads D; P, Y and U read D^-1
    const RbPos ps = rb_pos(d, tile - d.tile_base);
    constexpr int kColour = (KIND == RB_Y) ? 1 : 0;
    const double alpha = vs->alpha, beta = vs->beta;
    const bool first = vs->first != 0;
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
        } else if (KIND == RB_Y) {
            rb_nbrs(p_new, d, ci, s, in.u);
        } else if (KIND == RB_Q) {
            in.a = ld(p_new, ci);
            rb_nbrs(Y, d, ci, s, in.w);
        } else {
            in.a = ld(p_new, ci);
            in.b = ld(X, ci);
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
        } else if (KIND == RB_Y) {
            st(Y, ci, rb_sum(tabs.ftab, code, in.u) * di.iv);
        } else if (KIND == RB_Q) {
            const double pv = (double)in.a;
            const double qv = di.d * pv - rb_sum(tabs.ftab, code, in.w);
            st(Q, ci, (float)qv);
            acc[0].add(pv, qv);
        } else {
            const double pv = (double)in.a;
            st(X, ci, (float)((double)in.b + alpha * pv));
            const float rs = (float)((double)in.c - alpha * (double)in.e);
            st(Rr, ci, rs);
            acc[0].add(rs, rs);
            acc[1].add((double)rs, di.iv * (double)rs);
        }
    };
    RbIn i0, i1, i2;
    load(ps.z0, i0);
    load(ps.z0 + 1, i1);
    // no early exit: a step past the last plane loads the first slot and does nothing, which
    // keeps the loop body straight (the compiler then waits for each set's own loads only)
    for (int z = ps.z0; z < ps.z1; z += 3) {
        load(z + 2, i2);
        proc(z, i0);
        load(z + 3, i0);
        proc(z + 1, i1);
        load(z + 4, i1);
        proc(z + 2, i2);
    }
    if (KIND == RB_Q) block_reduce_store<AccDD, 1, kRbTX * kRbTY>(reinterpret_cast<AccDD(&)[1]>(acc), L.part + (long long)tile * kPart);
    if (KIND == RB_U) block_reduce_store<AccDD, 2, kRbTX * kRbTY>(acc, L.part + (long long)tile * kPart);
}

// Q with the black Y planes staged in LDS: a block owns kRbTX slots x kRbTY rows of red
// voxels and marches z, keeping black planes z - 1, z, z + 1 over its tile plus a one-slot,
// one-row halo in a four-plane ring.  Per plane a thread loads at most two Y values (the halo
// plane in block-wide order) instead of six neighbours.  With four ring planes one barrier per
// plane suffices: the plane written at step z + 1 is not among those read at step z, and the
// one written at step z + 2 can only be written once every thread has passed step z + 1's
// barrier.  Same values, roundings and summation order as k_rb_it<RB_Q>.  (Staging p for the
// Y pass the same way measured slower than k_rb_it<RB_Y>: its loads are 4-byte.)
__global__ __launch_bounds__(kRbTX* kRbTY) void k_rb_q(RbLaunch L) {
    using U = double;
    constexpr int W = kRbTX + 2, PL = W * (kRbTY + 2);
    constexpr int NT = kRbTX * kRbTY;
    constexpr int RND = (PL + NT - 1) / NT;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const RbDesc d = L.vol[v];
    const VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    __shared__ U ring[4 * PL];
    const Coef<double> c = L.coef[v];
    rb_load_tables<0>(L, v, c, tabs);   // Q reads D only
    const int lt = tile - d.tile_base;
    const int k0 = (lt % d.tiles_x) * kRbTX;
    const int y0 = ((lt / d.tiles_x) % d.tiles_y) * kRbTY;
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
        const double qv = di.d * p64 - rb_sum(tabs.ftab, code, o);
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

// Y and Q fused: q = S p = D_r p - A_rb (D_b^-1 A_br p) for a block's red tile without
// storing Y.  A block owns kRbTX slots x kRbTY rows of red voxels and marches z.  It keeps red
// p planes over the tile plus a two-slot, two-row halo in a four-plane LDS ring, computes the
// black Y of planes over the tile plus a one-slot, one-row halo from them into a second
// four-plane ring, and forms q from that.  Per red voxel the pass reads p and the two codes
// and writes q (10 bytes instead of the 30 of the separate Y and Q passes: Y is FP64, written
// once and read back); the halo Y values are computed twice by neighbouring blocks instead.
// Step z computes Y(z + 1) from p(z .. z + 2), stores p(z + 3) over p(z - 1), passes one
// barrier and forms q(z) from Y(z - 1 .. z + 1): the slot written into holds a plane that was
// last read before the previous step's barrier.  Y and q take the same values, roundings and
// summation order as k_rb_it<RB_Y> followed by k_rb_it<RB_Q>, so results are bit-identical.
__global__ __launch_bounds__(kRbTX* kRbTY) void k_rb_yq(RbLaunch L) {
    constexpr int NT = kRbTX * kRbTY;
    constexpr int PW = kRbTX + 4, PPL = PW * (kRbTY + 4);   // p plane: slots k0-2 .. k0+65, rows y0-2 .. y0+5
    constexpr int YW = kRbTX + 2, YPL = YW * (kRbTY + 2);   // Y plane: slots k0-1 .. k0+64, rows y0-1 .. y0+4
    constexpr int PR = (PPL + NT - 1) / NT;
    constexpr int YR = (YPL + NT - 1) / NT;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const RbDesc d = L.vol[v];
    const VolState<double>* vs = L.st + v;
    if (vs->status != ST_RUNNING) return;
    __shared__ RbTabs tabs;
    __shared__ float pring[4 * PPL];
    __shared__ double yring[4 * YPL];
    const Coef<double> c = L.coef[v];
    rb_load_tables<2>(L, v, c, tabs);   // Y reads D_b^-1, Q reads D_r
    const int lt = tile - d.tile_base;
    const int k0 = (lt % d.tiles_x) * kRbTX;
    const int y0 = ((lt / d.tiles_x) % d.tiles_y) * kRbTY;
    const int z0 = (lt / (d.tiles_x * d.tiles_y)) * d.tz;
    const int z1 = min(z0 + d.tz, d.nz);
    const bool odd = (vs->phase & 1) != 0;
    const float* P = L.p[odd ? 0 : 1] + d.off_c;
    const uint8_t* code_r = d.code_c[0];
    const uint8_t* code_b = d.code_c[1];
    const int tid = (int)threadIdx.x;
    // p halo points of this thread: element within a plane, or ~0u outside the padded layout
    // (such points feed no Y that is computed)
    unsigned pofs[PR];
#pragma unroll
    for (int j = 0; j < PR; ++j) {
        const int i = tid + j * NT;
        const int kk = k0 - 2 + i % PW, yy = y0 - 2 + i / PW;
        pofs[j] = (i < PPL && kk >= -1 && kk <= d.hx && yy >= -1 && yy <= d.ny)
                      ? (unsigned)(kk + 1) + (unsigned)d.px * (unsigned)(yy + 1) : ~0u;
    }
    // element 0 of the padded layout is a ghost (+0): loads outside the layout read it
    auto p_load = [&](int zc, float (&h)[PR]) {
#pragma unroll
        for (int j = 0; j < PR; ++j) {
            const bool ok = pofs[j] != ~0u && zc >= -1 && zc <= d.nz;
            h[j] = ld(P, ok ? pofs[j] + d.pp * (unsigned)(zc + 1) : 0u);
        }
    };
    auto p_store = [&](int zc, const float (&h)[PR]) {
        float* dst = pring + (zc & 3) * PPL;
#pragma unroll
        for (int j = 0; j < PR; ++j) {
            const int i = tid + j * NT;
            if (i < PPL) dst[i] = h[j];
        }
    };
    // black Y points of this thread: slot kb, row yb of the Y plane
    int ykb[YR], yyb[YR];
#pragma unroll
    for (int j = 0; j < YR; ++j) {
        const int i = tid + j * NT;
        ykb[j] = k0 - 1 + i % YW;
        yyb[j] = y0 - 1 + i / YW;
    }
    auto y_valid = [&](int j, int zb, int& s, int& x) {
        const int i = tid + j * NT;
        s = (1 + yyb[j] + zb) & 1;
        x = 2 * ykb[j] + s;
        return i < YPL && ykb[j] >= 0 && yyb[j] >= 0 && yyb[j] < d.ny && zb >= 0 && zb < d.nz && x < d.nx;
    };
    auto yc_load = [&](int zb, int (&cb)[YR]) {
#pragma unroll
        for (int j = 0; j < YR; ++j) {
            int s, x;
            cb[j] = ld(code_b, y_valid(j, zb, s, x) ? rb_index(d, ykb[j], yyb[j], zb) : 0u);
        }
    };
    auto y_plane = [&](int zb, const int (&cb)[YR]) {
        const float* pm = pring + ((zb - 1) & 3) * PPL;
        const float* p0 = pring + (zb & 3) * PPL;
        const float* pp1 = pring + ((zb + 1) & 3) * PPL;
        double* dst = yring + (zb & 3) * YPL;
#pragma unroll
        for (int j = 0; j < YR; ++j) {
            const int i = tid + j * NT;
            if (i >= YPL) continue;
            int s, x;
            double yv = 0.0;   // ghosts and slots past the row end hold +0, as in the Y vector
            if (y_valid(j, zb, s, x)) {
                const int pc = (ykb[j] - k0 + 2) + (yyb[j] - y0 + 2) * PW;
                float u[6];
                u[0] = pm[pc];
                u[1] = p0[pc - PW];
                u[2] = p0[pc - 1 + s];
                u[3] = p0[pc + s];
                u[4] = p0[pc + PW];
                u[5] = pp1[pc];
                const int bx = box_flags(x, yyb[j], zb, d.nx, d.ny, d.nz);
                const DiagInv<double> di = diag_at<double>(tabs.tab, cb[j], bx, d.primary, c);
                yv = rb_sum(tabs.ftab, cb[j], u) * di.iv;
            }
            dst[i] = yv;
        }
    };
    const int k = k0 + (tid % kRbTX), y = y0 + (tid / kRbTX);
    const bool row_ok = k < d.hx && y < d.ny;
    const int ycol = (k - k0 + 1) + (y - y0 + 1) * YW;   // this thread's slot in a Y plane
    const int pcol = (k - k0 + 2) + (y - y0 + 2) * PW;   // and in a p plane
    auto q_valid = [&](int z, int& s, int& x) {
        s = (y + z) & 1;
        x = 2 * k + s;
        return row_ok && z < z1 && x < d.nx;
    };
    auto rc_load = [&](int z) {
        int s, x;
        return (int)ld(code_r, q_valid(z, s, x) ? rb_index(d, k, y, z) : 0u);
    };
    AccDD acc[1];
    acc[0].init();
    // p of this thread's red voxel in plane z, read before the step's barrier: the slot is
    // refilled with p(z + 4) in the next step, possibly while a slower thread forms q(z)
    auto p_own = [&](int z) { return pring[(z & 3) * PPL + pcol]; };
    auto q_plane = [&](int z, int code, float pv) {
        int s, x;
        if (!q_valid(z, s, x)) return;
        const int bx = box_flags(x, y, z, d.nx, d.ny, d.nz);
        const DiagInv<double> di = diag_at<double>(tabs.tab, code, bx, d.primary, c);
        const double* pl = yring + (z & 3) * YPL + ycol;
        double o[6];
        o[0] = yring[((z - 1) & 3) * YPL + ycol];
        o[1] = pl[-YW];
        o[2] = pl[s - 1];
        o[3] = pl[s];
        o[4] = pl[YW];
        o[5] = yring[((z + 1) & 3) * YPL + ycol];
        const double p64 = (double)pv;
        const double qv = di.d * p64 - rb_sum(tabs.ftab, code, o);
        st(L.q + d.off_c, rb_index(d, k, y, z), (float)qv);
        acc[0].add(p64, qv);
    };
    // prologue: p(z0-2 .. z0+1), then Y(z0-1) and Y(z0), then p(z0+2) over p(z0-2)
    {
        float h[PR];
        int cb[YR];
        for (int zc = z0 - 2; zc <= z0 + 1; ++zc) {
            p_load(zc, h);
            p_store(zc, h);
        }
        __syncthreads();
        yc_load(z0 - 1, cb);
        y_plane(z0 - 1, cb);
        yc_load(z0, cb);
        y_plane(z0, cb);
        p_load(z0 + 2, h);
        __syncthreads();
        p_store(z0 + 2, h);
        __syncthreads();
    }
    float ha[PR];
    int cba[YR], cbb[YR];
    int rca, rcb;
    yc_load(z0 + 1, cba);
    rca = rc_load(z0);
    // unrolled by 2 over alternating register sets, no early exit (a step past the last plane
    // computes a Y plane and stores a p plane nobody reads, and forms no q)
    for (int z = z0; z < z1; z += 2) {
        p_load(z + 3, ha);
        yc_load(z + 2, cbb);
        rcb = rc_load(z + 1);
        y_plane(z + 1, cba);
        const float pva = p_own(z);
        p_store(z + 3, ha);
        __syncthreads();
        q_plane(z, rca, pva);
        p_load(z + 4, ha);
        yc_load(z + 3, cba);
        rca = rc_load(z + 2);
        y_plane(z + 2, cbb);
        const float pvb = p_own(z + 1);
        p_store(z + 4, ha);
        __syncthreads();
        q_plane(z + 1, rcb, pvb);
    }
    block_reduce_store<AccDD, 1, NT>(acc, L.part + (long long)tile * kPart);
}

// The passes outside the Krylov loop, one thread per slot of a row, marching tz planes:
//   CODE:    per-colour stencil codes from the full-grid codes (set-up)
//   UPD:     x64 += x (kept unless discarded), x = 0                      [update / fresh start]
//   Y64:     black x64b = D_b^-1 (b_b - A_br x64), partial |b_b|^2       [update / fresh start]
//   R64:     red r = b_S - S x64 = (b_r - D_r x64) - A_rb x64b in FP64, stored in FP32;
//            partials |r64|^2, |b_r|^2, |r32|^2, r32.(D^-1 r32)           [update / fresh start]
//   SCATTER: both colours of x64 / x64b into the full-grid solution        [end]
// On the first update x64 is the reference's initial guess x0, which depends only on the
// primary coordinate: Y64 evaluates it for the red neighbours, R64 stores it.
template <int KIND>
__global__ __launch_bounds__(kRbTX* kRbTY) void k_rb(RbLaunch L) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= L.ntiles) return;
    const int v = L.tile_vol[tile];
    const RbDesc d = L.vol[v];
    VolState<double>* vs = L.st + v;
    if (KIND == RB_UPD || KIND == RB_Y64 || KIND == RB_R64) {
        const int status = vs->status;
        if (status != ST_UPDATE && status != ST_FRESH) return;
    }
    constexpr bool kTables = (KIND == RB_Y64 || KIND == RB_R64);
    __shared__ RbTabs tabs;
    const Coef<double> c = L.coef[v];
    if (kTables) rb_load_tables(L, v, c, tabs);
    const RbPos ps = rb_pos(d, tile - d.tile_base);
    constexpr int kColour = (KIND == RB_Y64) ? 1 : 0;
    const bool first_res = vs->refinements == 0;   // x64 not yet set: use x0
    const bool keep = vs->refinements > 0 && !vs->discard;
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
                if (keep) X64[ci] = X64[ci] + (double)X[ci];
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
    if (KIND == RB_Y64) block_reduce_store<AccDD, 1, kRbTX * kRbTY>(reinterpret_cast<AccDD(&)[1]>(acc), L.part_b + (long long)tile * kPart);
    if (KIND == RB_R64) block_reduce_store<AccDD, 4, kRbTX * kRbTY>(acc, L.part + (long long)tile * kPart);
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
            mixres_update(s, tot);
        }
        return;
    }
    if (s->status != ST_RUNNING) return;
    constexpr int NP = (FIN == F_CG2) ? 2 : 1;
    double tot[NP];
    reduce_volume<NP, kFinThreads>(L.part, d.tile_base, d.ntiles, tot);
    if (threadIdx.x == 0) {
        double ta[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NP; ++k) ta[k] = tot[k];
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

template <int KIND>
void rb_launch(const RbLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL((k_rb<KIND>), dim3(L.ntiles), dim3(kRbTX * kRbTY), 0, s, L);
}

template <int KIND>
void rb_launch_it(const RbLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL((k_rb_it<KIND>), dim3(L.ntiles), dim3(kRbTX * kRbTY), 0, s, L);
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
    if (L.staged == 2)
        hipLaunchKernelGGL(k_rb_yq, dim3(L.ntiles), dim3(kRbTX * kRbTY), 0, s, L);
    else if (L.staged)
        hipLaunchKernelGGL(k_rb_q, dim3(L.ntiles), dim3(kRbTX * kRbTY), 0, s, L);
    else
        rb_launch_it<RB_Q>(L, s);
}

hipError_t rb_iter(const RbLaunch& L, hipStream_t s) {
    rb_launch_it<RB_P>(L, s);
    if (L.staged != 2) rb_launch_it<RB_Y>(L, s);   // the fused pass computes Y itself
    rb_launch_q(L, s);
    hipLaunchKernelGGL((k_rb_fin<F_CG1>), dim3(L.nvol), dim3(kFinThreads), 0, s, L);
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
            if (L.staged != 2) rb_launch_it<RB_Y>(L, s);
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
