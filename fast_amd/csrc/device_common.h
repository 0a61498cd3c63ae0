// Device helpers shared by the gfx950 kernels of the full-grid solve (tau_kernels.hip) and of
// the red-black Schur-complement solve (rb_kernels.hip): the branch-free row pieces of the
// operator, the diagonal tables, double-double accumulation and deterministic reductions,
// vector addressing, and the per-volume scalar recurrences.  Everything is internal to the
// including translation unit.
#pragma once
#include "tau_internal.h"

#include <cfloat>

namespace ft {
namespace {

constexpr int kPart = 8;         // partial-sum doubles per tile: up to 4 sums as (hi, lo) pairs
constexpr int kFinThreads = 256; // threads of a per-volume finalize block

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
}  // namespace ft
