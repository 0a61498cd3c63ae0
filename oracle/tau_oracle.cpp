/*
 * tau_oracle.cpp -- CPU restatement of the reference's `fast tau --cpu` path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness checker for the
 * HIP library in fast_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * What it restates (all citations relative to the reference tree):
 *   - porosity          fastlib/src/volume/VolumeMeasures.cpp:288-292,
 *                       fastlib/src/cuda/Volume.cu:612-614 (count of mask==0),
 *                       fast/src/TauModule.h:62 (value rounded to float)
 *   - system assembly   fastlib/src/solver/Diffusion.cpp:16-35 (params),
 *                       :51-62 (clamped neighbour), :64-153 (host CSR, column
 *                       order Z-,Y-,X-,diag,X+,Y+,Z+, zero coefficients not
 *                       inserted), fastlib/src/cuda/Diffusion.cuh:54-148
 *                       (coefficients, Dirichlet/zero-flux faces, rhs, x0),
 *                       fastlib/src/volume/VolumeMeasures.cpp:61-69 (h = float
 *                       1/maxDim), fastlib/include/fastlib/volume/
 *                       VolumeMeasures.h:22 + fast/src/TauModule.h:70 (the
 *                       diffusivities are a glm::vec2, i.e. float 1.0f/0.001f)
 *   - BiCGSTAB          fastlib/src/solver/BICGSTAB.cpp:70-98 -> Eigen::BiCGSTAB
 *                       with DiagonalPreconditioner (Eigen 3.3.x
 *                       IterativeLinearSolvers/BiCGSTAB.h, restated below; the
 *                       reference's GPU mirror is fastlib/src/cuda/BICGSTAB.cu:24-230)
 *   - CG                fastlib/src/solver/CG.cpp:75-104 -> Eigen::ConjugateGradient
 *                       <RowMajor, Lower|Upper> (Eigen 3.3.x ConjugateGradient.h,
 *                       restated below; GPU mirror fastlib/src/cuda/CG.cu:20-171)
 *   - flux + tau        fastlib/src/volume/VolumeMeasures.cpp:157-159,188-280
 *                       (useNonUniform: :228-231 pore voxels only, :242-245 half
 *                       spacing towards a solid neighbour, :261-264 plane mean
 *                       over the pore count)
 *   - reachable pores   fastlib/src/cuda/VolumeCCL.cu:86-170 (toggle spans),
 *                       :204-298 (6-connectivity by span overlap), :326-375,547
 *                       (labels renumbered by the component's minimum linear
 *                       index), :735-818 (per-face label table),
 *                       fastlib/src/volume/VolumeSegmentation.cpp:14-24,66-90
 *
 * Eigen is not in the reference tree and not in this image; its algorithm is
 * restated from the published Eigen 3.3 sources.  The summation order of
 * Eigen's dot products is SIMD-dependent and unknowable here, so it is not
 * imitated.  Instead every dot product and every flux-plane sum is computed
 * order-independently: exact products (Dekker's TwoProduct) accumulated in
 * double-double and rounded once.  Any summation order then rounds to the same
 * double (barring a sum within ~2^-80 of a rounding boundary), which is what
 * lets the GPU path reproduce this oracle's Krylov trajectory bit for bit.  The
 * operator, rhs, x0, Jacobi inverse, elementwise updates and scalar
 * recurrences are restated exactly.
 *
 * Pins (tests/test_oracle_*.py): closed forms (all-pore tau = 1, slabs
 * parallel to flow, slabs across flow via an exact tridiagonal solve), the
 * exact-rational fixtures in tests/golden/ (dense elimination on tiny masks,
 * BFS labelling), symmetry/flip invariants and brute-force CCL.
 */
#include <cstdint>
#include <cstring>
#include <cmath>
#include <cfloat>
#include <vector>
#include <limits>
#include <algorithm>
#include <omp.h>

namespace {

enum { X_POS = 0, X_NEG = 1, Y_POS = 2, Y_NEG = 3, Z_POS = 4, Z_NEG = 5, DIR_NONE = 6 };

inline int dirIndex(int d) { return d / 2; }            // PrimitiveTypes.h:153-167
inline int dirSgn(int d) { return -((d % 2) * 2 - 1); } // PrimitiveTypes.h:169-171
inline int dirOf(int index, int sgn) {                   // PrimitiveTypes.h:173-177
    sgn = (sgn + 1) / 2;
    sgn = 1 - sgn;
    return index * 2 + sgn;
}

struct Dims {
    int r[3];
    size_t n() const { return (size_t)r[0] * r[1] * r[2]; }
    size_t lin(int x, int y, int z) const { return (size_t)x + (size_t)r[0] * ((size_t)y + (size_t)r[1] * z); }
};

// DiffusionSysParams (Diffusion.cuh:23-33) as set by setupSysParams (Diffusion.cpp:16-35)
struct SysParams {
    double concBegin, concEnd;
    double cellDim[3];
    double faceArea[3];
    int dirPrimary;
    int dir;
};

SysParams setupSysParams(const Dims& d, int dir) {
    SysParams s;
    const int maxDim = std::max(d.r[0], std::max(d.r[1], d.r[2]));
    const float h = 1.0f / (float)maxDim;               // vec3(1.0f / maxDim)
    const double high = 1.0, low = 0.0;
    s.concBegin = (dirSgn(dir) == 1) ? high : low;
    s.concEnd = (dirSgn(dir) == 1) ? low : high;
    for (int k = 0; k < 3; ++k) s.cellDim[k] = (double)h;
    s.faceArea[0] = s.cellDim[1] * s.cellDim[2];
    s.faceArea[1] = s.cellDim[0] * s.cellDim[2];
    s.faceArea[2] = s.cellDim[0] * s.cellDim[1];
    s.dirPrimary = dirIndex(dir);
    s.dir = dir;
    return s;
}

// d0 = double(coeffs[0]), d1 = double(coeffs[1]) with coeffs a float vec2 {1.0, 0.001}
const double kD0 = (double)1.0f;
const double kD1 = (double)0.001f;

template <class T>
struct Row {
    T v[7];   // indexed by Dir, DIR_NONE = diagonal
    T rhs;
    T x0;
};

// getSystemTopKernel (Diffusion.cuh:54-148) for one voxel.
template <class T>
Row<T> systemRow(const SysParams& p, const T diff[7], const int vox[3], const int res[3]) {
    Row<T> out;
    T coeffs[7];
    bool use[7];
    coeffs[DIR_NONE] = T(0);
    use[DIR_NONE] = true;
    for (int j = 0; j < DIR_NONE; ++j) {
        const int k = dirIndex(j);
        const int sgn = dirSgn(j);
        const T Dface = diff[j];
        T cellDist[3] = {T(p.cellDim[0]), T(p.cellDim[1]), T(p.cellDim[2])};
        use[j] = true;
        if ((vox[k] == 0 && sgn == -1) || (vox[k] == res[k] - 1 && sgn == 1)) {
            cellDist[k] = p.cellDim[k] * T(0.5);
            use[j] = false;
        }
        coeffs[j] = (Dface * p.faceArea[k]) / cellDist[k];
        if (use[j] || k == p.dirPrimary) coeffs[DIR_NONE] -= coeffs[j];
    }
    const int primaryRes = res[p.dirPrimary];
    T rhs = T(0);
    if (vox[p.dirPrimary] == 0) {
        const int d = dirOf(p.dirPrimary, -1);
        rhs -= coeffs[d] * p.concBegin;
    } else if (vox[p.dirPrimary] == primaryRes - 1) {
        const int d = dirOf(p.dirPrimary, 1);
        rhs -= coeffs[d] * p.concEnd;
    }
    out.rhs = rhs;
    const int primaryVox = vox[p.dirPrimary];
    if (dirSgn(p.dir) == 1)
        out.x0 = 1.0f - (primaryVox / T(primaryRes + 1));
    else
        out.x0 = (primaryVox / T(primaryRes + 1));
    for (int j = 0; j < DIR_NONE; ++j)
        if (!use[j]) coeffs[j] = T(0);
    for (int i = 0; i < 7; ++i) out.v[i] = coeffs[i];
    return out;
}

// Row-major CSR, each row sorted by column (Eigen inner vectors are sorted).
template <class T>
struct CSR {
    size_t n = 0;
    std::vector<int64_t> rowStart;   // n + 1
    std::vector<int64_t> col;
    std::vector<T> val;
    std::vector<T> b, x0;
};

template <class T>
void assemble(const uint8_t* mask, const Dims& d, int dir, int threads, CSR<T>& A) {
    const SysParams p = setupSysParams(d, dir);
    const size_t N = d.n();
    A.n = N;
    A.rowStart.assign(N + 1, 0);
    A.b.assign(N, T(0));
    A.x0.assign(N, T(0));
    // fixed 7 slots per row then compacted: keeps the parallel fill simple
    std::vector<int64_t> col7(N * 7, -1);
    std::vector<T> val7(N * 7, T(0));
    std::vector<int> cnt(N, 0);
    const int64_t stride[3] = {1, d.r[0], (int64_t)d.r[0] * d.r[1]};
    // Column order of insertion (Diffusion.cpp:88-90)
    const int colOrder[7] = {Z_NEG, Y_NEG, X_NEG, DIR_NONE, X_POS, Y_POS, Z_POS};
    auto dfun = [](uint8_t a) -> T { return (a > 0) ? T(kD1) : T(kD0); };
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.r[2]; ++z) {
        for (int y = 0; y < d.r[1]; ++y) {
            for (int x = 0; x < d.r[0]; ++x) {
                const int pos[3] = {x, y, z};
                const size_t i = d.lin(x, y, z);
                T diff[7];
                diff[DIR_NONE] = dfun(mask[i]);
                for (int j = 0; j < 6; ++j) {
                    int q[3] = {x, y, z};
                    const int k = dirIndex(j);
                    const int nv = q[k] + dirSgn(j);
                    if (nv >= 0 && nv < d.r[k]) q[k] = nv;   // clampedIndex (Diffusion.cpp:51-62)
                    diff[j] = (dfun(mask[d.lin(q[0], q[1], q[2])]) + diff[DIR_NONE]) * T(0.5);
                }
                const Row<T> row = systemRow<T>(p, diff, pos, d.r);
                int c = 0;
                for (int k = 0; k < 7; ++k) {
                    const int dd = colOrder[k];
                    if (row.v[dd] != T(0)) {
                        int64_t cc = (int64_t)i;
                        if (dd != DIR_NONE) cc += dirSgn(dd) * stride[dirIndex(dd)];
                        col7[i * 7 + c] = cc;
                        val7[i * 7 + c] = row.v[dd];
                        ++c;
                    }
                }
                cnt[i] = c;
                A.b[i] = row.rhs;
                A.x0[i] = row.x0;
            }
        }
    }
    for (size_t i = 0; i < N; ++i) A.rowStart[i + 1] = A.rowStart[i] + cnt[i];
    A.col.resize(A.rowStart[N]);
    A.val.resize(A.rowStart[N]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)N; ++i) {
        for (int c = 0; c < cnt[i]; ++c) {
            A.col[A.rowStart[i] + c] = col7[i * 7 + c];
            A.val[A.rowStart[i] + c] = val7[i * 7 + c];
        }
    }
}

// Elementwise vector updates (Eigen runs them as single-thread SIMD loops) are split over the
// OpenMP threads: each element gets the same unfused operations in the same order, so the
// results are bit-identical to a serial loop, and 512^3 solves take minutes less per hundred
// iterations.
#define ELEMWISE _Pragma("omp parallel for schedule(static)")

// Eigen sparse (row-major) * dense: tmp = 0; tmp += a * x over the row's entries in order.
template <class T>
void spmv(const CSR<T>& A, const T* x, T* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)A.n; ++i) {
        T tmp = T(0);
        for (int64_t e = A.rowStart[i]; e < A.rowStart[i + 1]; ++e) tmp += A.val[e] * x[A.col[e]];
        y[i] = tmp;
    }
}

// DiagonalPreconditioner::factorize (Eigen): 1/diag, or 1 if the diagonal is missing or zero.
template <class T>
std::vector<T> invDiag(const CSR<T>& A) {
    std::vector<T> inv(A.n, T(1));
    for (size_t i = 0; i < A.n; ++i) {
        for (int64_t e = A.rowStart[i]; e < A.rowStart[i + 1]; ++e) {
            if (A.col[e] == (int64_t)i) {
                if (A.val[e] != T(0)) inv[i] = T(1) / A.val[e];
                break;
            }
        }
    }
    return inv;
}

// ---- order-independent sums: exact products, double-double accumulation ----------
struct DD {
    double hi = 0.0, lo = 0.0;
    // (hi, lo) += (p, e): TwoSum of the high parts, then renormalise (Dekker 1971)
    void add2(double p, double e) {
        const double s = hi + p;
        const double v = s - hi;
        double t = (hi - (s - v)) + (p - v);
        t += lo + e;
        hi = s + t;
        lo = t - (hi - s);
    }
    double value() const { return hi + lo; }
};

// Veltkamp split and Dekker's exact product a*b = p + e (no FMA needed)
inline void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    const double ca = 134217729.0 * a, cb = 134217729.0 * b;
    const double ah = ca - (ca - a), al = a - ah;
    const double bh = cb - (cb - b), bl = b - bh;
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

inline void dd_add_prod(DD& acc, double a, double b) {
    double p, e;
    two_prod(a, b, p, e);
    acc.add2(p, e);
}
inline void dd_add_prod(DD& acc, float a, float b) { acc.add2((double)a * (double)b, 0.0); }   // exact in double

template <class T>
T dot(const std::vector<T>& a, const std::vector<T>& b) {
    const int64_t n = (int64_t)a.size();
    const int nt = omp_get_max_threads();
    std::vector<DD> part(nt);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        DD acc;
        for (int64_t i = lo; i < hi; ++i) dd_add_prod(acc, a[i], b[i]);
        part[t] = acc;
    }
    DD tot;
    for (const DD& p : part) tot.add2(p.hi, p.lo);
    return T(tot.value());
}

template <class T>
T sqnorm(const std::vector<T>& a) { return dot(a, a); }

struct SolveOut {
    long iterations;
    double error;
    int success;     // Eigen ComputationInfo::Success
    long restarts;
};

// Eigen::internal::bicgstab (Eigen 3.3) with DiagonalPreconditioner.
template <class T>
SolveOut bicgstab(const CSR<T>& A, std::vector<T>& x, T tol, long maxIters) {
    const size_t n = A.n;
    const std::vector<T> inv = invDiag(A);
    std::vector<T> tmpv(n);
    std::vector<T> r(n);
    spmv(A, x.data(), tmpv.data());
    for (size_t i = 0; i < n; ++i) r[i] = A.b[i] - tmpv[i];
    std::vector<T> r0 = r;
    T r0_sqnorm = sqnorm(r0);
    const T rhs_sqnorm = sqnorm(A.b);
    SolveOut o{0, 0.0, 1, 0};
    if (rhs_sqnorm == T(0)) {
        std::fill(x.begin(), x.end(), T(0));
        o.iterations = 0;
        o.error = (double)tol;   // Eigen leaves tol_error untouched -> m_error == tolerance
        o.success = 1;
        return o;
    }
    T rho = 1, alpha = 1, w = 1;
    std::vector<T> v(n, T(0)), p(n, T(0)), y(n), z(n), s(n), t(n);
    const T tol2 = tol * tol * rhs_sqnorm;
    const T eps2 = std::numeric_limits<T>::epsilon() * std::numeric_limits<T>::epsilon();
    long i = 0, restarts = 0;
    while (sqnorm(r) > tol2 && i < maxIters) {
        const T rho_old = rho;
        rho = dot(r0, r);
        if (std::abs(rho) < eps2 * r0_sqnorm) {
            spmv(A, x.data(), tmpv.data());
            ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) r[k] = A.b[k] - tmpv[k];
            r0 = r;
            rho = r0_sqnorm = sqnorm(r);
            if (restarts++ == 0) i = 0;
        }
        const T beta = (rho / rho_old) * (alpha / w);
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) p[k] = r[k] + beta * (p[k] - w * v[k]);
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) y[k] = inv[k] * p[k];
        spmv(A, y.data(), v.data());
        alpha = rho / dot(r0, v);
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) s[k] = r[k] - alpha * v[k];
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) z[k] = inv[k] * s[k];
        spmv(A, z.data(), t.data());
        const T tmp = sqnorm(t);
        if (tmp > T(0))
            w = dot(t, s) / tmp;
        else
            w = T(0);
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) x[k] += alpha * y[k] + w * z[k];
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) r[k] = s[k] - w * t[k];
        ++i;
    }
    const T err = std::sqrt(sqnorm(r) / rhs_sqnorm);
    o.iterations = i;
    o.error = (double)err;
    o.success = (err <= tol) ? 1 : 0;
    o.restarts = restarts;
    return o;
}

// Eigen::internal::conjugate_gradient (Eigen 3.3), full matrix (Lower|Upper), Jacobi.
template <class T>
SolveOut cg(const CSR<T>& A, std::vector<T>& x, T tol, long maxIters) {
    const size_t n = A.n;
    const std::vector<T> inv = invDiag(A);
    std::vector<T> tmp(n), residual(n);
    spmv(A, x.data(), tmp.data());
    for (size_t i = 0; i < n; ++i) residual[i] = A.b[i] - tmp[i];
    const T rhsNorm2 = sqnorm(A.b);
    SolveOut o{0, 0.0, 1, 0};
    if (rhsNorm2 == T(0)) {
        std::fill(x.begin(), x.end(), T(0));
        o.iterations = 0;
        o.error = 0.0;
        o.success = 1;
        return o;
    }
    const T considerAsZero = std::numeric_limits<T>::min();
    const T threshold = std::max(tol * tol * rhsNorm2, considerAsZero);
    T residualNorm2 = sqnorm(residual);
    if (residualNorm2 < threshold) {
        o.iterations = 0;
        o.error = (double)std::sqrt(residualNorm2 / rhsNorm2);
        o.success = (T(o.error) <= tol) ? 1 : 0;
        return o;
    }
    std::vector<T> p(n), z(n);
    for (size_t i = 0; i < n; ++i) p[i] = inv[i] * residual[i];
    T absNew = dot(residual, p);
    long i = 0;
    while (i < maxIters) {
        spmv(A, p.data(), tmp.data());
        const T alpha = absNew / dot(p, tmp);
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) x[k] += alpha * p[k];
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) residual[k] -= alpha * tmp[k];
        residualNorm2 = sqnorm(residual);
        if (residualNorm2 < threshold) break;
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) z[k] = inv[k] * residual[k];
        const T absOld = absNew;
        absNew = dot(residual, z);
        const T beta = absNew / absOld;
        ELEMWISE for (int64_t k = 0; k < (int64_t)n; ++k) p[k] = z[k] + beta * p[k];
        ++i;
    }
    const T err = std::sqrt(residualNorm2 / rhsNorm2);
    o.iterations = i;
    o.error = (double)err;
    o.success = (err <= tol) ? 1 : 0;
    return o;
}

// Flux integral and tau (VolumeMeasures.cpp:188-280); nonuniform = useNonUniform.
template <class T>
T fluxTau(const uint8_t* mask, const Dims& d, int dir, const T* c, T porosity, double* dcOut, bool nonuniform) {
    const int primary = dirIndex(dir);
    const int sec[2] = {(primary + 1) % 3, (primary + 2) % 3};
    const int n = d.r[sec[0]] * d.r[sec[1]];
    int dpos[3] = {0, 0, 0};
    dpos[primary] = dirSgn(dir);
    const float coeffs[2] = {1.0f, 0.001f};
    T dcSumOverPlanes = 0;
    for (int k = 0; k < d.r[primary]; ++k) {
        DD sumD;   // order-independent plane sum (the reference sums sequentially in T)
        size_t poreCount = 0;
        for (int j = 0; j < d.r[sec[1]]; ++j) {
            for (int i = 0; i < d.r[sec[0]]; ++i) {
                int pos[3];
                pos[primary] = k;
                pos[sec[0]] = i;
                pos[sec[1]] = j;
                const size_t vi = d.lin(pos[0], pos[1], pos[2]);
                if (mask[vi] == 0) poreCount++;
                if (mask[vi] != 0 && nonuniform) continue;   // :228-231
                const T coeffHere = coeffs[(mask[vi] == 0) ? 0 : 1];
                const T concHere = c[vi];
                T concPrev = T(0.0);
                T concDx = T(0.5);
                const int q[3] = {pos[0] + dpos[0], pos[1] + dpos[1], pos[2] + dpos[2]};
                if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < d.r[0] && q[1] < d.r[1] && q[2] < d.r[2]) {
                    concPrev = c[d.lin(q[0], q[1], q[2])];
                    concDx = T(1.0);
                    if (nonuniform && mask[d.lin(q[0], q[1], q[2])] != 0) concDx = T(0.5);   // :242-245
                }
                const T term = coeffHere * (concHere - concPrev) / concDx;
                sumD.add2((double)term, 0.0);
            }
        }
        const T dcPlane = nonuniform ? T(sumD.value()) / T(poreCount) : T(sumD.value()) / n;   // :260-264
        dcSumOverPlanes += dcPlane;
    }
    const T dc = dcSumOverPlanes / d.r[primary];
    if (dcOut) *dcOut = (double)dc;
    return porosity / (dc * d.r[primary]);
}

// ---- reachable-pore labels -------------------------------------------------

// Voxels that the reference's span scan puts inside a span (VolumeCCL.cu:86-170):
// starting from prev = background, every change of value along x toggles "open".
void spanSet(const uint8_t* mask, const Dims& d, uint8_t background, std::vector<uint8_t>& S) {
    S.assign(d.n(), 0);
    for (int z = 0; z < d.r[2]; ++z)
        for (int y = 0; y < d.r[1]; ++y) {
            uint8_t prev = background;
            bool open = false;
            for (int x = 0; x < d.r[0]; ++x) {
                const size_t i = d.lin(x, y, z);
                if (mask[i] != prev) open = !open;
                S[i] = open ? 1 : 0;
                prev = mask[i];
            }
        }
}

}  // namespace

extern "C" {

struct oracle_result {
    double tau;
    double porosity;
    double error;
    double dc;
    long iterations;
    int status;    // 0 success, 1 no convergence, 2 bad argument
    int restarts;
};

/* #(mask == 0): Volume::sumZeroElems (Volume.cpp:261-272). */
uint64_t oracle_count_zero(const uint8_t* mask, int nx, int ny, int nz) {
    const size_t N = (size_t)nx * ny * nz;
    uint64_t c = 0;
    for (size_t i = 0; i < N; ++i) c += (mask[i] == 0);
    return c;
}

/* float(getPorosity<double>(volume)) as stored by TauModule.h:62. */
float oracle_porosity(const uint8_t* mask, int nx, int ny, int nz) {
    const double N = (double)((size_t)nx * ny * nz);
    return (float)((double)oracle_count_zero(mask, nx, ny, nz) / N);
}

/* The assembled system (FP64) for inspection: per voxel the 7 coefficients in Dir order
 * (X_POS..Z_NEG, diag; zero where not inserted), the rhs and x0. Any pointer may be NULL. */
void oracle_system_f64(const uint8_t* mask, int nx, int ny, int nz, int dir,
                       double* coef7, double* rhs, double* x0) {
    Dims d{{nx, ny, nz}};
    const SysParams p = setupSysParams(d, dir);
    auto dfun = [](uint8_t a) -> double { return (a > 0) ? kD1 : kD0; };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int pos[3] = {x, y, z};
                const size_t i = d.lin(x, y, z);
                double diff[7];
                diff[DIR_NONE] = dfun(mask[i]);
                for (int j = 0; j < 6; ++j) {
                    int q[3] = {x, y, z};
                    const int k = dirIndex(j);
                    const int nv = q[k] + dirSgn(j);
                    if (nv >= 0 && nv < d.r[k]) q[k] = nv;
                    diff[j] = (dfun(mask[d.lin(q[0], q[1], q[2])]) + diff[DIR_NONE]) * 0.5;
                }
                const Row<double> row = systemRow<double>(p, diff, pos, d.r);
                if (coef7) for (int j = 0; j < 7; ++j) coef7[i * 7 + j] = row.v[j];
                if (rhs) rhs[i] = row.rhs;
                if (x0) x0[i] = row.x0;
            }
}

/* y = A x through the assembled CSR (Eigen row-major product order). */
void oracle_apply_f64(const uint8_t* mask, int nx, int ny, int nz, int dir, const double* x, double* y) {
    Dims d{{nx, ny, nz}};
    CSR<double> A;
    assemble<double>(mask, d, dir, 0, A);
    spmv(A, x, y);
}
void oracle_apply_f32(const uint8_t* mask, int nx, int ny, int nz, int dir, const float* x, float* y) {
    Dims d{{nx, ny, nz}};
    CSR<float> A;
    assemble<float>(mask, d, dir, 0, A);
    spmv(A, x, y);
}

/* Eigen Jacobi inverse of the assembled diagonal (FP64). */
void oracle_invdiag_f64(const uint8_t* mask, int nx, int ny, int nz, int dir, double* inv) {
    Dims d{{nx, ny, nz}};
    CSR<double> A;
    assemble<double>(mask, d, dir, 0, A);
    std::vector<double> v = invDiag(A);
    std::memcpy(inv, v.data(), v.size() * sizeof(double));
}

/* getTortuosity<T> (VolumeMeasures.cpp:42-285) on the host path with porosity precomputed
 * as TauModule does.  solver: 0 = BiCGSTAB, 1 = CG (DiffusionSolverType).  precision:
 * 0 = double, 1 = float.  porosity < 0 computes float(#zero/N).  conc_out (nullable)
 * receives the solution as double.  max_iter_run > 0 caps the iterations actually run
 * (used only to time a bounded sample; the result is then not a converged tau). */
int oracle_tau(const uint8_t* mask, int nx, int ny, int nz, int dir, int solver, int precision,
               double tol, long max_iter, int threads, double porosity, double* conc_out,
               oracle_result* res, int nonuniform) {
    std::memset(res, 0, sizeof(*res));
    if (!mask || nx <= 0 || ny <= 0 || nz <= 0 || dir < 0 || dir > 5 || solver < 0 || solver > 1) {
        res->status = 2;
        return 2;
    }
    Dims d{{nx, ny, nz}};
    if (threads > 0) omp_set_num_threads(threads);
    const float por = (porosity < 0) ? oracle_porosity(mask, nx, ny, nz) : (float)porosity;
    res->porosity = (double)por;
    auto run = [&](auto tag) {
        using T = decltype(tag);
        CSR<T> A;
        assemble<T>(mask, d, dir, threads, A);
        std::vector<T> x = A.x0;
        SolveOut so = (solver == 0) ? bicgstab<T>(A, x, T(tol), max_iter) : cg<T>(A, x, T(tol), max_iter);
        res->iterations = so.iterations;
        res->error = so.error;
        res->restarts = (int)so.restarts;
        if (conc_out)
            for (size_t i = 0; i < A.n; ++i) conc_out[i] = (double)x[i];
        if (!so.success) {
            res->status = 1;
            res->tau = 0.0;
            return;
        }
        double dc = 0;
        const T tau = fluxTau<T>(mask, d, dir, x.data(), T((double)por), &dc, nonuniform != 0);
        res->dc = dc;
        res->tau = (double)tau;
        res->status = 0;
    };
    if (precision == 1)
        run(float(0));
    else
        run(double(0));
    return res->status;
}

/* Flux/tau of a given FP64 concentration field (flux formula alone). */
double oracle_flux_tau_f64(const uint8_t* mask, int nx, int ny, int nz, int dir, const double* c, double porosity,
                           int nonuniform) {
    Dims d{{nx, ny, nz}};
    return fluxTau<double>(mask, d, dir, c, porosity, nullptr, nonuniform != 0);
}

/* Connected-component labels of the span set (6-connected), renumbered 1..K by the
 * component's minimum linear index; background 0.  Returns numLabels = K + 1, like
 * VolumeCCL_Label (VolumeCCL.cu:381-620). */
uint32_t oracle_ccl(const uint8_t* mask, int nx, int ny, int nz, uint8_t background, uint32_t* labels) {
    Dims d{{nx, ny, nz}};
    std::vector<uint8_t> S;
    spanSet(mask, d, background, S);
    const size_t N = d.n();
    std::fill(labels, labels + N, 0u);
    std::vector<uint32_t> queue;
    queue.reserve(1024);
    uint32_t next = 0;
    for (size_t i = 0; i < N; ++i) {
        if (!S[i] || labels[i]) continue;
        ++next;
        labels[i] = next;
        queue.clear();
        queue.push_back((uint32_t)i);
        for (size_t h = 0; h < queue.size(); ++h) {
            const size_t v = queue[h];
            const int x = (int)(v % nx);
            const int y = (int)((v / nx) % ny);
            const int z = (int)(v / ((size_t)nx * ny));
            const int nb[6][3] = {{x + 1, y, z}, {x - 1, y, z}, {x, y + 1, z}, {x, y - 1, z}, {x, y, z + 1}, {x, y, z - 1}};
            for (int k = 0; k < 6; ++k) {
                const int a = nb[k][0], b = nb[k][1], c = nb[k][2];
                if (a < 0 || b < 0 || c < 0 || a >= nx || b >= ny || c >= nz) continue;
                const size_t u = d.lin(a, b, c);
                if (S[u] && !labels[u]) {
                    labels[u] = next;
                    queue.push_back((uint32_t)u);
                }
            }
        }
    }
    return next + 1;
}

/* Per-face label table (VolumeCCL_BoundaryLabels, VolumeCCL.cu:753-818): table[f*numLabels + l]
 * = 1 if label l > 0 occurs on face f (Dir order), row 6 = AND of the six rows. */
void oracle_boundary_table(const uint32_t* labels, int nx, int ny, int nz, uint32_t numLabels, uint8_t* table) {
    Dims d{{nx, ny, nz}};
    std::memset(table, 0, (size_t)7 * numLabels);
    for (int f = 0; f < 6; ++f) {
        const int k = dirIndex(f);
        const int fixed = (dirSgn(f) == 1) ? d.r[k] - 1 : 0;
        const int s0 = (k + 1) % 3, s1 = (k + 2) % 3;
        for (int b = 0; b < d.r[s1]; ++b)
            for (int a = 0; a < d.r[s0]; ++a) {
                int pos[3];
                pos[k] = fixed;
                pos[s0] = a;
                pos[s1] = b;
                const uint32_t l = labels[d.lin(pos[0], pos[1], pos[2])];
                if (l > 0) table[(size_t)f * numLabels + l] = 1;
            }
    }
    for (uint32_t l = 0; l < numLabels; ++l) {
        uint8_t all = 1;
        for (int f = 0; f < 6; ++f) all &= table[(size_t)f * numLabels + l];
        table[(size_t)6 * numLabels + l] = all;
    }
}

/* Reachable-pore mask for the axis of `dir`: 255 where the voxel's label touches both faces
 * of that axis (VolumeCCL::getDirMask(begin,end), VolumeSegmentation.cpp:14-24, rendered as
 * VolumeCCL_GenerateVolume, VolumeCCL.cu:821-860), else 0.  Returns numLabels. */
uint32_t oracle_reachable(const uint8_t* mask, int nx, int ny, int nz, int dir, uint8_t background,
                          uint8_t* out, uint32_t* labels_out) {
    Dims d{{nx, ny, nz}};
    const size_t N = d.n();
    std::vector<uint32_t> lab(N);
    const uint32_t numLabels = oracle_ccl(mask, nx, ny, nz, background, lab.data());
    std::vector<uint8_t> table((size_t)7 * numLabels);
    oracle_boundary_table(lab.data(), nx, ny, nz, numLabels, table.data());
    const int k = dirIndex(dir);
    const int fa = 2 * k, fb = 2 * k + 1;
    for (size_t i = 0; i < N; ++i) {
        const uint32_t l = lab[i];
        out[i] = (table[(size_t)fa * numLabels + l] && table[(size_t)fb * numLabels + l]) ? 255 : 0;
    }
    if (labels_out) std::memcpy(labels_out, lab.data(), N * sizeof(uint32_t));
    return numLabels;
}

}  // extern "C"
