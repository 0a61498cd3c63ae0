// Internal interface between the host driver (tau_api.cpp) and the gfx950 kernels
// (tau_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace ft {

// Per-volume solve status kept on the device.
// ST_UPDATE / ST_FRESH are the mixed form's requests for a reliable update (x64 += d and the
// FP32 residual replaced by the FP64 true residual) or a fresh BiCGSTAB start from x64.
enum : int { ST_RUNNING = 0, ST_DONE = 1, ST_RESTART = 2, ST_IDLE = 3, ST_UPDATE = 4, ST_FRESH = 5 };

// Stencil coefficients of one (direction, precision) system.  The operator is matrix-free:
// a face between voxels of solid flags (o, n) has coefficient ci[o+n]; a box face on the
// primary axis has the half-cell coefficient cb[o]; lateral box faces have none.
// The values are computed on the host exactly as the reference assembles them
// (fastlib/src/cuda/Diffusion.cuh:85-105, fastlib/src/solver/Diffusion.cpp:118-125).
template <class T>
struct Coef {
    T ci0, ci1, ci2;   // interior faces: pore-pore, pore-solid, solid-solid
    T cb0, cb1;        // primary box face of a pore / solid voxel
    T cbeg, cend;      // Dirichlet concentrations at the begin / end primary face
    T dflux0, dflux1;  // own-cell diffusivity of pore / solid in the flux integral
};

// One volume of a (possibly batched) solve.
struct VolDesc {
    int nx, ny, nz;
    int dir, primary, sgn;
    int tiles_x, tiles_y, tiles_z, tz;  // tile grid; tz = planes per tile (z chunk)
    int tile_base, ntiles;              // range in the launch's tile list
    int plane_base;                     // flux: first plane slot of this volume
    int flux_nu;                        // flux: TortuosityParams::useNonUniform
    long long off;                      // element offset in the concatenated vectors
    const uint8_t* code;                // per-voxel stencil code (bits 0-5 nbr solid, bit 6 own solid)
};

// Device-resident scalars of one solve (A = accumulation type).
template <class A>
struct VolState {
    A rho, rho_old, alpha, omega, beta;
    A r0sq, bsq, tol2, rsq, absnew;
    A tol, thresh, eps2;
    // mixed form (FP32 Krylov process with FP64 reliable updates)
    double r64sq, b64sq, otol;     // last FP64 |r|^2, |b|^2, relative tolerance on the true residual
    double rsq_ref, delta2;        // |r|^2 at the last update; update when |r|^2 < delta2 * rsq_ref
    double restart_cos;            // mixed BiCGSTAB: reset r0 = r when |r0.r| < restart_cos |r0| |r|
    double blowup2;                // mixed BiCGSTAB: drop the correction when |r|^2 > blowup2 * rsq_ref
    double rho_rec;                // mixed BiCGSTAB: r0.r of the recursive residual at the update
    long long maxiter;             // Krylov iteration cap
    int iter, status, restarts, first;
    int phase;                     // ping-pong parity counter (never reset)
    int zero_rhs;
    int mode;                      // 0 = plain solve, 1 = mixed form
    int refinements;               // mixed: FP64 residual replacements so far
    int discard;                   // mixed: drop d (non-finite) at the next update
    int reshadow;                  // mixed: the next update also resets r0 = r (Eigen's restart)
    int solver;                    // mixed: 0 = BiCGSTAB, 1 = CG
    int xpend;                     // red-black CG: steps alpha p that x still lacks (0..2), the
                                   // last xpend iterations'; the P pass adds two, the update any
    double xalpha[2];              // their alphas, oldest first
};

// Vectors of one (batched) solve, concatenated over its volumes (element offset VolDesc::off).
struct VecSet {
    void* x;
    void* r;
    void* r0;
    void* p[2];
    void* v[2];
    void* s;       // BiCGSTAB's s = r - alpha v; output of the apply diagnostic
    double* x64;   // mixed refinement: FP64 solution
};

// Arithmetic type of a pass: the storage type, except in the mixed form (FP32 vectors,
// FP64 accumulation) where every value is widened to FP64 after the load and rounded to FP32
// only on the store.  The row sum of the 7-point operator cancels by a factor of order N^2 on
// smooth fields, so FP32 arithmetic (or FP32-rounded coefficients) would perturb A x by
// ~1e-3 relative at 256^3 and break BiCGSTAB; with FP64 arithmetic only storage rounds.
template <class T, class A>
struct CmpOf {
    using type = T;
};
template <>
struct CmpOf<float, double> {
    using type = double;
};

// Storage type of BiCGSTAB's directions p and v.  In the mixed form they stay FP64: on
// ill-conditioned volumes BiCGSTAB's p and v grow to 1e5..1e7 times the residual between
// updates, and rounding them to FP32 (relative error 6e-8 of their own size) would perturb
// the residual by order one; the residual, shadow residual, s and the correction stay FP32.
template <class T, class A>
struct DirOf {
    using type = T;
};
template <>
struct DirOf<float, double> {
    using type = double;
};

// Launch description shared by all passes of one batched solve.
template <class T, class A>
struct Launch {
    using C = typename CmpOf<T, A>::type;
    int tx, ty;               // block shape: one of the compiled shapes
    int ntiles, nvol;
    const VolDesc* vol;       // device
    const int* tile_vol;      // device: tile -> volume
    const Coef<C>* coef;      // device: per volume, in the arithmetic type
    VolState<A>* st;          // device: per volume
    double* part;             // device: ntiles * 8 partial sums ((hi, lo) pairs)
    C* tab;                   // device: per volume, kTabEntries diagonals then as many inverses
    VecSet vs;
};

// Diagonal tables: entry (code | box_x << 7) holds the diagonal of a row with stencil code
// `code` whose only box faces are the x faces in box_x (bit 0: x = nx-1, bit 1: x = 0);
// entry kTabEntries + i holds 1 / entry i.  Rows on a y or z box face compute theirs.
constexpr int kTabEntries = 512;

// ---- launchers (all asynchronous on the given stream) ----
// Diagonal tables of every volume of a launch (L.tab), from L.coef and the volumes' primary axes.
template <class T, class A>
hipError_t build_tables(const Launch<T, A>& L, hipStream_t s);

hipError_t build_code(const uint8_t* mask, uint8_t* code, int nx, int ny, int nz,
                      unsigned long long* zero_count, hipStream_t s);

template <class T, class A>
struct Passes {
    // r = b - A x0, x = x0, r0 = r, p = v = 0; BiCGSTAB loop head.
    static hipError_t init_bicg(const Launch<T, A>&, hipStream_t);
    // r = b - A x0, x = x0; CG set-up.
    static hipError_t init_cg(const Launch<T, A>&, hipStream_t);
    // BiCGSTAB restart of volumes in ST_RESTART: r = b - A x, r0 = r.
    static hipError_t restart(const Launch<T, A>&, hipStream_t);
    // One BiCGSTAB iteration: 3 stencil passes, each followed by its per-volume finalize.
    static hipError_t bicg_iter(const Launch<T, A>&, hipStream_t);
    // Same, recording ev[2k] / ev[2k+1] around stencil pass k (k = 0, 1, 2).
    static hipError_t bicg_iter_timed(const Launch<T, A>&, hipStream_t, hipEvent_t* ev);
    // One PCG iteration: 2 stencil passes, each followed by its per-volume finalize.
    static hipError_t cg_iter(const Launch<T, A>&, hipStream_t);
    // Same, recording ev[2k] / ev[2k+1] around stencil pass k (k = 0, 1).
    static hipError_t cg_iter_timed(const Launch<T, A>&, hipStream_t, hipEvent_t* ev);
    // Stencil kernel of pass k alone (k = 0, 1 for CG; 0, 1, 2 for BiCGSTAB), no finalize
    // (timing: the volumes' scalars and statuses stay as they are).
    static hipError_t pass_only(const Launch<T, A>&, int solver, int k, hipStream_t);
    // y = A x (diagnostic; same device code as the solver).
    static hipError_t apply(const Launch<T, A>&, const T* x, T* y, hipStream_t);
};

// Mixed form.  One BiCGSTAB process runs with FP32 vectors and stencil (T = float) and FP64
// reductions and scalars (A = double) on the correction d; x64 accumulates the solution in
// FP64.  For every volume in ST_UPDATE or ST_FRESH:
//   mixed_update:   x64 += d (unless discarded), d = 0;
//   mixed_residual: x64 = x0 on the first step; r64 = b - A x64 with the FP64 operator
//                   (coef64); r = float(r64) replaces the recursive residual; on ST_FRESH
//                   also r0 = r, p = v = 0.  Then the FP64 test |r64| <= otol |b| and either
//                   ST_IDLE (finished) or ST_RUNNING with the BiCGSTAB scalars carried over.
hipError_t mixed_residual(const Launch<float, double>&, const Coef<double>* coef64, hipStream_t);
hipError_t mixed_update(const Launch<float, double>&, hipStream_t);

// ---- red-black reduced system (mixed form, CG) --------------------------------------------
// The 7-point operator couples each voxel only to voxels of the other colour (colour =
// (x + y + z) & 1).  Eliminating the black voxels leaves the Schur complement on the red ones,
//   S x_r = b_S,  S p = D_r p - A_rb D_b^-1 A_br p,  b_S = b_r - A_rb D_b^-1 b_b,
// and x_b = D_b^-1 (b_b - A_br x_r) afterwards.  Jacobi-preconditioned CG on S needs about
// half the iterations of Jacobi-PCG on the full system, on half the unknowns.  Vectors of one
// colour are stored compactly with a ghost layer: row (y, z) holds the hx = ceil(nx / 2)
// voxels x = 2k + s, s = (colour + y + z) & 1, at element (k + 1) + px (y + 1) + pp (z + 1),
// px = hx + 2, pp = px (ny + 2).  Ghost elements and slots with x >= nx hold 0 and are never
// written, so neighbour reads need no bounds tests.
struct RbDesc {
    int nx, ny, nz, hx;
    int dir, primary, sgn, px;           // px: row pitch of the padded layout
    int tiles_x, tiles_y, tiles_z, tz;   // tile grid over (k, y, z); tz = planes per tile
    int tile_base, ntiles;
    unsigned pp, pad0;                   // pp: plane pitch of the padded layout
    long long off_c;                     // element offset in the per-colour vectors
    long long off;                       // element offset in the full-grid solution
    const uint8_t* code;                 // full-grid stencil code
    const uint8_t* code_c[2];            // per-colour stencil codes (padded layout; red, black)
};

struct RbLaunch {
    int ntiles, nvol;
    int ty;           // rows per tile (2, 4, 8, 9, 10, 12 or 13) of every red-black pass of this launch set
    int ahead;        // fused Y+Q pass: planes its loads run ahead (2 or 3)
    int staged;       // 2: Y and Q fused (no Y vector); 1: Q pass with the Y planes staged in LDS; 0: six
                      // Y loads per voxel
    int fold_alpha;   // 1: the U pass reduces the Q pass's partials into alpha itself (one-volume
                      // launch sets); 0: a finalize launch between Q and U does
    const RbDesc* vol;
    const int* tile_vol;   // volume of each launched tile
    const int* tile_id;    // its index in the batch (tile_base + position); null: the launch index
    const Coef<double>* coef;
    VolState<double>* st;
    double* part;     // ntiles * 8: red-voxel partial sums
    double* part_b;   // ntiles * 8: black-voxel partial sums
    double* tab;      // per volume 2 * kTabEntries: diagonals, inverses
    float* x;         // red: FP32 correction since the last reliable update
    float* r;         // red: residual
    float* p[2];      // red: search direction (ping-pong by VolState::phase)
    float* q;         // red: S p
    double* y;        // black: D_b^-1 A_br p (FP64: S p cancels strongly)
    double* x64;      // red: FP64 solution
    double* x64b;     // black: FP64 solution recovered from x64
    double* xfull;    // full grid: FP64 solution (for the flux integral)
};

constexpr int kRbTX = 64;   // threads per block along k (one wave per row of a tile)

// Row pitch of the padded per-colour layout: at least hx + 2 (a ghost slot on either side of
// a row) and a multiple of 4, so every plane pitch is too and a 4-byte word of a code array
// holds the same four slots of a row in every plane.
inline int rb_row_pitch(int hx) { return (hx + 2 + 3) & ~3; }

// Diagonal tables and per-colour codes (after the vectors and codes were zero-filled).
hipError_t rb_setup(const RbLaunch&, hipStream_t);
// One CG iteration on S for the volumes in ST_RUNNING (4 passes, 2 per-volume finalizes).
hipError_t rb_iter(const RbLaunch&, hipStream_t);
// Reliable update / start for the volumes in ST_UPDATE or ST_FRESH: x64 += x, x = 0, then the
// FP64 residual r = b_S - S x64 (x64 = x0 on the first step) and the FP64 convergence test.
hipError_t rb_update(const RbLaunch&, hipStream_t);
// Streaming kernels for the measured HBM ceiling: kind 0 c = a + s b, 1 b = a, 2 a = s, over
// n4 float4 per buffer; blocks = 0 launches one float4 per thread, else a striding grid.
hipError_t stream_kernel(int kind, float4* a, float4* b, float4* c, size_t n4, unsigned blocks, hipStream_t s);
// One pass of rb_iter alone (0 = P, 1 = Y, 2 = Q, 3 = U), for timing.
hipError_t rb_pass_only(const RbLaunch&, int k, hipStream_t);
// Both colours scattered into the full-grid solution (the last update left x64b = x_b(x64)).
hipError_t rb_finish(const RbLaunch&, hipStream_t);

// Flux integral of the concentration `conc` (element offsets VolDesc::off) over a list of
// planes: entry i is plane plane_k[i] (along the primary axis) of volume plane_vol[i];
// plane_out[i] = (sum over that plane of D (c - c_next) / dx) / n, in T.
template <class T>
hipError_t flux_planes(int count, const VolDesc* vol, const int* plane_vol, const int* plane_k, const Coef<T>* coef,
                       const T* conc, double* plane_out, hipStream_t);

}  // namespace ft
