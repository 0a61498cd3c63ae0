// Internal interface between the host driver (tau_api.cpp) and the gfx950 kernels
// (tau_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace ft {

// Per-volume solve status kept on the device.
enum : int { ST_RUNNING = 0, ST_DONE = 1, ST_RESTART = 2, ST_IDLE = 3 };

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
    int plane_base, pad;                // flux: first plane slot of this volume
    long long off;                      // element offset in the concatenated vectors
    const uint8_t* code;                // per-voxel stencil code (bits 0-5 nbr solid, bit 6 own solid)
};

// Device-resident scalars of one solve (A = accumulation type).
template <class A>
struct VolState {
    A rho, rho_old, alpha, omega, beta;
    A r0sq, bsq, tol2, rsq, absnew;
    A tol, thresh, eps2;
    A tol_floor;                   // mixed: smallest inner relative tolerance asked for
    // mixed refinement (FP64 outer loop)
    double r64sq, b64sq, otol;     // last FP64 |r|^2, |b|^2, outer relative tolerance
    long long maxiter;             // Krylov iteration cap (inner cap in mixed mode)
    long long total_iter;          // mixed: inner iterations summed over refinement cycles
    int iter, status, restarts, first;
    int phase;                     // ping-pong parity counter (never reset)
    int zero_rhs;
    int mode;                      // 0 = plain solve, 1 = inner solve of mixed refinement
    int refinements, max_refine;
    int pad0;
};

// Vectors of one (batched) solve, concatenated over its volumes (element offset VolDesc::off).
struct VecSet {
    void* x;
    void* r;
    void* r0;
    void* p[2];
    void* v[2];
    void* t;
    double* x64;   // mixed refinement: FP64 solution
};

// Launch description shared by all passes of one batched solve.
template <class T, class A>
struct Launch {
    int tx, ty;               // block shape: one of the compiled shapes
    int ntiles, nvol;
    const VolDesc* vol;       // device
    const int* tile_vol;      // device: tile -> volume
    const Coef<T>* coef;      // device: per volume
    VolState<A>* st;          // device: per volume
    double* part;             // device: ntiles * 8 partial sums ((hi, lo) pairs)
    VecSet vs;
};

// ---- launchers (all asynchronous on the given stream) ----
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
    // One BiCGSTAB iteration (3 fused passes + 3 per-volume scalar updates).
    static hipError_t bicg_iter(const Launch<T, A>&, hipStream_t);
    // Same, recording ev[2k] / ev[2k+1] around pass k (k = 0, 1, 2) for timing.
    static hipError_t bicg_iter_timed(const Launch<T, A>&, hipStream_t, hipEvent_t* ev);
    // One PCG iteration (2 fused passes + 2 scalar updates).
    static hipError_t cg_iter(const Launch<T, A>&, hipStream_t);
    // y = A x (diagnostic; same device code as the solver).
    static hipError_t apply(const Launch<T, A>&, const T* x, T* y, hipStream_t);
};

// Mixed precision.  The inner system is FP32 (T = float, A = double); the outer residual
// is FP64 on x64 with the FP64 coefficients coef64 (per volume).  For every volume in
// ST_DONE: on its first refinement step x64 = x0; then r64 = b - A x64, the outer test, and
// (unless converged) a fresh inner BiCGSTAB on d with r = r0 = float(r64), d = p = v = 0.
hipError_t mixed_residual(const Launch<float, double>&, const Coef<double>* coef64, hipStream_t);
// x64 += d for volumes whose inner solve finished (status ST_DONE).
hipError_t mixed_update(const Launch<float, double>&, hipStream_t);

// Flux integral of the concentration `conc` (element offsets VolDesc::off):
// plane_out[plane_base + k] = (sum over plane k of D (c - c_next) / dx) / n, in T.
template <class T>
hipError_t flux_planes(int total_planes, const VolDesc* vol, const int* plane_vol, const Coef<T>* coef,
                       const T* conc, double* plane_out, hipStream_t);

}  // namespace ft
