/*
 * fast_tau.h -- C ABI of libfasttau, the MI355X-native tortuosity (`fast tau`) solver and
 * reactive area density (`fast alpha`).
 *
 * The reference exposes this path as a C++ template ABI of libfastlib.so:
 *   template<T> T fast::getTortuosity(const Volume&, const TortuosityParams&,
 *                                     DiffusionSolverType, Volume* conc, size_t* iters)
 *       fastlib/include/fastlib/volume/VolumeMeasures.h:51-58,
 *       fastlib/src/volume/VolumeMeasures.cpp:42-285
 *   template<T> T fast::getPorosity(const Volume&)       VolumeMeasures.h:60-61
 *   fast::getVolumeCCL / generateBoundaryConnectedVolume  VolumeSegmentation.h:33-47
 *   template<T> T fast::getReactiveAreaDensity(const Volume&, ivec3, float),
 *   template<T> std::array<T,6> fast::getReactiveAreaDensityTensor(const VolumeCCL&, ivec3, float)
 *                                                          VolumeMeasures.h:64-77
 *   fast::loadVolumeBinary / saveVolumeBinary              VolumeIO.h:15, VolumeIO.cpp:256-305
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions: return 0 on success, a negative FAST_E_* code on error; nothing
 * calls exit(); no global mutable state (all device state lives in a fast_ctx);
 * all caller buffers are host memory unless the name says _device.  One fast_ctx
 * may be used by one host thread at a time; different contexts are independent.
 * Volumes are uint8, x fastest then y then z (linearIndex, utility/Types.h:26-32);
 * 0 = pore, any nonzero = solid (Diffusion.cpp:94-96).
 */
#ifndef FAST_TAU_H
#define FAST_TAU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
#define FAST_OK 0
#define FAST_E_BADARG (-1)   /* invalid argument / unsupported value */
#define FAST_E_OOM (-2)      /* device memory (or HBM budget) exhausted */
#define FAST_E_NOCONV (-3)   /* solver did not reach the tolerance (tau reported as 0) */
#define FAST_E_HIP (-4)      /* HIP runtime error */
#define FAST_E_NODEV (-5)    /* no usable GPU */
#define FAST_E_IO (-6)       /* file I/O error */
#define FAST_E_BUSY (-7)     /* server queue full / request not finished */

/* ---- enums mirroring the reference ---------------------------------------- */
/* Dir, fastlib/include/fastlib/utility/PrimitiveTypes.h:142-150 */
enum fast_dir { FAST_X_POS = 0, FAST_X_NEG = 1, FAST_Y_POS = 2, FAST_Y_NEG = 3, FAST_Z_POS = 4, FAST_Z_NEG = 5 };
/* DiffusionSolverType, VolumeMeasures.h:15-18 */
enum fast_solver { FAST_BICGSTAB = 0, FAST_CG = 1 };
/* Precision: FP64 and FP32 are the reference's `-p double|float` (TauModule.h:19,125-129);
 * MIXED is this library's form: FP32 vectors (BiCGSTAB's directions p and v in FP64), FP64
 * arithmetic, dot products and scalars, the solution accumulated in FP64 and the residual
 * periodically replaced by the FP64 true residual (iterative refinement inside one Krylov
 * process); CG runs on the red-black Schur complement unless mixed_full_grid is set. */
enum fast_precision { FAST_FP64 = 0, FAST_FP32 = 1, FAST_MIXED = 2 };
/* Solve status (Solver<T>::Status, fastlib/include/fastlib/solver/Solver.h:47-51, extended) */
enum fast_status { FAST_SOLVED = 0, FAST_NO_CONVERGENCE = 1, FAST_DIVERGED = 2, FAST_FAILED = 3 };

/* TortuosityParams (VolumeMeasures.h:20-34) + SolveParams (Solver.h:39-43). */
typedef struct fast_tau_params {
    int dir;              /* fast_dir */
    int solver;           /* fast_solver */
    int precision;        /* fast_precision */
    double tolerance;     /* relative residual, reference default 1e-6 (TauModule.h:16,137) */
    long max_iter;        /* reference default 10000 (TauModule.h:17) */
    double d_pore;        /* coeffs[0]; the reference CLI passes float 1.0f (TauModule.h:70) */
    double d_solid;       /* coeffs[1]; the reference CLI passes float 0.001f */
    double porosity;      /* < 0: computed on the device as float(#zero/N) (TauModule.h:62) */
    double mixed_inner_tol; /* MIXED only: residual reduction between FP64 reliable updates (0 = 0.1) */
    double mixed_outer_tol; /* MIXED only: FP64 true relative residual target (0 = `tolerance`) */
    int verbose;
    int mixed_full_grid;  /* MIXED + CG only: 0 (default) = CG on the red-black Schur complement
                             (about half the iterations); 1 = CG on the full grid */
    int flux_nonuniform;  /* TortuosityParams::useNonUniform (VolumeMeasures.h:27): the flux sums
                             pore voxels only, with half spacing towards a solid neighbour, and a
                             plane's mean divides by its pore count (VolumeMeasures.cpp:228-264) */
    int reserved[1];
} fast_tau_params;

/* Output of one solve: getTortuosity's return value and iterationsOut, plus status. */
typedef struct fast_tau_result {
    double tau;           /* 0 on failure / non-convergence, like the reference */
    double porosity;      /* the porosity used in tau = porosity / (dc * N_primary) */
    double dc;            /* mean boundary flux (VolumeMeasures.cpp:274) */
    double rel_residual;  /* solver error: sqrt(|r|^2/|b|^2) of the recursive residual */
    double true_rel_residual; /* MIXED: FP64 true residual at exit; otherwise -1 */
    long iterations;      /* Krylov iterations */
    int status;           /* fast_status */
    int restarts;         /* BiCGSTAB restarts (rho breakdown; MIXED: fresh starts) */
    int refinements;      /* MIXED: FP64 true-residual replacements (reliable updates) */
    int reserved;
    double solve_ms;      /* device time of the solve (events on the solve stream) */
} fast_tau_result;

/* Fill `p` with the reference CLI defaults: X_POS, CG, FP64, tol 1e-6, 10000 iterations,
 * D = {1.0f, 0.001f}, porosity computed. */
void fast_tau_params_default(fast_tau_params* p);

/* ---- context --------------------------------------------------------------- */
typedef struct fast_ctx fast_ctx;

/* Create a context on HIP device `device`.  `hbm_budget` (bytes) bounds the device memory
 * the context may hold (0 = 90% of free memory).  Replaces the reference's implicit global
 * CUDA state (Diffusion.cu:8 __constant__ params, default stream).
 * Environment: FAST_TAU_GRAPHS=0 launches every polling round directly instead of replaying
 * a captured HIP graph (same kernels, same results); FAST_TAU_TUNING names a tile-shape table;
 * FAST_TAU_MIXED_RESTART sets the mixed BiCGSTAB's shadow-residual reset threshold (a cosine,
 * default 0: only Eigen's rule resets it); FAST_TAU_MIXED_BLOWUP g drops a mixed BiCGSTAB
 * correction whose residual grew g times since the last reliable update (default: only a
 * non-finite residual does);
 * FAST_TAU_GROUP_MB (default 320) is the red-black state size of the consecutive sub-batches a
 * batch of large volumes runs in (0 = always one launch; results are the same either way);
 * FAST_TAU_UPDATE_EVERY (default 4) is the number of mixed-form iterations between launches of
 * the reliable-update step (results are the same for every value); FAST_TAU_RB_TZ forces the
 * z depth of the red-black tiles (default: nz * tiles per plane / 2048, clamped to 4..32);
 * FAST_TAU_RB_STAGED=0 selects the red-black Q pass that reads its Y neighbours from memory
 * instead of staging them in LDS (results are bit-identical). */
int fast_ctx_create(int device, size_t hbm_budget, fast_ctx** out);
void fast_ctx_destroy(fast_ctx* ctx);
/* Last error message of this context (never NULL). */
const char* fast_ctx_last_error(const fast_ctx* ctx);
/* Bytes of device memory currently held by the context. */
size_t fast_ctx_device_bytes(const fast_ctx* ctx);
/* Force the tile shape of subsequent solves: a block of tx*ty threads owns a tx*ty column
 * brick and marches tz planes (tx,ty one of 64x8, 64x4, 64x16, 32x8, 32x16, 32x32, 16x16;
 * tz 1..256).
 * tx = 0 restores the automatic choice (the autotuned table, else a size rule).  Used by
 * the autotuner (fast_amd/autotune.py).  No reference counterpart. */
int fast_ctx_set_tile_shape(fast_ctx* ctx, int tx, int ty, int tz);
/* Force the tile shape of subsequent red-black (mixed CG) solves: ty rows of 64 slots per
 * tile (2, 4, 8, 9, 10, 12 or 13), tz planes per tile (1..256; 0 = the size rule) and the fused Y+Q pass's
 * prefetch distance in planes (2 or 3).  ty = 0 restores the automatic choice (the autotuned
 * "rb_" table entries, else 4 rows, the size rule and 2).  Used by the autotuner
 * (fast_amd/autotune.py).  No reference counterpart. */
int fast_ctx_set_rb_shape(fast_ctx* ctx, int ty, int tz, int ahead);
/* The red-black tile shape a solve of an nx x ny x nz volume would run with (forced, tuned or
 * automatic): shape[0] = rows per tile, shape[1] = planes per tile, shape[2] = prefetch.
 * No reference counterpart. */
int fast_ctx_rb_shape(const fast_ctx* ctx, int nx, int ny, int nz, int* shape);
/* Polling-round length in Krylov iterations (captured as one HIP graph); 0 = default 16. */
int fast_ctx_set_round(fast_ctx* ctx, int iterations);

/* ---- device-resident volumes ---------------------------------------------- */
/* A mask uploaded once and kept in HBM with its per-voxel stencil code; several solves
 * (directions, precisions, solvers) reuse it.  Replaces Volume + Texture3DPtr::commit
 * (Volume.cpp:19-43, DataPtr.cpp:209-242,377-407). */
typedef struct fast_volume fast_volume;

int fast_volume_create(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz, fast_volume** out);
/* Same, from a mask already in device memory (not retained: it is copied). */
int fast_volume_create_device(fast_ctx* ctx, const uint8_t* d_mask, int nx, int ny, int nz, fast_volume** out);
void fast_volume_destroy(fast_volume* vol);
/* #(mask == 0) on the device: Volume::sumZeroElems (Volume.cpp:261-272). */
int fast_volume_count_zero(fast_volume* vol, uint64_t* count);

/* ---- the drop-in ---------------------------------------------------------- */
/* getPorosity<double> (VolumeMeasures.cpp:288-292) computed on the device. */
int fast_porosity(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz, double* porosity);

/* getTortuosity<T> (VolumeMeasures.cpp:42-285): assemble (matrix-free), solve, flux, tau.
 * Volumes of up to 2^29 - 1 voxels (e.g. 812^3) are solved; larger ones return FAST_E_BADARG.
 * conc_out (nullable, N doubles) receives the concentration field (the reference's
 * concetrationOutput).  Returns FAST_OK, FAST_E_NOCONV (result->tau = 0) or an error. */
int fast_tau_solve(fast_ctx* ctx, const uint8_t* mask, int nx, int ny, int nz,
                   const fast_tau_params* params, fast_tau_result* result, double* conc_out);
/* Same on a device-resident volume (no host transfer of the mask). */
int fast_tau_solve_volume(fast_volume* vol, const fast_tau_params* params, fast_tau_result* result,
                          double* conc_out);

/* Many solves in one persistent batched solve (uneven sizes allowed): every volume shares
 * the same kernels, launched once per pass over the concatenated tile list and replayed
 * from a HIP graph.  Per-volume results are bit-identical to solving the volume alone.
 * params[i] applies to vols[i].  No reference counterpart: the reference solves volumes
 * one by one (fast/src/SliceModule.h:126-133). */
int fast_tau_solve_batch(fast_ctx* ctx, int count, fast_volume* const* vols,
                         const fast_tau_params* params, fast_tau_result* results);

/* ---- reachable pores (integer, bit-exact) ---------------------------------- */
/* 6-connected labels of the reference's span set (background value `background`, 255 in
 * the reference callers), numbered 1..K by each component's minimum linear index, 0 for
 * background: VolumeCCL_Label (VolumeCCL.cu:381-620) via getVolumeCCL (VolumeSegmentation.cpp:33-64).
 * labels_out: N uint32 (nullable); num_labels = K + 1.  Labels are 32-bit voxel indices, which
 * every volume fits (fast_volume_create takes fewer than 2^31 voxels). */
int fast_ccl_labels(fast_volume* vol, uint8_t background, uint32_t* labels_out, uint32_t* num_labels);
/* Reachable-pore mask for the axis of `dir`: 255 where the component touches both faces of
 * that axis, else 0 (VolumeCCL::getDirMask(begin,end), VolumeSegmentation.cpp:14-24, and
 * VolumeCCL_GenerateVolume, VolumeCCL.cu:821-860).  out: N bytes. */
int fast_reachable_mask(fast_volume* vol, int dir, uint8_t background, uint8_t* out, uint32_t* num_labels);
/* Per-face label table (VolumeCCL_BoundaryLabels, VolumeCCL.cu:753-818): 7 rows of
 * num_labels bytes (rows in Dir order, row 6 = present on all faces).  table: 7*num_labels. */
int fast_ccl_boundary_table(fast_volume* vol, uint8_t background, uint8_t* table, uint32_t table_labels);
/* Per-label segment statistics (getCCLSegmentInfo, VolumeSegmentation.cpp:135-178) of the
 * labels for `background`: out[l] for every label l (0, the background, included) holds the
 * voxel count, the bounding box (min inclusive, max exclusive: the reference's maxBB + 1) and
 * whether the box touches a face of the volume.  out: num_labels entries (num_labels from
 * fast_ccl_labels; a different count is refused). */
typedef struct fast_segment_info {
    int32_t min_bb[3];    /* x, y, z; the volume's dims for a label with no voxels */
    int32_t max_bb[3];    /* exclusive */
    int64_t voxel_num;
    int32_t at_boundary;  /* max_bb == dim or min_bb == 0 on some axis */
    int32_t label_id;
} fast_segment_info;
int fast_ccl_segment_info(fast_volume* vol, uint8_t background, fast_segment_info* out, uint32_t num_labels);

/* ---- reactive area density (`fast alpha`) ---------------------------------- */
/* Marching-cubes area of the `isovalue` level set (0.5 in the reference callers) of the
 * volume read as its linearly filtered texture: each voxel corner takes the mean of the 8
 * voxels around it (indices clamped into the volume), byte/255, and a cube corner below the
 * isovalue is inside.  Coordinates are normalised per axis, so the volume spans the unit
 * cube and the area is the area density in units of 1/L.  Triangles are Bourke's table;
 * per-cube areas in the reference's float arithmetic, summed in double (one rounding).
 * getReactiveAreaDensity<double>(mask, dim, 0.5f) (VolumeMeasures.cpp:294-300) ->
 * getVolumeArea (VolumeSurface.cpp:9-36) -> VolumeSurface_MarchingCubesArea
 * (VolumeSurface.cu:203-258, 507-535) and Volume::sum. */
int fast_alpha_area(fast_volume* vol, float isovalue, double* area);
/* fast_alpha_area without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The kernels' block partials and result live in a static 1 MiB
 * device array (one per device, set up when the context is created, held until the process
 * ends), so volumes needing more than 65,535 blocks of 64x4x16 voxels return FAST_E_BADARG
 * (512^3 takes 32,768 blocks and fits; 512x512x1024 takes 65,536 and does not).  Same kernels and the same result as fast_alpha_area, bit for bit.
 * Calls that overlap on one device race on the shared array (one context serialises its own
 * calls). */
int fast_alpha_area_noalloc(fast_volume* vol, float isovalue, double* area);
/* The six per-face areas of getReactiveAreaDensityTensor<double>(getVolumeCCL(vol, 255))
 * (VolumeMeasures.cpp:303-333): for face f (Dir order x+, x-, y+, y-, z+, z-), the area
 * above on the volume that is 255 on the pore voxels whose component touches face f and 0
 * elsewhere (generateBoundaryConnectedVolume, VolumeSegmentation.cpp:66-90).  area: 6. */
int fast_alpha_area_tensor(fast_volume* vol, float isovalue, double area[6]);
/* Triangle mesh of the volume padded with one zero voxel per side, as `fast alpha --mesh`
 * builds it (AlphaModule.h:145-167: withZeroPadding(1, 1), getVolumeAreaMesh(padded, false);
 * VolumeSurface.cpp:39-74, VolumeSurface.cu:260-505): marching cubes over the padded grid,
 * positions in a box centred at the origin with spacing 1/maxDim of the padded dims, three
 * vertices per triangle in cube order then table order, each carrying its triangle's unit
 * normal.  *nverts = number of vertices; when verts is non-null and capacity (in vertices) is
 * at least that, verts receives 6 floats per vertex (x, y, z, nx, ny, nz).  Meshes of 2^32
 * vertices or more (the reference's unsigned vertex count) and padded volumes of 2^31 cubes or
 * more are refused. */
int fast_alpha_mesh(fast_volume* vol, float isovalue, float* verts, size_t capacity, size_t* nverts);
/* The same mesh written as the reference's OBJ (CUDA_VBO::saveObj, DataPtr.cpp:684-740):
 * vertices merged by equal position in order of first appearance, normals summed and
 * normalised, `v` / `vn` / `f` lines.  No file is written for an empty mesh (*nverts = 0). */
int fast_alpha_save_obj(fast_volume* vol, float isovalue, const char* path, size_t* nverts);

/* ---- serving: continuous batching ------------------------------------------ */
/* A request loop that admits tau solves of mixed sizes by HBM budget, keeps masks and
 * Krylov state resident, lets requests join the running batch at iteration boundaries
 * and leave on convergence.  No reference counterpart. */
typedef struct fast_server fast_server;
typedef struct fast_server_config {
    size_t hbm_budget;      /* bytes of the Krylov-state arena (0 = half the context budget, at
                             most 16 GiB); a request takes 46 B per element of its padded
                             per-colour layout, (ceil(nx/2)+2)(ny+2)(nz+2) elements (about
                             23 B per voxel), with CG; 56 B per voxel with BiCGSTAB */
    int max_active;         /* max concurrently solving requests (0 = 64) */
    int iters_per_round;    /* Krylov iterations between admission points (0 = 8) */
    int use_bicgstab;       /* 0: CG (default), 1: BiCGSTAB */
    int reserved[3];
} fast_server_config;

int fast_server_create(fast_ctx* ctx, const fast_server_config* cfg, fast_server** out);
/* Requests are solved in the mixed form (FP32 vectors, FP64 arithmetic, reliable updates) to
 * params->tolerance on the FP64 true residual: CG on the red-black Schur complement, or
 * BiCGSTAB on the full grid (use_bicgstab); params->solver, ->precision and ->mixed_full_grid
 * are ignored.  A request's result is bit-identical to fast_tau_solve_volume on the same volume
 * with precision FAST_MIXED and the server's solver (mixed_full_grid = 0), whatever else is in
 * flight.  The server owns its context while it runs: do not use the context from other
 * threads meanwhile. */
/* Copy the mask and queue a request; returns a ticket.  Never blocks on the GPU. */
int fast_server_submit(fast_server* srv, const uint8_t* mask, int nx, int ny, int nz,
                       const fast_tau_params* params, uint64_t* ticket);
/* Wait up to timeout_ms (negative = forever) for any finished request; returns FAST_E_BUSY
 * on timeout.  latency_ms = completion - submission (host monotonic clock). */
int fast_server_wait(fast_server* srv, int timeout_ms, uint64_t* ticket, fast_tau_result* result,
                     double* latency_ms);
/* Requests submitted and not yet returned by fast_server_wait. */
int fast_server_pending(fast_server* srv);
void fast_server_destroy(fast_server* srv);

/* ---- input generation and .bin I/O (host only) ------------------------------- */
/* Reference .bin format (VolumeIO.cpp:256-305): int32 PrimitiveType (UCHAR = 2), int32[3]
 * dims, 1 byte flag, raw voxels. */
int fast_load_bin_dims(const char* path, int* nx, int* ny, int* nz);
int fast_load_bin(const char* path, uint8_t* mask, size_t capacity);
int fast_save_bin(const char* path, const uint8_t* mask, int nx, int ny, int nz);
/* TIFF folder (loadTiffFolder, fastlib/src/volume/VolumeIO.cpp:195-253): every file whose name
 * ends in ".tiff" or ".tif" (case-sensitive, VolumeIO.cpp:307-310), in sorted name order
 * (listDir, VolumeIO.cpp:79-111), is one z slice; nx, ny come from the first.  Baseline
 * uncompressed 8-bit grayscale strips in either byte order, the first image of each file.
 * FAST_E_IO for anything else (compression, other sample sizes, tiles, BigTIFF) and for slices
 * of another size; the reference's reader copies other sample sizes into the uint8 slice
 * unchecked. */
int fast_tiff_folder_dims(const char* folder, int* nx, int* ny, int* nz);
int fast_load_tiff_folder(const char* folder, uint8_t* mask, size_t capacity);
/* Volume::binarize on uint8 (fastlib/src/cuda/Volume.cu:152-182): t = uchar(threshold * 255),
 * v = v < t ? 0 : 255.  The CLI binarizes TIFF input with threshold 1 (VolumeIOModule.h:55-60). */
void fast_binarize(uint8_t* mask, size_t n, float threshold);
/* `.sph` input (VolumeIOModule.h:86-96): loadSpheres (VolumeGenerator.cpp:273-286: a count,
 * then x y z r per sphere in normalized coordinates) rasterized at nx x ny x nz by
 * rasterizeSpheres (VolumeGenerator.cpp:210-251): 255 where float(i)/dim lies strictly inside a
 * sphere (float arithmetic), scanning int((dim-1)(c -+ r)) clamped to the box. */
int fast_rasterize_sph(const char* path, int nx, int ny, int nz, uint8_t* mask);
/* Subvolume (Volume::getSubvolume, fastlib/src/volume/Volume.cpp:129-160, used by the CLI's
 * --sub/--origin/--randsub/--aspect crops, VolumeIOModule.h:159-245): dst (sx*sy*sz) = the box
 * [o, o + s) of src. */
int fast_crop(const uint8_t* src, int nx, int ny, int nz, int ox, int oy, int oz, int sx, int sy, int sz,
              uint8_t* dst);

/* Seeded generators (255 = solid, 0 = pore): bodies are added with centres uniform in the
 * box, overlap allowed, until the solid fraction reaches `solid_fraction`; a voxel is solid
 * when its centre (i + 1/2) lies inside a body.  Ellipsoid semi-axes are (a, a*r, a*r) with
 * a uniformly random orientation (unit quaternion).  Returns the number of bodies placed. */
long fast_generate_spheres(uint8_t* mask, int nx, int ny, int nz, double radius,
                           double solid_fraction, uint64_t seed);
long fast_generate_ellipsoids(uint8_t* mask, int nx, int ny, int nz, double a, double r,
                              double solid_fraction, uint64_t seed);

/* ---- diagnostics (test hooks on the same kernels the solver uses) ---------------- */
/* y = A x with the matrix-free operator for direction `dir` (FP64). */
int fast_debug_apply_f64(fast_volume* vol, int dir, const double* x, double* y);
/* y = A x in FP32 (the operator of `-p float` and of the MIXED inner solve). */
int fast_debug_apply_f32(fast_volume* vol, int dir, const float* x, float* y);
/* Right-hand side b, initial guess x0 and Jacobi inverse 1/diag(A) (FP64). */
int fast_debug_system_f64(fast_volume* vol, int dir, double* b, double* x0, double* invdiag);
/* Device times (ms, HIP events on the solve stream) on a solve set up on `vol` (precision
 * FAST_FP64, FAST_FP32, or FAST_MIXED for the FP32-storage / FP64-arithmetic passes): the mean
 * duration of each stencil pass kernel, launched `iters` times back to back, and the mean
 * period of a whole iteration (passes and per-volume finalizes) over `iters` iterations.
 * ms: 4 doubles; BiCGSTAB {pass1, pass2, pass3, iteration}, CG {pass1, pass2, 0, iteration}. */
int fast_debug_time_passes(fast_volume* vol, int dir, int precision, int solver, int iters, double* ms);
/* The same for mixed CG on the red-black Schur complement: ms[0..3] = mean duration of the
 * passes P (p = D^-1 r + beta p), Y (black: D^-1 A p), Q (red: q = S p, p.q) and U (x, r
 * update, r.r, r.D^-1 r), each launched `iters` times back to back; ms[4] = mean period of a
 * whole iteration.  ms: 5 doubles. */
int fast_debug_time_rb_passes(fast_volume* vol, int dir, int iters, double* ms);
/* Measured HBM ceiling: streaming kernels over three device buffers of `bytes` each (16-byte
 * accesses), each launched `iters` times back to back after one untimed launch (HIP events):
 * ms[0] c = a + s b with one float4 per thread, ms[1] the same on 8192 striding blocks,
 * ms[2] b = a, ms[3] a = s.  Bytes moved per launch: 3, 3, 2 and 1 times `bytes`.  No
 * reference counterpart.  ms: 4 doubles. */
int fast_debug_stream(fast_ctx* ctx, size_t bytes, int iters, double* ms);
/* `fast alpha` kernels, each run `iters` times back to back on the context's stream (HIP
 * events) after one untimed run: ms[0] = whole-volume area (area pass + final sum), ms[1] = the six per-face areas
 * from the cached labels (face codes + area pass + final sums), ms[2] = the six per-face
 * areas from scratch (CCL labels, face table, face codes, area).  ms: 3 doubles. */
int fast_debug_time_alpha(fast_volume* vol, int iters, double* ms);

#ifdef __cplusplus
}
#endif
#endif
