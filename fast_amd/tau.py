"""Host-side mirror of the reference's tortuosity interface over libfasttau's C ABI.

Names, argument meaning and error behaviour follow the reference:
  - Dir                     fastlib/include/fastlib/utility/PrimitiveTypes.h:142-150
  - DiffusionSolverType     fastlib/include/fastlib/volume/VolumeMeasures.h:15-18
  - TortuosityParams        VolumeMeasures.h:20-34
  - getTortuosity(...)      VolumeMeasures.h:51-58, fastlib/src/volume/VolumeMeasures.cpp:42-285
                            (returns 0 on any failure or non-convergence)
  - getPorosity(...)        VolumeMeasures.h:60-61, VolumeMeasures.cpp:288-292
  - getDirs(...)            fast/src/utility.h:84-113
  - load/save .bin          fastlib/src/volume/VolumeIO.cpp:256-305
Every compute call runs the HIP kernels of libfasttau; there is no CPU solver here.
"""
from __future__ import annotations

import ctypes as C
import enum
import weakref
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from ._lib import (FAST_E_BUSY, FAST_E_NOCONV, FAST_OK, FastError, ServerConfig, TauParams, TauResult, check, lib)


class Dir(enum.IntEnum):
    X_POS = 0
    X_NEG = 1
    Y_POS = 2
    Y_NEG = 3
    Z_POS = 4
    Z_NEG = 5


class DiffusionSolverType(enum.IntEnum):
    DSOLVER_BICGSTAB = 0
    DSOLVER_CG = 1


class Precision(enum.IntEnum):
    FP64 = 0
    FP32 = 1
    MIXED = 2


def _as_precision(p) -> int:
    if isinstance(p, str):
        p = p.lower()
        if p in ("double", "d", "fp64", "f64"):
            return int(Precision.FP64)
        if p in ("float", "single", "f", "s", "fp32", "f32"):
            return int(Precision.FP32)
        if p in ("mixed", "m"):
            return int(Precision.MIXED)
        raise ValueError(f"unknown precision {p!r}")
    return int(p)


@dataclass
class TortuosityParams:
    """TortuosityParams (VolumeMeasures.h:20-34).  coeffs are rounded to float like the
    reference's glm::vec2."""
    dir: Dir = Dir.X_NEG
    coeffs: tuple = (1.0, 0.001)
    tolerance: float = 1e-6
    maxIter: int = 10000
    verbose: bool = False
    oversubscribeGPUMemory: bool = False
    useNonUniform: bool = False
    porosityPrecomputed: bool = False
    porosity: float = -1.0
    onDevice: bool = True
    cpuThreads: int = -1
    # additions of this library
    mixed_inner_tol: float = 0.0
    mixed_outer_tol: float = 0.0


def getDirs(flag: str) -> list:
    """Direction list of a `-d` flag string (fast/src/utility.h:84-113)."""
    if flag == "all":
        return [Dir.X_POS, Dir.X_NEG, Dir.Y_POS, Dir.Y_NEG, Dir.Z_POS, Dir.Z_NEG]
    if flag == "pos":
        return [Dir.X_POS, Dir.Y_POS, Dir.Z_POS]
    if flag == "neg":
        return [Dir.X_NEG, Dir.Y_NEG, Dir.Z_NEG]
    if flag in ("xneg", "yneg", "zneg"):
        return [{"xneg": Dir.X_NEG, "yneg": Dir.Y_NEG, "zneg": Dir.Z_NEG}[flag]]
    out = [{"x": Dir.X_POS, "y": Dir.Y_POS, "z": Dir.Z_POS}[c] for c in flag if c in "xyz"]
    return out or [Dir.X_POS]


def _u8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _mask_xyz(mask: np.ndarray) -> tuple:
    """Volumes are numpy arrays indexed [z, y, x] (x fastest in memory, like the reference)."""
    if mask.ndim != 3:
        raise ValueError("mask must be a 3-D array indexed [z, y, x]")
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nz, ny, nx = m.shape
    return m, nx, ny, nz


class Context:
    """A device context (one HIP device, one stream, a device-memory pool)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        self._volumes = weakref.WeakSet()
        check(lib().fast_ctx_create(device, hbm_budget, C.byref(self._h)), "fast_ctx_create")

    @property
    def handle(self):
        return self._h

    def close(self):
        # volumes first: their handles point into this context
        for v in list(self._volumes):
            v.close()
        if self._h:
            lib().fast_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return lib().fast_ctx_last_error(self._h).decode(errors="replace")

    def device_bytes(self) -> int:
        return lib().fast_ctx_device_bytes(self._h)

    def set_rb_shape(self, ty: int = 0, tz: int = 0, ahead: int = 2):
        """Force the red-black tile shape of later mixed CG solves (rows per tile 2/4/8/9/10/12/13,
        planes per tile, 0 = size rule; fused-pass prefetch 2/3); ty = 0 restores the
        tuned/automatic choice.  Results do not depend on it."""
        check(lib().fast_ctx_set_rb_shape(self._h, ty, tz, ahead), "fast_ctx_set_rb_shape", self._h)

    def rb_shape(self, dims) -> tuple:
        """(rows per tile, planes per tile, prefetch) a red-black solve of a volume of `dims`
        (nx, ny, nz) runs with."""
        out = (C.c_int * 3)()
        check(lib().fast_ctx_rb_shape(self._h, int(dims[0]), int(dims[1]), int(dims[2]), out), "fast_ctx_rb_shape",
              self._h)
        return tuple(out)

    def volume(self, mask: np.ndarray) -> "Volume":
        return Volume(self, mask)


class Volume:
    """A mask resident in HBM (with its per-voxel stencil code)."""

    def __init__(self, ctx: Context, mask: np.ndarray):
        m, nx, ny, nz = _mask_xyz(mask)
        self.ctx = ctx
        self.shape = m.shape
        self.dims = (nx, ny, nz)
        self._h = C.c_void_p()
        check(lib().fast_volume_create(ctx.handle, _u8(m), nx, ny, nz, C.byref(self._h)), "fast_volume_create",
              ctx.handle)
        ctx._volumes.add(self)

    @property
    def handle(self):
        return self._h

    @property
    def n(self) -> int:
        return int(np.prod(self.dims))

    def close(self):
        if self._h:
            if self.ctx.handle:
                lib().fast_volume_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def count_zero(self) -> int:
        c = C.c_uint64()
        check(lib().fast_volume_count_zero(self._h, C.byref(c)), "fast_volume_count_zero")
        return int(c.value)

    def solve(self, params: TauParams, conc: bool = False):
        res = TauResult()
        out = np.empty(self.shape, dtype=np.float64) if conc else None
        rc = lib().fast_tau_solve_volume(self._h, C.byref(params), C.byref(res),
                                         out.ctypes.data_as(C.POINTER(C.c_double)) if conc else None)
        if rc not in (FAST_OK, FAST_E_NOCONV):
            check(rc, "fast_tau_solve_volume", self.ctx.handle)
        return res, out

    def ccl_labels(self, background: int = 255):
        lab = np.empty(self.shape, dtype=np.uint32)
        nl = C.c_uint32()
        check(lib().fast_ccl_labels(self._h, background, lab.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(nl)),
              "fast_ccl_labels", self.ctx.handle)
        return lab, int(nl.value)

    def boundary_table(self, background: int = 255) -> np.ndarray:
        _, nl = self.ccl_labels(background)
        t = np.empty((7, nl), dtype=np.uint8)
        check(lib().fast_ccl_boundary_table(self._h, background, _u8(t), nl), "fast_ccl_boundary_table",
              self.ctx.handle)
        return t

    def segment_info(self, background: int = 255) -> list:
        """Per-label statistics of the CCL labels (getCCLSegmentInfo,
        VolumeSegmentation.cpp:135-178): a dict per label (label 0 included) with minBB
        (x, y, z), maxBB (exclusive), voxelNum, atBoundary and labelID."""
        from ._lib import SegmentInfo
        _, nl = self.ccl_labels(background)
        arr = (SegmentInfo * nl)()
        check(lib().fast_ccl_segment_info(self._h, background, arr, nl), "fast_ccl_segment_info", self.ctx.handle)
        return [{"minBB": tuple(r.min_bb), "maxBB": tuple(r.max_bb), "voxelNum": int(r.voxel_num),
                 "atBoundary": bool(r.at_boundary), "labelID": int(r.label_id)} for r in arr]

    def reachable_mask(self, d: int, background: int = 255):
        out = np.empty(self.shape, dtype=np.uint8)
        nl = C.c_uint32()
        check(lib().fast_reachable_mask(self._h, int(d), background, _u8(out), C.byref(nl)), "fast_reachable_mask",
              self.ctx.handle)
        return out, int(nl.value)

    def area(self, isovalue: float = 0.5) -> float:
        """Marching-cubes area of the mask's isovalue level set in unit-cube coordinates
        (getVolumeArea + Volume::sum, VolumeSurface.cpp:9-36, VolumeMeasures.cpp:294-300)."""
        out = C.c_double()
        check(lib().fast_alpha_area(self._h, float(isovalue), C.byref(out)), "fast_alpha_area", self.ctx.handle)
        return float(out.value)

    def area_tensor(self, isovalue: float = 0.5) -> np.ndarray:
        """Per-face areas (Dir order) of the pore space connected to each face
        (getReactiveAreaDensityTensor, VolumeMeasures.cpp:303-333)."""
        out = np.empty(6, dtype=np.float64)
        check(lib().fast_alpha_area_tensor(self._h, float(isovalue), out.ctypes.data_as(C.POINTER(C.c_double))),
              "fast_alpha_area_tensor", self.ctx.handle)
        return out

    def mesh(self, isovalue: float = 0.5) -> np.ndarray:
        """Triangle mesh of the zero-padded volume as `fast alpha --mesh` builds it
        (getVolumeAreaMesh, VolumeSurface.cpp:39-74): (n_vertices, 6) float32, position and
        triangle normal, three vertices per triangle."""
        n = C.c_size_t()
        check(lib().fast_alpha_mesh(self._h, float(isovalue), None, 0, C.byref(n)), "fast_alpha_mesh", self.ctx.handle)
        out = np.empty((n.value, 6), dtype=np.float32)
        if n.value:
            check(lib().fast_alpha_mesh(self._h, float(isovalue), out.ctypes.data_as(C.POINTER(C.c_float)), n.value,
                                        C.byref(n)), "fast_alpha_mesh", self.ctx.handle)
        return out

    def save_obj(self, path: str, isovalue: float = 0.5) -> int:
        """The mesh as the reference's OBJ (CUDA_VBO::saveObj, DataPtr.cpp:684-740); returns the
        number of triangle vertices (0: empty mesh, no file written)."""
        n = C.c_size_t()
        check(lib().fast_alpha_save_obj(self._h, float(isovalue), str(path).encode(), C.byref(n)),
              "fast_alpha_save_obj", self.ctx.handle)
        return int(n.value)

    def apply(self, d: int, x: np.ndarray, precision="double") -> np.ndarray:
        """y = A x with the solver's matrix-free operator (diagnostic)."""
        if _as_precision(precision) == Precision.FP32:
            xx = np.ascontiguousarray(x, dtype=np.float32)
            y = np.empty_like(xx)
            check(lib().fast_debug_apply_f32(self._h, int(d), xx.ctypes.data_as(C.POINTER(C.c_float)),
                                             y.ctypes.data_as(C.POINTER(C.c_float))), "fast_debug_apply_f32",
                  self.ctx.handle)
            return y
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(xx)
        check(lib().fast_debug_apply_f64(self._h, int(d), xx.ctypes.data_as(C.POINTER(C.c_double)),
                                         y.ctypes.data_as(C.POINTER(C.c_double))), "fast_debug_apply_f64",
              self.ctx.handle)
        return y

    def system(self, d: int):
        """(b, x0, 1/diag) of the FP64 system for direction d (diagnostic)."""
        b = np.empty(self.shape, dtype=np.float64)
        x0 = np.empty_like(b)
        inv = np.empty_like(b)
        dp = C.POINTER(C.c_double)
        check(lib().fast_debug_system_f64(self._h, int(d), b.ctypes.data_as(dp), x0.ctypes.data_as(dp),
                                          inv.ctypes.data_as(dp)), "fast_debug_system_f64", self.ctx.handle)
        return b, x0, inv


def make_params(d=Dir.X_POS, solver=DiffusionSolverType.DSOLVER_BICGSTAB, precision="double", tol=1e-6,
                max_iter=10000, coeffs=(1.0, 0.001), porosity=-1.0, inner_tol=0.0, outer_tol=0.0,
                verbose=False, full_grid=False, nonuniform=False) -> TauParams:
    p = TauParams()
    lib().fast_tau_params_default(C.byref(p))
    p.dir = int(d)
    p.solver = int(solver)
    p.precision = _as_precision(precision)
    p.tolerance = float(tol)
    p.max_iter = int(max_iter)
    p.d_pore = float(coeffs[0])
    p.d_solid = float(coeffs[1])
    p.porosity = float(porosity)
    p.mixed_inner_tol = float(inner_tol)
    p.mixed_outer_tol = float(outer_tol)
    p.verbose = int(bool(verbose))
    p.mixed_full_grid = int(bool(full_grid))
    p.flux_nonuniform = int(bool(nonuniform))
    return p


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def getPorosity(mask: np.ndarray, ctx: Optional[Context] = None) -> float:
    """T(#zero) / T(N) computed on the device (VolumeMeasures.cpp:288-292)."""
    ctx = ctx or default_context()
    m, nx, ny, nz = _mask_xyz(mask)
    out = C.c_double()
    check(lib().fast_porosity(ctx.handle, _u8(m), nx, ny, nz, C.byref(out)), "fast_porosity", ctx.handle)
    return float(out.value)


def getTortuosity(mask, params: TortuosityParams, solverType=DiffusionSolverType.DSOLVER_BICGSTAB,
                  concetrationOutput: Optional[np.ndarray] = None, iterationsOut: Optional[list] = None,
                  precision="double", ctx: Optional[Context] = None) -> float:
    """getTortuosity<T> (VolumeMeasures.cpp:42-285).  `mask` is a [z, y, x] uint8 array or a
    resident Volume.  Returns 0.0 on failure or non-convergence, like the reference; the
    iteration count is appended to `iterationsOut` when given."""
    if not params.onDevice:
        raise NotImplementedError("this library has no CPU solver (onDevice=False)")
    ctx = ctx or (mask.ctx if isinstance(mask, Volume) else default_context())
    vol = mask if isinstance(mask, Volume) else Volume(ctx, mask)
    por = params.porosity if params.porosityPrecomputed else -1.0
    p = make_params(params.dir, solverType, precision, params.tolerance, params.maxIter, params.coeffs, por,
                    params.mixed_inner_tol, params.mixed_outer_tol, params.verbose, nonuniform=params.useNonUniform)
    try:
        res, conc = vol.solve(p, conc=concetrationOutput is not None)
    except FastError:
        return 0.0
    if concetrationOutput is not None:
        concetrationOutput[...] = conc
    if iterationsOut is not None:
        iterationsOut.append(int(res.iterations))
    return float(res.tau)


def getReactiveAreaDensity(mask, isovalue: float = 0.5, ctx: Optional[Context] = None) -> float:
    """getReactiveAreaDensity<double>(mask, mask.dim(), isovalue) (VolumeMeasures.cpp:294-300):
    the marching-cubes area of the mask (a [z, y, x] uint8 array or a resident Volume)."""
    ctx = ctx or (mask.ctx if isinstance(mask, Volume) else default_context())
    vol = mask if isinstance(mask, Volume) else Volume(ctx, mask)
    try:
        return vol.area(isovalue)
    finally:
        if vol is not mask:
            vol.close()


def getReactiveAreaDensityTensor(mask, isovalue: float = 0.5, ctx: Optional[Context] = None) -> np.ndarray:
    """getReactiveAreaDensityTensor<double>(getVolumeCCL(mask, 255)) (VolumeMeasures.cpp:303-333):
    six areas in Dir order, each of the pore space connected to that face."""
    ctx = ctx or (mask.ctx if isinstance(mask, Volume) else default_context())
    vol = mask if isinstance(mask, Volume) else Volume(ctx, mask)
    try:
        return vol.area_tensor(isovalue)
    finally:
        if vol is not mask:
            vol.close()


def filterSegmentInfo(sinfo: list, table: np.ndarray, d: int = 6, invertMask: bool = True) -> list:
    """filterSegmentInfo (VolumeSegmentation.cpp:182-205): the segments whose label is on row
    `d` of the per-face table (Dir order, 6 = on every face), or with `invertMask` the labels
    >= 1 that are not."""
    row = np.asarray(table)[d].astype(bool)
    keep = row.copy()
    if invertMask:
        keep[1:] = ~row[1:]
    return [s for s, k in zip(sinfo, keep) if k]


def getCCLSegmentRatio(sinfo: list, n_voxels: int) -> float:
    """getCCLSegmentRatio (VolumeSegmentation.cpp:207-215): float(sum of voxelNum) / N."""
    return float(np.float32(sum(s["voxelNum"] for s in sinfo)) / np.float32(n_voxels))


def solve_batch(ctx: Context, vols: Sequence[Volume], params: Sequence[TauParams]) -> list:
    """One persistent batched solve of many volumes (uneven sizes allowed)."""
    n = len(vols)
    arr_v = (C.c_void_p * n)(*[v.handle for v in vols])
    arr_p = (TauParams * n)(*params)
    arr_r = (TauResult * n)()
    rc = lib().fast_tau_solve_batch(ctx.handle, n, arr_v, arr_p, arr_r)
    if rc not in (FAST_OK, FAST_E_NOCONV):
        check(rc, "fast_tau_solve_batch", ctx.handle)
    return [arr_r[i] for i in range(n)]


class Server:
    """Continuous-batching tau solves: requests of any size join the running batch at round
    boundaries and leave when their FP64 true residual meets the tolerance (mixed form).
    The server owns `ctx` while it is open."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_active: int = 0, iters_per_round: int = 0,
                 use_bicgstab: bool = False):
        cfg = ServerConfig(hbm_budget=hbm_budget, max_active=max_active, iters_per_round=iters_per_round,
                           use_bicgstab=int(use_bicgstab))
        self.ctx = ctx
        self._h = C.c_void_p()
        check(lib().fast_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)), "fast_server_create", ctx.handle)

    def submit(self, mask: np.ndarray, params: TauParams) -> int:
        m, nx, ny, nz = _mask_xyz(mask)
        t = C.c_uint64()
        check(lib().fast_server_submit(self._h, _u8(m), nx, ny, nz, C.byref(params), C.byref(t)),
              "fast_server_submit", self.ctx.handle)
        return int(t.value)

    def wait(self, timeout_ms: int = -1):
        """(ticket, TauResult, latency_ms) of the next finished request, or None on timeout."""
        t, res, lat = C.c_uint64(), TauResult(), C.c_double()
        rc = lib().fast_server_wait(self._h, timeout_ms, C.byref(t), C.byref(res), C.byref(lat))
        if rc == FAST_E_BUSY:
            return None
        if rc not in (FAST_OK, FAST_E_NOCONV):
            check(rc, "fast_server_wait", self.ctx.handle)
        return int(t.value), res, float(lat.value)

    def pending(self) -> int:
        return int(lib().fast_server_pending(self._h))

    def close(self):
        if self._h:
            lib().fast_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- volumes: .bin I/O and generators (host) -----------------------------------------------
def load_bin(path: str) -> np.ndarray:
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    check(lib().fast_load_bin_dims(str(path).encode(), C.byref(nx), C.byref(ny), C.byref(nz)), "fast_load_bin_dims")
    m = np.empty((nz.value, ny.value, nx.value), dtype=np.uint8)
    check(lib().fast_load_bin(str(path).encode(), _u8(m), m.size), "fast_load_bin")
    return m


def save_bin(path: str, mask: np.ndarray) -> None:
    m, nx, ny, nz = _mask_xyz(mask)
    check(lib().fast_save_bin(str(path).encode(), _u8(m), nx, ny, nz), "fast_save_bin")


def loadTiffFolder(folder: str) -> np.ndarray:
    """loadTiffFolder (VolumeIO.cpp:195-253): the .tif/.tiff slices of `folder` in sorted name
    order as a [z, y, x] uint8 volume (raw values; the CLI binarizes them)."""
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    check(lib().fast_tiff_folder_dims(str(folder).encode(), C.byref(nx), C.byref(ny), C.byref(nz)),
          "fast_tiff_folder_dims")
    m = np.empty((nz.value, ny.value, nx.value), dtype=np.uint8)
    check(lib().fast_load_tiff_folder(str(folder).encode(), _u8(m), m.size), "fast_load_tiff_folder")
    return m


def binarize(mask: np.ndarray, threshold: float = 1.0) -> np.ndarray:
    """Volume::binarize on uint8 (Volume.cu:152-182): v < uchar(threshold * 255) -> 0, else 255."""
    m = np.ascontiguousarray(mask, dtype=np.uint8).copy()
    lib().fast_binarize(_u8(m), m.size, threshold)
    return m


def rasterizeSpheres(path: str, shape) -> np.ndarray:
    """A `.sph` sphere list (VolumeGenerator.cpp:273-286) rasterized at `shape` = (nx, ny, nz)
    (rasterizeSpheres, VolumeGenerator.cpp:210-251), as a [z, y, x] volume."""
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    m = np.empty((nz, ny, nx), dtype=np.uint8)
    check(lib().fast_rasterize_sph(str(path).encode(), nx, ny, nz, _u8(m)), "fast_rasterize_sph")
    return m


def getSubvolume(mask: np.ndarray, origin, size) -> np.ndarray:
    """Volume::getSubvolume (Volume.cpp:129-160): the box [origin, origin + size), (x, y, z)."""
    m, nx, ny, nz = _mask_xyz(mask)
    sx, sy, sz = size
    out = np.empty((sz, sy, sx), dtype=np.uint8)
    check(lib().fast_crop(_u8(m), nx, ny, nz, origin[0], origin[1], origin[2], sx, sy, sz, _u8(out)), "fast_crop")
    return out


def generate_spheres(shape, radius: float, solid_fraction: float, seed: int) -> np.ndarray:
    """[z, y, x] volume of overlapping spheres (255 = solid) filled to `solid_fraction`."""
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    m = np.empty((nz, ny, nx), dtype=np.uint8)
    if lib().fast_generate_spheres(_u8(m), nx, ny, nz, radius, solid_fraction, seed) < 0:
        raise ValueError("invalid generator arguments")
    return m


def generate_ellipsoids(shape, a: float, r: float, solid_fraction: float, seed: int) -> np.ndarray:
    """[z, y, x] volume of randomly oriented ellipsoids, semi-axes (a, a r, a r)."""
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    m = np.empty((nz, ny, nx), dtype=np.uint8)
    if lib().fast_generate_ellipsoids(_u8(m), nx, ny, nz, a, r, solid_fraction, seed) < 0:
        raise ValueError("invalid generator arguments")
    return m
