"""ctypes loader of the CPU oracle (oracle/tau_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product (fast_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class OracleResult(C.Structure):
    _fields_ = [
        ("tau", C.c_double),
        ("porosity", C.c_double),
        ("error", C.c_double),
        ("dc", C.c_double),
        ("iterations", C.c_long),
        ("status", C.c_int),
        ("restarts", C.c_int),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        u8, dp, fp, u32 = C.POINTER(C.c_uint8), C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        h.oracle_count_zero.restype = C.c_uint64
        h.oracle_count_zero.argtypes = [u8, C.c_int, C.c_int, C.c_int]
        h.oracle_porosity.restype = C.c_float
        h.oracle_porosity.argtypes = [u8, C.c_int, C.c_int, C.c_int]
        h.oracle_system_f64.restype = None
        h.oracle_system_f64.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        h.oracle_apply_f64.restype = None
        h.oracle_apply_f64.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        h.oracle_apply_f32.restype = None
        h.oracle_apply_f32.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp]
        h.oracle_invdiag_f64.restype = None
        h.oracle_invdiag_f64.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        h.oracle_tau.restype = C.c_int
        h.oracle_tau.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_long,
                                 C.c_int, C.c_double, dp, C.POINTER(OracleResult), C.c_int]
        h.oracle_flux_tau_f64.restype = C.c_double
        h.oracle_flux_tau_f64.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_int]
        h.oracle_ccl.restype = C.c_uint32
        h.oracle_ccl.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_uint8, u32]
        h.oracle_boundary_table.restype = None
        h.oracle_boundary_table.argtypes = [u32, C.c_int, C.c_int, C.c_int, C.c_uint32, u8]
        h.oracle_reachable.restype = C.c_uint32
        h.oracle_reachable.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint8, u8, u32]
        h.oracle_mc_table.restype = None
        h.oracle_mc_table.argtypes = [C.POINTER(C.c_uint64)]
        h.oracle_mc_cube_area.restype = C.c_double
        h.oracle_mc_cube_area.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float]
        h.oracle_mc_area.restype = C.c_double
        h.oracle_mc_area.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_float]
        h.oracle_mc_mesh.restype = C.c_size_t
        h.oracle_mc_mesh.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_float, fp, C.c_size_t]
        h.oracle_mc_area_tensor.restype = None
        h.oracle_mc_area_tensor.argtypes = [u8, C.c_int, C.c_int, C.c_int, C.c_float, dp]
        _lib = h
    return _lib


def _m(mask):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nz, ny, nx = m.shape
    return m, m.ctypes.data_as(C.POINTER(C.c_uint8)), nx, ny, nz


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def porosity(mask) -> float:
    m, p, nx, ny, nz = _m(mask)
    return float(lib().oracle_porosity(p, nx, ny, nz))


def count_zero(mask) -> int:
    m, p, nx, ny, nz = _m(mask)
    return int(lib().oracle_count_zero(p, nx, ny, nz))


def system(mask, d):
    """(coef7 [..., 7] in Dir order + diag, rhs, x0) of the assembled FP64 system."""
    m, p, nx, ny, nz = _m(mask)
    coef = np.empty(m.shape + (7,), dtype=np.float64)
    rhs = np.empty(m.shape, dtype=np.float64)
    x0 = np.empty(m.shape, dtype=np.float64)
    lib().oracle_system_f64(p, nx, ny, nz, int(d), _dp(coef), _dp(rhs), _dp(x0))
    return coef, rhs, x0


def apply(mask, d, x, precision="double"):
    m, p, nx, ny, nz = _m(mask)
    if precision in ("float", "single"):
        xx = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty_like(xx)
        lib().oracle_apply_f32(p, nx, ny, nz, int(d), xx.ctypes.data_as(C.POINTER(C.c_float)),
                               y.ctypes.data_as(C.POINTER(C.c_float)))
        return y
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(xx)
    lib().oracle_apply_f64(p, nx, ny, nz, int(d), _dp(xx), _dp(y))
    return y


def invdiag(mask, d):
    m, p, nx, ny, nz = _m(mask)
    out = np.empty(m.shape, dtype=np.float64)
    lib().oracle_invdiag_f64(p, nx, ny, nz, int(d), _dp(out))
    return out


def tau(mask, d, solver=0, precision=0, tol=1e-6, max_iter=10000, threads=0, porosity=-1.0, conc=False,
        nonuniform=False):
    """Restated getTortuosity<T> on the CPU.  solver 0 = BiCGSTAB, 1 = CG; precision 0 = double,
    1 = float; nonuniform = TortuosityParams::useNonUniform.  Returns (OracleResult, conc or None)."""
    m, p, nx, ny, nz = _m(mask)
    r = OracleResult()
    c = np.empty(m.shape, dtype=np.float64) if conc else None
    lib().oracle_tau(p, nx, ny, nz, int(d), int(solver), int(precision), float(tol), int(max_iter), int(threads),
                     float(porosity), _dp(c) if conc else None, C.byref(r), int(bool(nonuniform)))
    return r, c


def flux_tau(mask, d, c, porosity, nonuniform=False):
    m, p, nx, ny, nz = _m(mask)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    return float(lib().oracle_flux_tau_f64(p, nx, ny, nz, int(d), _dp(cc), float(porosity), int(bool(nonuniform))))


def ccl(mask, background=255):
    m, p, nx, ny, nz = _m(mask)
    lab = np.empty(m.shape, dtype=np.uint32)
    nl = lib().oracle_ccl(p, nx, ny, nz, background, lab.ctypes.data_as(C.POINTER(C.c_uint32)))
    return lab, int(nl)


def boundary_table(labels, num_labels):
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    nz, ny, nx = lab.shape
    t = np.empty((7, num_labels), dtype=np.uint8)
    lib().oracle_boundary_table(lab.ctypes.data_as(C.POINTER(C.c_uint32)), nx, ny, nz, num_labels,
                                t.ctypes.data_as(C.POINTER(C.c_uint8)))
    return t


def reachable(mask, d, background=255):
    m, p, nx, ny, nz = _m(mask)
    out = np.empty(m.shape, dtype=np.uint8)
    lab = np.empty(m.shape, dtype=np.uint32)
    nl = lib().oracle_reachable(p, nx, ny, nz, int(d), background, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                lab.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out, lab, int(nl)


def mc_table():
    """The triangle table as lists of edge ids per cube configuration (256 lists)."""
    w = (C.c_uint64 * 256)()
    lib().oracle_mc_table(w)
    out = []
    for c in range(256):
        v = int(w[c])
        n = v >> 60
        out.append([(v >> (4 * j)) & 15 for j in range(3 * n)])
    return out


def mc_cube_area(mask, x, y, z, isovalue=0.5):
    m, p, nx, ny, nz = _m(mask)
    return lib().oracle_mc_cube_area(p, nx, ny, nz, int(x), int(y), int(z), float(isovalue))


def mc_area(mask, isovalue=0.5):
    """getReactiveAreaDensity: marching-cubes area of the mask (alpha_oracle.cpp)."""
    m, p, nx, ny, nz = _m(mask)
    return lib().oracle_mc_area(p, nx, ny, nz, float(isovalue))


def mc_area_tensor(mask, isovalue=0.5):
    """getReactiveAreaDensityTensor over getVolumeCCL(mask, 255): six per-face areas."""
    m, p, nx, ny, nz = _m(mask)
    out = np.empty(6, dtype=np.float64)
    lib().oracle_mc_area_tensor(p, nx, ny, nz, float(isovalue), _dp(out))
    return out


def mc_mesh(mask, isovalue=0.5):
    """The --mesh triangles: (n_vertices, 6) float32 array of position and triangle normal."""
    m, p, nx, ny, nz = _m(mask)
    n = lib().oracle_mc_mesh(p, nx, ny, nz, float(isovalue), None, 0)
    out = np.empty((n, 6), dtype=np.float32)
    lib().oracle_mc_mesh(p, nx, ny, nz, float(isovalue), out.ctypes.data_as(C.POINTER(C.c_float)), n)
    return out
