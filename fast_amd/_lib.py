"""ctypes binding of libfasttau.so (the C ABI declared in include/fast_tau.h).

The library is built in-tree by fast_amd/build.py.  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfasttau.so"

FAST_OK = 0
FAST_E_BADARG = -1
FAST_E_OOM = -2
FAST_E_NOCONV = -3
FAST_E_HIP = -4
FAST_E_NODEV = -5
FAST_E_IO = -6
FAST_E_BUSY = -7

ERROR_NAMES = {
    FAST_E_BADARG: "FAST_E_BADARG",
    FAST_E_OOM: "FAST_E_OOM",
    FAST_E_NOCONV: "FAST_E_NOCONV",
    FAST_E_HIP: "FAST_E_HIP",
    FAST_E_NODEV: "FAST_E_NODEV",
    FAST_E_IO: "FAST_E_IO",
    FAST_E_BUSY: "FAST_E_BUSY",
}


class TauParams(C.Structure):
    _fields_ = [
        ("dir", C.c_int),
        ("solver", C.c_int),
        ("precision", C.c_int),
        ("tolerance", C.c_double),
        ("max_iter", C.c_long),
        ("d_pore", C.c_double),
        ("d_solid", C.c_double),
        ("porosity", C.c_double),
        ("mixed_inner_tol", C.c_double),
        ("mixed_outer_tol", C.c_double),
        ("verbose", C.c_int),
        ("mixed_full_grid", C.c_int),
        ("flux_nonuniform", C.c_int),
        ("reserved", C.c_int * 1),
    ]


class TauResult(C.Structure):
    _fields_ = [
        ("tau", C.c_double),
        ("porosity", C.c_double),
        ("dc", C.c_double),
        ("rel_residual", C.c_double),
        ("true_rel_residual", C.c_double),
        ("iterations", C.c_long),
        ("status", C.c_int),
        ("restarts", C.c_int),
        ("refinements", C.c_int),
        ("reserved", C.c_int),
        ("solve_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_size_t),
        ("max_active", C.c_int),
        ("iters_per_round", C.c_int),
        ("use_bicgstab", C.c_int),
        ("reserved", C.c_int * 3),
    ]


# (name, restype, argtypes) for every entry point of include/fast_tau.h
class SegmentInfo(C.Structure):
    """fast_segment_info: getCCLSegmentInfo's CCLSegmentInfo (VolumeSegmentation.h:58-64)."""
    _fields_ = [
        ("min_bb", C.c_int32 * 3),
        ("max_bb", C.c_int32 * 3),
        ("voxel_num", C.c_int64),
        ("at_boundary", C.c_int32),
        ("label_id", C.c_int32),
    ]


_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p
PROTOTYPES = [
    ("fast_tau_params_default", None, [C.POINTER(TauParams)]),
    ("fast_ctx_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(_vp)]),
    ("fast_ctx_destroy", None, [_vp]),
    ("fast_ctx_last_error", C.c_char_p, [_vp]),
    ("fast_ctx_device_bytes", C.c_size_t, [_vp]),
    ("fast_ctx_set_tile_shape", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("fast_ctx_set_rb_shape", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("fast_ctx_rb_shape", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    ("fast_ctx_set_round", C.c_int, [_vp, C.c_int]),
    ("fast_volume_create", C.c_int, [_vp, _u8p, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("fast_volume_create_device", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("fast_volume_destroy", None, [_vp]),
    ("fast_volume_count_zero", C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    ("fast_porosity", C.c_int, [_vp, _u8p, C.c_int, C.c_int, C.c_int, _dp]),
    ("fast_tau_solve", C.c_int, [_vp, _u8p, C.c_int, C.c_int, C.c_int, C.POINTER(TauParams), C.POINTER(TauResult), _dp]),
    ("fast_tau_solve_volume", C.c_int, [_vp, C.POINTER(TauParams), C.POINTER(TauResult), _dp]),
    ("fast_tau_solve_batch", C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(TauParams), C.POINTER(TauResult)]),
    ("fast_ccl_labels", C.c_int, [_vp, C.c_uint8, _u32p, _u32p]),
    ("fast_reachable_mask", C.c_int, [_vp, C.c_int, C.c_uint8, _u8p, _u32p]),
    ("fast_ccl_boundary_table", C.c_int, [_vp, C.c_uint8, _u8p, C.c_uint32]),
    ("fast_ccl_segment_info", C.c_int, [_vp, C.c_uint8, C.POINTER(SegmentInfo), C.c_uint32]),
    ("fast_alpha_area", C.c_int, [_vp, C.c_float, C.POINTER(C.c_double)]),
    ("fast_alpha_area_noalloc", C.c_int, [_vp, C.c_float, C.POINTER(C.c_double)]),
    ("fast_debug_stream", C.c_int, [_vp, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    ("fast_alpha_mesh", C.c_int, [_vp, C.c_float, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_size_t)]),
    ("fast_alpha_save_obj", C.c_int, [_vp, C.c_float, C.c_char_p, C.POINTER(C.c_size_t)]),
    ("fast_alpha_area_tensor", C.c_int, [_vp, C.c_float, C.POINTER(C.c_double)]),
    ("fast_debug_time_alpha", C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
    ("fast_server_create", C.c_int, [_vp, C.POINTER(ServerConfig), C.POINTER(_vp)]),
    ("fast_server_submit", C.c_int, [_vp, _u8p, C.c_int, C.c_int, C.c_int, C.POINTER(TauParams), C.POINTER(C.c_uint64)]),
    ("fast_server_wait", C.c_int, [_vp, C.c_int, C.POINTER(C.c_uint64), C.POINTER(TauResult), _dp]),
    ("fast_server_pending", C.c_int, [_vp]),
    ("fast_server_destroy", None, [_vp]),
    ("fast_load_bin_dims", C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fast_load_bin", C.c_int, [C.c_char_p, _u8p, C.c_size_t]),
    ("fast_tiff_folder_dims", C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fast_load_tiff_folder", C.c_int, [C.c_char_p, _u8p, C.c_size_t]),
    ("fast_binarize", None, [_u8p, C.c_size_t, C.c_float]),
    ("fast_rasterize_sph", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _u8p]),
    ("fast_crop", C.c_int, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                            _u8p]),
    ("fast_save_bin", C.c_int, [C.c_char_p, _u8p, C.c_int, C.c_int, C.c_int]),
    ("fast_generate_spheres", C.c_long, [_u8p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64]),
    ("fast_generate_ellipsoids", C.c_long, [_u8p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64]),
    ("fast_debug_apply_f64", C.c_int, [_vp, C.c_int, _dp, _dp]),
    ("fast_debug_apply_f32", C.c_int, [_vp, C.c_int, _fp, _fp]),
    ("fast_debug_system_f64", C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    ("fast_debug_time_passes", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    ("fast_debug_time_rb_passes", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
]

_lib = None


def lib() -> C.CDLL:
    """Load libfasttau.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python fast_amd/build.py`")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, res, args in PROTOTYPES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class FastError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {ERROR_NAMES.get(code, code)}")
        self.code = code


def check(code: int, what: str, ctx=None) -> int:
    if code != FAST_OK:
        msg = what
        if ctx:
            detail = lib().fast_ctx_last_error(ctx)
            if detail:
                msg += f" ({detail.decode(errors='replace')})"
        raise FastError(code, msg)
    return code
