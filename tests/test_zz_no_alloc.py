"""fast_alpha_area_noalloc: the whole-volume marching-cubes area with its block partials in
the library's static device scratch instead of allocated memory.  Through the C ABI, two
calls in a row (the second on a smaller volume, so stale partials of the first call sit
beyond its range) against fast_alpha_area on the same volumes, bit for bit, and the size
limit refused with FAST_E_BADARG before any kernel runs."""
import ctypes as C

import pytest

import fast_amd as F
from fast_amd import _lib

pytestmark = pytest.mark.gpu


def _areas(ctx, mask):
    vol = ctx.volume(mask)
    try:
        ref, got = C.c_double(), C.c_double()
        print(f"no-alloc area: volume {mask.shape[::-1]}", flush=True)
        assert _lib.lib().fast_alpha_area(vol.handle, C.c_float(0.5), C.byref(ref)) == _lib.FAST_OK
        held = ctx.device_bytes()
        rc = _lib.lib().fast_alpha_area_noalloc(vol.handle, C.c_float(0.5), C.byref(got))
        assert rc == _lib.FAST_OK, ctx.last_error()
        assert ctx.device_bytes() == held
        return ref.value, got.value
    finally:
        vol.close()


def test_noalloc_area_equals_allocating_area_on_two_calls(ctx):
    big = F.generate_ellipsoids((200, 144, 120), 12.0, 0.35, 0.65, 3)
    small = F.generate_spheres((45, 33, 21), 4.0, 0.5, 4)
    ref_big, got_big = _areas(ctx, big)
    assert got_big == ref_big and ref_big > 0
    ref_small, got_small = _areas(ctx, small)
    assert got_small == ref_small and ref_small > 0
    # the same small volume again after the large one: nothing of the earlier call leaks in
    assert _areas(ctx, small) == (ref_small, ref_small)


def test_noalloc_area_refuses_volumes_beyond_the_scratch(ctx):
    # 1 x 8 x 524288 voxels: 1 * 2 * 32768 = 65536 blocks of 64x4x16, one more than fits
    import numpy as np
    m = np.zeros((524288, 8, 1), dtype=np.uint8)
    vol = ctx.volume(m)
    try:
        out = C.c_double(-1.0)
        rc = _lib.lib().fast_alpha_area_noalloc(vol.handle, C.c_float(0.5), C.byref(out))
        assert rc == _lib.FAST_E_BADARG
        assert "static scratch" in ctx.last_error()
        assert out.value == -1.0
        # the allocating entry point takes the same volume
        ref = C.c_double()
        assert _lib.lib().fast_alpha_area(vol.handle, C.c_float(0.5), C.byref(ref)) == _lib.FAST_OK
    finally:
        vol.close()
