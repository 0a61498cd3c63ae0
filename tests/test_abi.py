"""The C ABI library loads, exports every declared entry point, and its host-only pieces
(defaults, generators, .bin I/O, CLI) behave -- CPU only, no compute on a GPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fast_amd as F
from fast_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "fast_tau.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fast_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = _lib.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fast_\w+)", out))
    assert set(names) <= exported
    assert {p[0] for p in _lib.PROTOTYPES} == set(names)


def test_params_default_match_reference_cli():
    p = _lib.TauParams()
    _lib.lib().fast_tau_params_default(C.byref(p))
    # TauModule.h:15-23 defaults: dir x -> X_POS, CG, double, tol 1e-6, 10000 iterations
    assert p.dir == 0 and p.solver == 1 and p.precision == 0
    assert p.tolerance == 1e-6 and p.max_iter == 10000
    assert p.d_pore == 1.0 and p.d_solid == 0.001 and p.porosity < 0


def test_context_without_gpu_fails_cleanly():
    h = C.c_void_p()
    rc = _lib.lib().fast_ctx_create(0, 0, C.byref(h))
    assert rc in (_lib.FAST_OK, _lib.FAST_E_NODEV)
    if rc == _lib.FAST_OK:
        _lib.lib().fast_ctx_destroy(h)
    else:
        assert not h.value


def test_bad_arguments_rejected_without_device():
    assert _lib.lib().fast_tau_solve(None, None, 1, 1, 1, None, None, None) == _lib.FAST_E_BADARG
    assert _lib.lib().fast_volume_count_zero(None, None) == _lib.FAST_E_BADARG


def test_generators_deterministic_and_filled():
    a = F.generate_spheres(32, 6.0, 0.6, 0)
    b = F.generate_spheres(32, 6.0, 0.6, 0)
    c = F.generate_spheres(32, 6.0, 0.6, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert set(np.unique(a)) <= {0, 255}
    frac = (a != 0).mean()
    assert 0.6 <= frac < 0.6 + 4 / 3 * np.pi * 6 ** 3 / a.size + 1e-9
    e = F.generate_ellipsoids((40, 24, 16), 12.0, 0.35, 0.65, 3)
    assert e.shape == (16, 24, 40) and (e != 0).mean() >= 0.65
    # an ellipsoid with r = 1 is a sphere: the voxel set is rotation invariant
    s1 = F.generate_ellipsoids(24, 5.0, 1.0, 0.3, 9)
    assert (s1 != 0).mean() >= 0.3


def test_bin_roundtrip_reference_layout(tmp_path):
    m = F.generate_spheres((20, 12, 7), 3.0, 0.5, 2)
    p = tmp_path / "v.bin"
    F.save_bin(str(p), m)
    raw = p.read_bytes()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    assert list(hdr) == [2, 20, 12, 7]          # TYPE_UCHAR, dims
    assert len(raw) == 17 + m.size and raw[16] == 0
    assert np.array_equal(F.load_bin(str(p)), m)


def test_cli_usage_and_cpu_flag(tmp_path):
    exe = ROOT / "fast_amd" / "fast"
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and "fast tau" in out.stdout
    p = tmp_path / "v.bin"
    F.save_bin(str(p), np.zeros((4, 4, 4), np.uint8))
    out = subprocess.run([str(exe), "tau", str(p), "--cpu"], capture_output=True, text=True)
    assert out.returncode == 1 and "no CPU solver" in out.stderr


def test_get_dirs_matches_reference():
    D = F.Dir
    assert F.getDirs("x-") == [D.X_POS]
    assert F.getDirs("xyz") == [D.X_POS, D.Y_POS, D.Z_POS]
    assert F.getDirs("all") == list(D)
    assert F.getDirs("neg") == [D.X_NEG, D.Y_NEG, D.Z_NEG]
    assert F.getDirs("zneg") == [D.Z_NEG]
    assert F.getDirs("") == [D.X_POS]
