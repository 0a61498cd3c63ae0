"""512^3 needle-like ellipsoids (a = 12, r = 0.1, solid 0.55): the mixed form driven to a
true relative residual of 1e-8, against the CPU oracle's FP64 results committed in
tests/golden/config3_oracle.json and config3_converged_oracle.json
(tests/golden/make_oracle_fixtures.py config3, hours of CPU time; each record carries the
mask's SHA-256).  The GPU's FP64 CG reproduces the oracle's tau and iteration count bit for
bit at 1e-6 and at 1e-11; the mixed tau must be within 1e-5 of the oracle's converged tau
(directions without an oracle record use the GPU FP64 CG to 1e-11, whose recurrences are the
oracle's).  The reference's CPU path is VolumeMeasures.cpp:42-285."""
import hashlib
import json
from pathlib import Path

import pytest

from fast_amd import tau as T
from fast_amd import workloads as W

pytestmark = pytest.mark.gpu
CG = 1
GOLD = Path(__file__).resolve().parent / "golden"
DIRS = ["X_POS", "X_NEG", "Y_POS", "Y_NEG", "Z_POS", "Z_NEG"]


def _fixture(name):
    p = GOLD / name
    return json.loads(p.read_text()) if p.exists() else {}


FP64 = _fixture("config3_oracle.json")
CONV = _fixture("config3_converged_oracle.json")


@pytest.fixture(scope="module")
def needles(ctx):
    m = W.needles_512(0)
    digest = hashlib.sha256(m.tobytes()).hexdigest()
    for rec in list(FP64.values()) + list(CONV.values()):
        assert rec["sha256"] == digest
    vol = ctx.volume(m)
    yield vol
    vol.close()


@pytest.mark.parametrize("key", sorted(FP64))
def test_fp64_cg_equals_the_oracle_at_512(ctx, needles, key):
    ref = FP64[key]
    res, _ = needles.solve(T.make_params(ref["dir"], CG, "double", ref["tol"], ref["max_iter"]))
    print(f"512^3 needles {key} tol {ref['tol']}: gpu tau={res.tau!r} ({res.iterations} it, {res.solve_ms:.0f} ms), "
          f"oracle tau={ref['tau']!r} ({ref['iterations']} it, {ref['seconds']:.0f} s on the CPU)")
    assert res.status == 0 and ref["status"] == 0
    assert res.tau == ref["tau"] and res.iterations == ref["iterations"]


@pytest.mark.parametrize("d", [0, 2, 4], ids=["x", "y", "z"])
def test_mixed_1e8_matches_converged_fp64(ctx, needles, d):
    conv, _ = needles.solve(T.make_params(d, CG, "double", 1e-11, 40000))
    rec = CONV.get(f"CG_{DIRS[d]}")
    if rec is not None:   # the GPU's converged FP64 solve is the oracle's, bit for bit
        assert conv.tau == rec["tau"] and conv.iterations == rec["iterations"]
    mixed, _ = needles.solve(T.make_params(d, CG, "mixed", 1e-8, 40000))
    err = abs(mixed.tau - conv.tau) / conv.tau
    print(f"512^3 needles d={d}: converged tau={conv.tau!r} ({conv.iterations} it, {conv.solve_ms:.0f} ms, "
          f"{'oracle record' if rec else 'no oracle record'}), mixed tau={mixed.tau!r} ({mixed.iterations} it, "
          f"{mixed.refinements} updates, {mixed.solve_ms:.0f} ms, true res {mixed.true_rel_residual:.2e}) "
          f"rel err {err:.2e}")
    assert conv.status == 0 and mixed.status == 0
    assert mixed.true_rel_residual <= 1e-8
    assert err <= 1e-5
