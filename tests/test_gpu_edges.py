"""Edge cases of the solve: degenerate and tiny volumes in one batch, all-solid volumes, empty
and over-limit volumes (refused, never solved), in every precision and solver form."""
import numpy as np
import pytest

import fast_amd as F
from fast_amd import tau as T
from fast_amd._lib import FastError
from oracle import oracle as O

pytestmark = pytest.mark.gpu
BICG, CG = 0, 1


def _tiny_volumes(n, seed):
    rng = np.random.default_rng(seed)
    out = [np.zeros((1, 1, 1), np.uint8), np.full((1, 1, 2), 255, np.uint8), np.zeros((3, 1, 1), np.uint8)]
    while len(out) < n:
        shape = tuple(int(v) for v in rng.integers(1, 10, size=3))
        out.append(np.where(rng.random(shape) < 0.35, 255, 0).astype(np.uint8))
    return out


@pytest.mark.parametrize("prec,solver", [("double", BICG), ("double", CG), ("mixed", CG), ("mixed", BICG)])
def test_many_tiny_volumes_in_one_batch(ctx, prec, solver):
    """120 volumes of 1..9 voxels per axis (1x1x1 included), every direction, one batch: each
    result equals its solo solve bit for bit; FP64 equals the oracle bit for bit, the mixed form
    agrees with the oracle's converged tau."""
    masks = _tiny_volumes(120, 7)
    vols = [ctx.volume(m) for m in masks]
    tol = 1e-10
    ps = [T.make_params(i % 6, solver, prec, tol, 20000) for i in range(len(vols))]
    batch = T.solve_batch(ctx, vols, ps)
    for i, (m, v, p, b) in enumerate(zip(masks, vols, ps, batch)):
        solo, _ = v.solve(p)
        assert b.status == solo.status and b.iterations == solo.iterations, (i, m.shape)
        assert (np.isnan(b.tau) and np.isnan(solo.tau)) or b.tau == solo.tau, (i, m.shape, b.tau, solo.tau)
        if i % 10 == 0:
            ref, _ = O.tau(m, i % 6, solver=solver if prec == "double" else CG, tol=1e-13 if prec == "mixed" else tol)
            if prec == "double":
                same = b.tau == ref.tau or (np.isinf(b.tau) and np.isinf(ref.tau)) or (np.isnan(b.tau) and np.isnan(ref.tau))
                assert b.iterations == ref.iterations and same, (i, m.shape, b.tau, ref.tau)
            elif b.status == 0 and np.isfinite(ref.tau) and ref.tau != 0:
                assert abs(b.tau - ref.tau) <= 1e-7 * abs(ref.tau), (i, m.shape, b.tau, ref.tau)
    for v in vols:
        v.close()


@pytest.mark.parametrize("prec", ["double", "float", "mixed"])
def test_all_solid_tau_is_zero(ctx, prec):
    """Porosity 0: tau = 0 / (dc N) = 0, as the reference (VolumeMeasures.cpp:275)."""
    vol = ctx.volume(np.full((12, 10, 14), 255, np.uint8))
    for d in (0, 3, 4):
        res, _ = vol.solve(T.make_params(d, CG, prec, 1e-8, 10000))
        assert res.status == 0 and res.porosity == 0.0 and res.tau == 0.0, (prec, d, res.tau)
    vol.close()


def test_empty_volume_refused(ctx):
    with pytest.raises(FastError):
        ctx.volume(np.zeros((0, 4, 4), np.uint8))


def test_volume_over_the_solver_limit_refused(ctx):
    """2^29 voxels (one more than the limit of the 32-bit offset addressing): every solve form
    and the server refuse it with FAST_E_BADARG before allocating solver state."""
    m = np.zeros((512, 1024, 1024), np.uint8)
    vol = ctx.volume(m)
    try:
        for prec, solver in (("double", BICG), ("double", CG), ("float", CG), ("mixed", CG), ("mixed", BICG)):
            with pytest.raises(FastError) as e:
                vol.solve(T.make_params(0, solver, prec, 1e-6, 10))
            assert "limit" in str(e.value), str(e.value)
    finally:
        vol.close()
    with T.Server(ctx, max_active=2) as srv:
        with pytest.raises(FastError):
            srv.submit(m, T.make_params(0, CG, "mixed", 1e-6, 10))
