"""Serving on the request mix of the Poisson-stream workload: cubic volumes of edge 64, 128,
192, 256 and 384 (ellipsoids, r log-uniform on [0.1, 10], solid fraction uniform on
[0.5, 0.75]), directions x, y and z, all submitted while others are solving.  Each served
result must be bit-identical to the same mixed solve run alone, and its tau within 1e-5 of
the converged FP64 tau (CG to 1e-11, the FP64 path whose recurrences are bit-identical to
the oracle's).  The reference has no request loop; its many-volume loop is the sequential
fast/src/SliceModule.h:120-133, and a failed solve reports tau = 0 (TauModule.h:34-118)."""
import numpy as np
import pytest

from fast_amd import tau as T
from fast_amd import workloads as W

pytestmark = pytest.mark.gpu
CG = 1
TOL = 1e-9          # FP64 true relative residual of the served mixed solves
TAU_BAR = 1e-5      # |tau - tau_FP64| / tau_FP64

# (edge, pool index, direction): every edge of the mix, the 384^3 class included
REQUESTS = [(64, 0, 0), (128, 0, 2), (192, 0, 4), (256, 0, 0), (384, 0, 2), (64, 1, 4), (128, 1, 0),
            (64, 2, 2), (192, 1, 2), (128, 2, 4), (256, 1, 4), (64, 3, 0)]


@pytest.fixture(scope="module")
def requests():
    vols = {}
    for e, j, _ in REQUESTS:
        if (e, j) not in vols:
            vols[(e, j)] = W.request_volume(e, j)
    return vols


def test_served_results_equal_solo_and_meet_the_tau_bar(ctx, requests):
    mk = lambda d: T.make_params(d, CG, "mixed", TOL, 40000)
    with T.Server(ctx, max_active=8, iters_per_round=16) as srv:
        tickets = {}
        for i, (e, j, d) in enumerate(REQUESTS):
            tickets[srv.submit(requests[(e, j)].mask, mk(d))] = i
        print(f"config5: {len(REQUESTS)} requests submitted", flush=True)
        served = {}
        while len(served) < len(REQUESTS):
            r = srv.wait(120000)
            assert r is not None, "server timed out"
            t, res, lat = r
            served[tickets[t]] = (res, lat)
        assert srv.pending() == 0
    worst = 0.0
    for i, (e, j, d) in enumerate(REQUESTS):
        res, lat = served[i]
        vol = ctx.volume(requests[(e, j)].mask)
        solo, _ = vol.solve(mk(d))
        conv, _ = vol.solve(T.make_params(d, CG, "double", 1e-11, 60000))
        vol.close()
        err = abs(res.tau - conv.tau) / conv.tau
        worst = max(worst, err)
        print(f"req {i}: {e}^3 r={requests[(e, j)].r:.2f} dir {d} tau={res.tau!r} it={res.iterations} "
              f"lat={lat:.0f} ms | solo it={solo.iterations} | FP64 tau={conv.tau!r} ({conv.iterations} it) "
              f"rel err {err:.1e}", flush=True)
        assert res.status == 0 and solo.status == 0 and conv.status == 0
        assert res.tau == solo.tau and res.iterations == solo.iterations
        assert res.true_rel_residual == solo.true_rel_residual <= TOL
        assert err <= TAU_BAR
    print(f"config5: max tau rel err {worst:.1e}")
