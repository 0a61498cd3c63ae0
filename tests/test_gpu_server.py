"""Continuous-batching server: every request's result is bit-identical to solving its volume
alone in the mixed form (CG on the red-black Schur complement, or BiCGSTAB on the full grid),
whatever else was in flight, and admission respects the HBM budget."""
import numpy as np
import pytest

import fast_amd as F
from fast_amd import tau as T
from fast_amd._lib import FastError

pytestmark = pytest.mark.gpu
CG, BICG = 1, 0


def _requests():
    # uneven sizes and directions; seeds fixed
    out = []
    shapes = [(48, 40, 56), (32, 32, 32), (20, 64, 36), (64, 24, 24), (40, 40, 40), (17, 33, 29), (56, 48, 30)]
    for i, (nz, ny, nx) in enumerate(shapes):
        m = F.generate_spheres((nx, ny, nz), 4.0 + (i % 3), 0.35 + 0.05 * (i % 4), 10 + i)
        out.append((m, [0, 1, 2, 3, 4, 5][i % 6]))
    return out


@pytest.mark.parametrize("solver", [CG, BICG], ids=["cg", "bicgstab"])
def test_server_matches_solo(ctx, solver):
    reqs = _requests()
    solo = []
    for m, d in reqs:
        vol = ctx.volume(m)
        res, _ = vol.solve(T.make_params(d, solver, "mixed", 1e-7, 10000))
        vol.close()
        solo.append(res)
    # few slots and short rounds: requests join and leave while others are running
    with T.Server(ctx, max_active=3, iters_per_round=5, use_bicgstab=(solver == BICG)) as srv:
        tickets = {srv.submit(m, T.make_params(d, solver, "mixed", 1e-7, 10000)): i for i, (m, d) in enumerate(reqs)}
        got = {}
        while len(got) < len(reqs):
            r = srv.wait(60000)
            assert r is not None, "server timed out"
            t, res, lat = r
            got[tickets[t]] = (res, lat)
        assert srv.pending() == 0
    for i, (res, lat) in got.items():
        ref = solo[i]
        print(f"req {i} tau={res.tau!r} it={res.iterations} upd={res.refinements} lat={lat:.1f} ms | "
              f"solo tau={ref.tau!r} it={ref.iterations}")
        assert res.status == 0 and ref.status == 0
        assert res.tau == ref.tau and res.iterations == ref.iterations and res.refinements == ref.refinements
        assert res.true_rel_residual == ref.true_rel_residual
        assert res.porosity == ref.porosity
        assert lat > 0


def test_server_budget_admission(ctx):
    """An arena holding about two and a half 32^3 volumes still serves eight of them, in FIFO
    admission, and a request larger than the whole arena is refused at submit.  A red-black
    request takes 46 B per element of its padded per-colour layout, (16 + 2) x 34 x 34
    elements (64-aligned) at 32^3."""
    per_request = 46 * ((18 * 34 * 34 + 63) // 64 * 64)
    with T.Server(ctx, hbm_budget=int(2.5 * per_request), max_active=8, iters_per_round=4) as srv:
        masks = [F.generate_spheres(32, 3.0, 0.4, 100 + i) for i in range(8)]
        tickets = [srv.submit(m, T.make_params(4, CG, "mixed", 1e-6, 10000)) for m in masks]
        seen = []
        while len(seen) < len(masks):
            r = srv.wait(60000)
            assert r is not None
            seen.append(r[0])
            assert r[1].status == 0 and r[1].tau > 0
        assert sorted(seen) == sorted(tickets)
        big = np.zeros((64, 64, 64), np.uint8)
        with pytest.raises(FastError):
            srv.submit(big, T.make_params(0, CG, "mixed", 1e-6, 10000))
