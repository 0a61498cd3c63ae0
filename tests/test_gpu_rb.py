"""Mixed CG on the red-black Schur complement (the default mixed CG) against the CPU oracle's
converged solution and against mixed CG on the full grid.

The reduced system eliminates the black voxels exactly (in FP64), so it solves the same linear
system as the reference: its tau and concentration field converge to the same values, and its
FP64 true residual is that of the full system.  What differs is the Krylov process (CG on S,
about half the iterations), so the bars are accuracy bars, not bit-identity."""
import numpy as np
import pytest

import fast_amd as F
from fast_amd import tau as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
BICG, CG = 0, 1


def rand_mask(shape, frac, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random(shape) < frac, 255, 0).astype(np.uint8)


# (nz, ny, nx): odd and even extents, one-voxel rows and planes, a volume narrower than a tile
RAGGED = [
    ("odd_37x23x19", lambda: rand_mask((19, 23, 37), 0.45, 1)),
    ("even_40x30x20", lambda: F.generate_spheres((40, 30, 20), 3.0, 0.5, 2)),
    ("tall_1x5x70", lambda: rand_mask((70, 5, 1), 0.3, 3)),
    ("thin_67x3x1", lambda: rand_mask((1, 3, 67), 0.3, 4)),
    ("wide_131x17x11", lambda: rand_mask((11, 17, 131), 0.4, 5)),
    ("cube_9", lambda: rand_mask((9, 9, 9), 0.4, 6)),
]


@pytest.mark.parametrize("name,make", RAGGED, ids=[r[0] for r in RAGGED])
def test_reduced_matches_converged_oracle_on_ragged_shapes(ctx, name, make):
    """Every direction at outer tolerance 1e-11: tau within 1e-7 of the oracle's CG at 1e-13,
    the concentration field within 1e-6 (absolute; it lies in [0, 1])."""
    m = make()
    vol = ctx.volume(m)
    try:
        for d in range(6):
            ref, cref = O.tau(m, d, solver=CG, tol=1e-13, conc=True)
            res, c = vol.solve(T.make_params(d, CG, "mixed", 1e-11, 20000), conc=True)
            assert res.status == 0, (name, d, res.iterations)
            assert res.true_rel_residual <= 1e-11
            if np.isinf(ref.tau) or ref.tau == 0:
                assert res.tau == ref.tau or (np.isinf(ref.tau) and np.isinf(res.tau)), (name, d, res.tau, ref.tau)
                continue
            err = abs(res.tau - ref.tau) / ref.tau
            dc = np.abs(c - cref).max()
            print(f"{name} d={d}: tau={res.tau!r} oracle={ref.tau!r} err={err:.2e} it={res.iterations} "
                  f"(oracle {ref.iterations}) max|dc|={dc:.2e}")
            assert err <= 1e-7, (name, d)
            assert dc <= 1e-6, (name, d)
    finally:
        vol.close()


def test_reduced_needs_about_half_the_iterations(ctx):
    """64^3 spheres (radius 6, solid 0.60), X: CG on S converges in at most 0.6x the
    iterations of CG on the full grid, to the same tau."""
    m = F.generate_spheres(64, 6.0, 0.60, 0)
    vol = ctx.volume(m)
    tight, _ = O.tau(m, 0, solver=CG, tol=1e-12)
    for tol in (1e-6, 1e-8):
        rb, _ = vol.solve(T.make_params(0, CG, "mixed", tol, 10000))
        full, _ = vol.solve(T.make_params(0, CG, "mixed", tol, 10000, full_grid=True))
        print(f"tol {tol:g}: reduced it={rb.iterations} updates={rb.refinements} err="
              f"{abs(rb.tau - tight.tau) / tight.tau:.2e}; full it={full.iterations} err="
              f"{abs(full.tau - tight.tau) / tight.tau:.2e}")
        assert rb.status == 0 and full.status == 0
        assert rb.iterations <= 0.6 * full.iterations
        assert abs(rb.tau - tight.tau) / tight.tau <= (3e-5 if tol == 1e-6 else 5e-7)
    vol.close()


def test_reduced_closed_forms(ctx):
    """All pore: tau = 1 in every direction.  Parallel slabs along X: tau =
    float(phi) / (phi + 0.001f (1 - phi)) (slabs in parallel: the arithmetic mean)."""
    m = np.zeros((21, 16, 33), np.uint8)
    vol = ctx.volume(m)
    for d in range(6):
        res, _ = vol.solve(T.make_params(d, CG, "mixed", 1e-10, 10000))
        assert res.status == 0 and abs(res.tau - 1.0) < 1e-9, (d, res.tau)
    vol.close()
    m = np.zeros((30, 27, 45), np.uint8)
    m[:, ::3, :] = 255
    phi = (m == 0).mean()
    expected = float(np.float32(phi)) / (phi + float(np.float32(0.001)) * (1 - phi))
    vol = ctx.volume(m)
    res, _ = vol.solve(T.make_params(0, CG, "mixed", 1e-11, 10000))
    assert res.status == 0 and abs(res.tau - expected) < 1e-8 * expected, (res.tau, expected)
    vol.close()


def test_reduced_max_iter_reports_no_convergence(ctx):
    m = F.generate_spheres(64, 6.0, 0.60, 0)
    vol = ctx.volume(m)
    res, _ = vol.solve(T.make_params(0, CG, "mixed", 1e-8, 5))
    assert res.status == 1 and res.tau == 0.0 and res.iterations == 5
    vol.close()


def test_reduced_batch_equals_solo_uneven(ctx):
    """Volumes of different sizes and directions in one batch: each result is bit-identical
    to its solo solve (per-volume dot products reduce in a fixed order)."""
    shapes = [(33, 20, 17), (64, 64, 64), (9, 40, 31), (50, 21, 48)]
    vols, ps = [], []
    for i, sh in enumerate(shapes):
        vols.append(ctx.volume(F.generate_spheres(sh, 3.0 + i, 0.45, 40 + i)))
        ps.append(T.make_params(i % 6, CG, "mixed", 1e-8, 10000))
    batch = T.solve_batch(ctx, vols, ps)
    for v, p, b in zip(vols, ps, batch):
        solo, _ = v.solve(p)
        assert b.status == 0 and solo.status == 0
        assert b.tau == solo.tau and b.iterations == solo.iterations and b.refinements == solo.refinements
        v.close()


def test_y_and_q_forms_bit_identical(ctx, monkeypatch):
    """The three forms of the Y and Q passes compute the same values in the same order: Y and
    Q separate with Q reading the six Y neighbours from memory (0), Q staging the Y planes in
    LDS (1), and Y fused into Q (2, the default).  tau, iterations and the concentration field
    are identical, on cubes and on odd shapes whose tiles are ragged in every axis, with tiles
    of a few planes and of many."""
    cases = [F.generate_spheres(64, 6.0, 0.60, 0), rand_mask((19, 23, 37), 0.45, 1), rand_mask((11, 17, 131), 0.4, 5),
             rand_mask((70, 5, 1), 0.3, 3), rand_mask((33, 9, 66), 0.5, 7), rand_mask((1, 1, 40), 0.2, 8)]
    for m in cases:
        vol = ctx.volume(m)
        for d in (0, 3, 4):
            p = T.make_params(d, CG, "mixed", 1e-9, 10000)
            out = []
            for mode in ("0", "1", "2"):
                monkeypatch.setenv("FAST_TAU_RB_STAGED", mode)
                out.append(vol.solve(p, conc=True))
            monkeypatch.delenv("FAST_TAU_RB_STAGED")
            (a, ca) = out[0]
            assert a.status == 0
            for b, cb in out[1:]:
                assert b.status == 0
                assert a.tau == b.tau and a.iterations == b.iterations, (m.shape, d, a.tau, b.tau)
                assert np.array_equal(ca, cb), (m.shape, d)
        vol.close()


def test_compacted_batch_equals_solo(ctx):
    """Volumes that converge at different iterations: once the unfinished ones hold at most
    half of the launched tiles, later rounds launch only theirs (a new graph is captured).
    Every result stays bit-identical to its solo solve."""
    vols, ps = [], []
    for i in range(14):   # nx, ny, nz = 24, 40, 96: flow along x (24 planes) converges far sooner than along z
        m = F.generate_spheres((24, 40, 96), 2.5 + 0.25 * (i % 7), 0.4 + 0.02 * (i % 7), 70 + i)
        vols.append(ctx.volume(m))
        ps.append(T.make_params(0 if i % 2 == 0 else 4, CG, "mixed", 1e-9, 20000))
    solo = [v.solve(p)[0] for v, p in zip(vols, ps)]
    fast = max(r.iterations for i, r in enumerate(solo) if i % 2 == 0)
    slow = min(r.iterations for i, r in enumerate(solo) if i % 2 == 1)
    # the x half finishes first, several 16-iteration polling rounds before the z half: the
    # batch has a tail, and half the tiles are left to launch
    assert slow > fast + 48, (fast, slow)
    batch = T.solve_batch(ctx, vols, ps)
    for b, r in zip(batch, solo):
        assert b.status == 0 and r.status == 0
        assert b.tau == r.tau and b.iterations == r.iterations
    for v in vols:
        v.close()


def test_tile_shapes_bit_identical(ctx):
    """Rows per tile (2, 4, 8, 9, 10, 12, 13), planes per tile and the fused pass's prefetch distance change
    only the launch geometry: every dot product is rounded once from a double-double sum, so
    tau, iterations and the field are identical, on a cube and on shapes ragged in every axis."""
    cases = [F.generate_spheres(48, 5.0, 0.55, 2), rand_mask((21, 13, 29), 0.45, 11), rand_mask((66, 9, 5), 0.35, 12)]
    for m in cases:
        vol = ctx.volume(m)
        p = T.make_params(4, CG, "mixed", 1e-9, 10000)
        ref, cref = vol.solve(p, conc=True)
        assert ref.status == 0
        for ty, tz, ahead in ((2, 0, 3), (8, 0, 3), (4, 0, 2), (2, 3, 2), (8, 5, 3), (4, 32, 3), (10, 0, 2),
                              (13, 7, 3), (9, 0, 2), (9, 6, 3), (12, 0, 2), (12, 4, 3)):
            ctx.set_rb_shape(ty, tz, ahead)
            got, c = vol.solve(p, conc=True)
            ctx.set_rb_shape(0)
            assert got.tau == ref.tau and got.iterations == ref.iterations, (m.shape, ty, tz, ahead)
            assert np.array_equal(c, cref), (m.shape, ty, tz, ahead)
        vol.close()


def test_fused_form_bit_identical_with_forced_tile_depths(ctx, monkeypatch):
    """Tile depths of 1 to 7 planes (FAST_TAU_RB_TZ) put the fused pass's prologue and the
    unrolled march's spare step at every offset from the volume's last plane."""
    m = rand_mask((13, 21, 70), 0.45, 9)
    vol = ctx.volume(m)
    p = T.make_params(2, CG, "mixed", 1e-9, 10000)
    monkeypatch.setenv("FAST_TAU_RB_STAGED", "0")
    ref, cref = vol.solve(p, conc=True)
    monkeypatch.delenv("FAST_TAU_RB_STAGED")
    for tz in range(1, 8):
        monkeypatch.setenv("FAST_TAU_RB_TZ", str(tz))
        got, c = vol.solve(p, conc=True)
        assert got.tau == ref.tau and got.iterations == ref.iterations and np.array_equal(c, cref), tz
    monkeypatch.delenv("FAST_TAU_RB_TZ")
    vol.close()


def test_large_volume_batch_runs_in_cache_sized_groups(monkeypatch):
    """A batch of volumes that are each large for the last-level cache runs as consecutive
    sub-batches (here the group size is lowered to 8 MB so that 64^3 volumes count as large);
    every result stays bit-identical to its solo solve."""
    monkeypatch.setenv("FAST_TAU_GROUP_MB", "8")
    c = T.Context(0)
    try:
        vols = [c.volume(F.generate_spheres(64, 4.0 + i, 0.5, 60 + i)) for i in range(3)]
        ps = [T.make_params(d, CG, "mixed", 1e-8, 10000) for d in (0, 2, 5)]
        batch = T.solve_batch(c, vols, ps)
        for v, p, b in zip(vols, ps, batch):
            solo, _ = v.solve(p)
            assert b.status == 0 and b.tau == solo.tau and b.iterations == solo.iterations
            v.close()
    finally:
        c.close()
