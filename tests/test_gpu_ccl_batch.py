"""GPU reachable-pore labels (bit-exact) and batched solves (bit-identical to solo solves)."""
import json
from pathlib import Path

import numpy as np
import pytest

import fast_amd as F
from fast_amd import tau as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
BICG, CG = 0, 1


def test_ccl_golden_cases(ctx):
    for r in json.loads((GOLD / "ccl_cases.json").read_text()):
        m = np.array(r["mask"], dtype=np.uint8).reshape(r["shape_zyx"])
        vol = ctx.volume(m)
        lab, nl = vol.ccl_labels(255)
        assert nl == r["num_labels"], r["name"]
        assert np.array_equal(lab.ravel(), np.array(r["labels"], dtype=np.uint32)), r["name"]
        assert np.array_equal(vol.boundary_table(255).ravel(), np.array(r["table"], dtype=np.uint8)), r["name"]


@pytest.mark.parametrize("shape,frac,seed", [((64, 64, 64), None, 0), ((50, 70, 90), 0.6, 1), ((33, 17, 129), 0.5, 2),
                                             ((96, 96, 96), 0.69, 3)])
def test_ccl_and_reachable_match_oracle(ctx, shape, frac, seed):
    if frac is None:
        m = F.generate_spheres(64, 6.0, 0.60, 0)   # the 64^3 sphere volume
    else:
        rng = np.random.default_rng(seed)
        m = ((rng.random(shape) < frac) * 255).astype(np.uint8)
    vol = ctx.volume(m)
    lab, nl = vol.ccl_labels(255)
    olab, onl = O.ccl(m)
    assert nl == onl and np.array_equal(lab, olab)
    for d in (0, 2, 4):
        out, _ = vol.reachable_mask(d, 255)
        oout, _, _ = O.reachable(m, d)
        assert np.array_equal(out, oout), d


def test_ccl_thin_x_volume_with_more_than_2pow26_rows(ctx):
    """1 x 8192 x 8193 voxels: 67,117,056 x-rows, above 2^26, so the span pass (one wave per
    row) needs more waves than one launch may hold and strides over the rows."""
    rng = np.random.default_rng(11)
    m = ((rng.random((8193, 8192, 1)) < 0.35) * 255).astype(np.uint8)   # percolates in z
    print("thin-x CCL: 1 x 8192 x 8193", flush=True)
    vol = ctx.volume(m)
    lab, nl = vol.ccl_labels(255)
    olab, onl = O.ccl(m)
    assert nl == onl and np.array_equal(lab, olab)
    out, _ = vol.reachable_mask(4, 255)
    oout, _, _ = O.reachable(m, 4)
    assert np.array_equal(out, oout)
    vol.close()


@pytest.mark.parametrize("shape,seed", [((40, 36, 70), 4), ((9, 130, 66), 5)])
def test_ccl_multivalued_bytes_match_oracle(ctx, shape, seed):
    """Bytes 0, 100 and 255 in runs: every change of value toggles a span, so equal
    neighbouring bytes are the only sign of a continuing span (the union pass skips links
    already made one voxel earlier on that sign)."""
    rng = np.random.default_rng(seed)
    runs = rng.choice(np.array([0, 100, 255], np.uint8), size=shape)
    m = np.repeat(runs, 3, axis=2)[:, :, :shape[2]]
    vol = ctx.volume(m)
    lab, nl = vol.ccl_labels(255)
    olab, onl = O.ccl(m)
    assert nl == onl and np.array_equal(lab, olab)
    assert np.array_equal(vol.area_tensor(), O.mc_area_tensor(m))
    vol.close()


def _segments_from_labels(lab, nl):
    """getCCLSegmentInfo from the oracle's labels with numpy scatter-reductions."""
    nz, ny, nx = lab.shape
    flat = lab.ravel().astype(np.int64)
    z, y, x = (a.ravel() for a in np.indices(lab.shape))
    cnt = np.bincount(flat, minlength=nl)
    mn = np.tile(np.array([nx, ny, nz]), (nl, 1))
    mx = np.zeros((nl, 3), np.int64)
    for a, c in enumerate((x, y, z)):
        np.minimum.at(mn[:, a], flat, c)
        np.maximum.at(mx[:, a], flat, c)
    mx += 1
    dim = np.array([nx, ny, nz])
    at = (mx == dim).any(axis=1) | (mn == 0).any(axis=1)
    return cnt, mn, mx, at


@pytest.mark.parametrize("shape,frac,seed", [((40, 50, 70), 0.55, 7), ((1, 9, 300), 0.5, 8), ((64, 64, 64), None, 0)])
def test_segment_info_matches_oracle_labels(ctx, shape, frac, seed):
    if frac is None:
        m = F.generate_spheres(64, 6.0, 0.60, 0)
    else:
        m = ((np.random.default_rng(seed).random(shape) < frac) * 255).astype(np.uint8)
    vol = ctx.volume(m)
    info = vol.segment_info(255)
    olab, onl = O.ccl(m)
    cnt, mn, mx, at = _segments_from_labels(olab, onl)
    assert len(info) == onl
    for l, r in enumerate(info):
        assert r["labelID"] == l and r["voxelNum"] == cnt[l], l
        assert r["minBB"] == tuple(mn[l]) and r["maxBB"] == tuple(mx[l]), l
        assert r["atBoundary"] == bool(at[l]), l
    # segments that reach no face / not every face, and their volume ratios
    table = O.boundary_table(olab, onl)
    for d in range(7):
        f = F.filterSegmentInfo(info, table, d)
        keep = [l for l in range(1, onl) if not table[d][l]]
        assert [r["labelID"] for r in f] == keep
        assert F.getCCLSegmentRatio(f, m.size) == float(np.float32(cnt[keep].sum()) / np.float32(m.size))
    vol.close()


def test_batch_equals_solo_bit_exact(ctx):
    rng = np.random.default_rng(5)
    vols, params = [], []
    shapes = [(32, 48, 64), (80, 40, 48), (16, 16, 16), (48, 112, 96), (64, 64, 64)]
    for i, (nx, ny, nz) in enumerate(shapes):
        m = F.generate_ellipsoids((nx, ny, nz), 6.0, float(np.exp(rng.uniform(np.log(0.2), np.log(5)))), 0.55, 100 + i)
        vols.append(ctx.volume(m))
        params.append((i % 3) * 2)
    for prec, solver in (("double", 0), ("double", 1), ("mixed", 0), ("mixed", 1)):
        ps = [T.make_params(d, solver, prec, 1e-6, 10000) for d in params]
        batch = T.solve_batch(ctx, vols, ps)
        for v, p, b in zip(vols, ps, batch):
            solo, _ = v.solve(p)
            assert solo.status == 0 and b.status == 0
            assert solo.tau == b.tau and solo.iterations == b.iterations, (prec, solver, v.dims, solo.tau, b.tau)


SHAPES = [(64, 8), (64, 4), (64, 16), (32, 8), (32, 16), (32, 32), (16, 16)]


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_every_tile_shape_gives_identical_results(ctx, precision):
    """Dot products are rounded once from double-double sums, so the Krylov trajectory, tau
    and iteration count do not depend on the tile shape (block shape and planes per tile)."""
    from fast_amd._lib import lib
    m = F.generate_spheres((70, 45, 38), 4.0, 0.45, 21)
    vol = ctx.volume(m)
    ref = {}
    try:
        for (tx, ty) in SHAPES:
            for tz in (3, 16):
                assert lib().fast_ctx_set_tile_shape(ctx.handle, tx, ty, tz) == 0
                for solver in (BICG, CG):
                    res, _ = vol.solve(T.make_params(2, solver, precision, 1e-7, 10000, full_grid=True))
                    got = (res.tau, res.iterations)
                    ref.setdefault(solver, got)
                    assert got == ref[solver], ((tx, ty, tz), solver, got, ref[solver])
    finally:
        lib().fast_ctx_set_tile_shape(ctx.handle, 0, 0, 0)
        vol.close()
