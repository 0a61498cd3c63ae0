"""`fast alpha` on the GPU: the marching-cubes area kernels against the CPU oracle.

Both compute every cube in the reference's float arithmetic without contraction and sum the
per-cube double areas in double-double, so the totals must be equal bit for bit: for the
volume's bytes (getReactiveAreaDensity) and for the six boundary-connected volumes
(getReactiveAreaDensityTensor), on shapes that are ragged against the 64 x 4 x 16 block
tile, one voxel thin in some axis, or a single voxel, for binary and arbitrary bytes and for
other isovalues."""
import math

import numpy as np
import pytest

import fast_amd as F
from fast_amd._lib import FastError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_mask(shape, frac, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random(shape) < frac, 255, 0).astype(np.uint8)


# (nz, ny, nx)
SHAPES = [
    ("single_1x1x1", lambda: np.full((1, 1, 1), 255, np.uint8)),
    ("tile_64x4x16", lambda: rand_mask((16, 4, 64), 0.5, 1)),
    ("tile_plus_1_65x5x17", lambda: rand_mask((17, 5, 65), 0.45, 2)),
    ("odd_37x23x19", lambda: rand_mask((19, 23, 37), 0.45, 3)),
    ("tall_1x5x70", lambda: rand_mask((70, 5, 1), 0.5, 4)),
    ("thin_67x3x1", lambda: rand_mask((1, 3, 67), 0.5, 5)),
    ("wide_131x17x11", lambda: rand_mask((11, 17, 131), 0.4, 6)),
    ("spheres_64", lambda: F.generate_spheres(64, 5.0, 0.55, 7)),
    ("ellipsoids_96", lambda: F.generate_ellipsoids(96, 8.0, 0.3, 0.6, 8)),
]


@pytest.mark.parametrize("name,make", SHAPES, ids=[s[0] for s in SHAPES])
def test_area_and_tensor_equal_the_oracle(ctx, name, make):
    m = make()
    vol = ctx.volume(m)
    try:
        a = vol.area()
        ref = O.mc_area(m)
        t = vol.area_tensor()
        tref = O.mc_area_tensor(m)
        print(f"{name}: area {a!r} oracle {ref!r}; tensor {t.tolist()}")
        assert a == ref, (name, a, ref)
        assert np.array_equal(t, tref), (name, t, tref)
    finally:
        vol.close()


def test_arbitrary_bytes_and_other_isovalues(ctx):
    """Corner value = (sum of the 8 bytes) / 2040 for any bytes; isovalues 0 to 1."""
    rng = np.random.default_rng(11)
    for shape in [(23, 17, 70), (9, 33, 5)]:
        m = rng.integers(0, 256, size=shape, dtype=np.uint8)
        vol = ctx.volume(m)
        for iso in (0.0, 0.25, 0.5, 0.75, 1.0):
            assert vol.area(iso) == O.mc_area(m, iso), (shape, iso)
        vol.close()
    m = rand_mask((40, 30, 50), 0.5, 12)
    vol = ctx.volume(m)
    for iso in (0.3, 0.7):
        assert vol.area(iso) == O.mc_area(m, iso)
        assert np.array_equal(vol.area_tensor(iso), O.mc_area_tensor(m, iso))
    vol.close()


def test_closed_forms_on_the_device(ctx):
    """A half-space is one plane of area 1 on any axis; a solid box is the chamfered box."""
    m = np.zeros((40, 33, 70), np.uint8)
    m[:, :, :31] = 255
    assert abs(F.getReactiveAreaDensity(m, ctx=ctx) - 1.0) < 1e-6
    m = np.zeros((40, 33, 70), np.uint8)
    m[:17] = 255
    assert abs(F.getReactiveAreaDensity(m, ctx=ctx) - 1.0) < 1e-6
    n = 48
    m = np.zeros((n, n, n), np.uint8)
    a, b, c = 9, 14, 6
    m[10:10 + c, 11:11 + b, 12:12 + a] = 255
    A, B, C = a - 2, b - 2, c - 2
    box = 2 * (A * B + B * C + A * C) + 4 * math.sqrt(2) * (A + B + C) + 4 * math.sqrt(3)
    assert abs(F.getReactiveAreaDensity(m, ctx=ctx) * n * n / box - 1) < 2e-6
    m = np.zeros((20, 16, 12), np.uint8)
    m[5:9] = 255
    np.testing.assert_allclose(F.getReactiveAreaDensityTensor(m, ctx=ctx), [2, 2, 2, 2, 1, 1], rtol=0, atol=2e-6)


def test_bench_volume_256(ctx):
    """The 256^3 ellipsoid volume of the tau benchmark: area and tensor equal the oracle."""
    m = F.generate_ellipsoids(256, 12.0, 0.35, 0.65, 0)
    vol = ctx.volume(m)
    try:
        a = vol.area()
        t = vol.area_tensor()
        assert a == O.mc_area(m)
        assert np.array_equal(t, O.mc_area_tensor(m))
        print(f"256^3 ellipsoids: area {a!r} tensor {t.tolist()}")
    finally:
        vol.close()


def test_non_finite_isovalue_refused(ctx):
    vol = ctx.volume(rand_mask((8, 8, 8), 0.5, 1))
    with pytest.raises(FastError):
        vol.area(float("nan"))
    with pytest.raises(FastError):
        vol.area_tensor(float("inf"))
    vol.close()


MESH_SHAPES = [
    ("box_in_pore", lambda: np.pad(np.full((3, 4, 5), 255, np.uint8), ((2, 3), (1, 2), (3, 1)))),
    ("odd_19x23x37", lambda: rand_mask((19, 23, 37), 0.45, 3)),
    ("thin_67x3x1", lambda: rand_mask((1, 3, 67), 0.5, 5)),
    ("spheres_40", lambda: F.generate_spheres(40, 4.0, 0.5, 9)),
]


def _tri_areas(v):
    p = v[:, :3].astype(np.float64).reshape(-1, 3, 3)
    return np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])


@pytest.mark.parametrize("name,make", MESH_SHAPES, ids=[s[0] for s in MESH_SHAPES])
def test_mesh_equals_the_oracle_and_is_closed(ctx, name, make):
    """The --mesh triangles equal the oracle's bit for bit (NaN normals of degenerate triangles
    included); the surface of the zero-padded volume is closed (vector area 0); for a cube the
    triangles' total area equals the area kernel's on the padded volume."""
    m = make()
    vol = ctx.volume(m)
    try:
        got = vol.mesh()
        ref = O.mc_mesh(m)
        assert got.shape == ref.shape and got.shape[0] % 3 == 0, (name, got.shape, ref.shape)
        assert np.array_equal(got, ref, equal_nan=True), name
        cr = _tri_areas(got)
        total = 0.5 * np.linalg.norm(cr, axis=1).sum()
        assert total > 0
        assert np.abs(cr.sum(axis=0)).max() <= 1e-5 * total, name
        if len(set(m.shape)) == 1:
            # same triangles; the mesh's coordinates are centred (|c| up to 0.5), so each one
            # carries a rounding of up to 3e-8, a few 1e-6 of a voxel edge at 40^3
            padded = np.pad(m, 1)
            pvol = ctx.volume(padded)
            assert total == pytest.approx(pvol.area(), rel=1e-5)
            pvol.close()
    finally:
        vol.close()


def test_mesh_of_empty_volume(ctx):
    vol = ctx.volume(np.zeros((5, 6, 7), np.uint8))
    assert vol.mesh().shape == (0, 6)
    vol.close()
