"""The `fast alpha` oracle (oracle/alpha_oracle.cpp) pinned without the GPU.

The triangle table is pinned by mesh invariants that a wrong entry breaks (every cube
configuration cuts exactly its sign-change edges; the mesh of any zero-padded volume is a
closed, consistently oriented surface) and, where the reference tree is present, entry by
entry against the reference's table (fastlib/src/cuda/MCTable.cuh).  The geometry
(interpolation, positions, triangle areas, sums) is pinned by closed forms: planes through a
half-space or a slab have area 1 each in unit-cube coordinates on any axis and any extents,
a solid box gives the chamfered box of `_box_area`, the per-face tensor of a split volume
counts the planes each face's pore region touches, and a voxelised sphere approaches
4 pi r^2."""
import math
import os
import re
from collections import Counter

import numpy as np
import pytest

from oracle import oracle as O

EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
REF_TABLE = "/root/reference/fastlib/src/cuda/MCTable.cuh"


def test_table_cuts_exactly_the_sign_change_edges():
    tab = O.mc_table()
    assert tab[0] == [] and tab[255] == []
    for c, tris in enumerate(tab):
        assert len(tris) % 3 == 0 and len(tris) <= 15
        need = {e for e, (a, b) in enumerate(EDGES) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert set(tris) == need, c
        for i in range(0, len(tris), 3):
            assert len(set(tris[i:i + 3])) == 3, (c, tris[i:i + 3])


def _edge_id(x, y, z, e):
    a, b = EDGES[e]
    pa = (x + CORNERS[a][0], y + CORNERS[a][1], z + CORNERS[a][2])
    pb = (x + CORNERS[b][0], y + CORNERS[b][1], z + CORNERS[b][2])
    return min(pa, pb), max(pa, pb)


@pytest.mark.parametrize("seed", range(8))
def test_table_mesh_is_closed_and_oriented(seed):
    """Zero-padded random 6^3 corner grids: every directed mesh edge occurs once and its
    reverse once (a closed 2-manifold with consistent orientation)."""
    tab = O.mc_table()
    rng = np.random.default_rng(seed)
    n = 6
    bits = np.zeros((n + 2,) * 3, bool)
    bits[1:-1, 1:-1, 1:-1] = rng.random((n,) * 3) < rng.uniform(0.25, 0.75)
    directed = Counter()
    for z in range(n + 1):
        for y in range(n + 1):
            for x in range(n + 1):
                c = sum(int(bits[x + dx, y + dy, z + dz]) << m for m, (dx, dy, dz) in enumerate(CORNERS))
                t = tab[c]
                for i in range(0, len(t), 3):
                    g = [_edge_id(x, y, z, e) for e in t[i:i + 3]]
                    for k in range(3):
                        directed[(g[k], g[(k + 1) % 3])] += 1
    assert directed
    for (a, b), cnt in directed.items():
        assert cnt == 1 and directed.get((b, a), 0) == 1


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference tree not present")
def test_table_equals_the_references_entry_by_entry():
    src = open(REF_TABLE).read()
    body = src[src.index("triTable"):src.index("numVertsTable")]
    rows = []
    for r in re.findall(r"\{([^{}]*)\}", body):
        vals = [v.strip() for v in r.split(",") if v.strip()]
        if len(vals) == 16:
            rows.append([int(v) for v in vals if v != "X"])
    assert len(rows) == 256
    assert rows == O.mc_table()


@pytest.mark.parametrize("shape", [(20, 16, 12), (9, 31, 5), (3, 2, 40)])
def test_half_spaces_and_slabs_are_unit_planes(shape):
    nz, ny, nx = shape
    for axis, n in ((0, nz), (1, ny), (2, nx)):
        if n < 4:
            continue
        for k in (1, n // 2, n - 1):
            m = np.zeros(shape, np.uint8)
            idx = [slice(None)] * 3
            idx[axis] = slice(0, k)
            m[tuple(idx)] = 255
            assert abs(O.mc_area(m) - 1.0) < 1e-6, (shape, axis, k)
            assert abs(O.mc_area(255 - m) - 1.0) < 1e-6, (shape, axis, k)
        if n >= 6:
            m = np.zeros(shape, np.uint8)
            idx[axis] = slice(2, n - 2)
            m[tuple(idx)] = 255
            assert abs(O.mc_area(m) - 2.0) < 1e-6, (shape, axis)


def test_no_surface_for_uniform_volumes_and_isolated_voxels():
    """A corner is the mean of its 8 voxels, so a lone solid voxel (corners 1/8) or a lone
    pore voxel in solid (corners 7/8) never reaches the 0.5 level."""
    for v in (0, 255):
        assert O.mc_area(np.full((7, 8, 9), v, np.uint8)) == 0.0
    m = np.zeros((9, 9, 9), np.uint8)
    m[4, 4, 4] = 255
    assert O.mc_area(m) == 0.0
    assert O.mc_area(255 - m) == 0.0


def _box_area(a, b, c):
    """Marching-cubes surface of a solid a x b x c voxel box in pore (voxel units): the
    corners on the box's faces read 1/2 and lie on the surface, the corners on its edges
    (1/4) and vertices (1/8) lie outside, so the faces shrink by one voxel per side and are
    joined by chamfer strips of width sqrt(2) and corner triangles of side sqrt(2)."""
    A, B, C = a - 2, b - 2, c - 2
    return 2 * (A * B + B * C + A * C) + 4 * math.sqrt(2) * (A + B + C) + 8 * (math.sqrt(3) / 2)


@pytest.mark.parametrize("box", [(2, 2, 2), (5, 7, 9), (3, 10, 4), (12, 12, 12)])
def test_solid_box_is_a_chamfered_box(box):
    a, b, c = box
    n = 32
    m = np.zeros((n, n, n), np.uint8)
    m[9:9 + c, 7:7 + b, 5:5 + a] = 255
    got = O.mc_area(m) * n * n
    assert abs(got / _box_area(a, b, c) - 1) < 2e-6, (box, got)


def test_tensor_counts_the_planes_each_face_region_touches():
    """A solid slab across z splits the pore space in two: faces x+-, y+- touch both regions
    (two planes), z- only the lower and z+ only the upper one (one plane each).  A solid
    block in the upper region holding a sealed pore cavity adds its outer surface to z+ and
    to the basic area; the cavity adds its own surface to the basic area only."""
    m = np.zeros((20, 16, 12), np.uint8)
    m[5:9] = 255
    np.testing.assert_allclose(O.mc_area_tensor(m), [2, 2, 2, 2, 1, 1], rtol=0, atol=2e-6)
    m[12:19, 3:13, 2:10] = 255
    m[14:17, 6:10, 4:8] = 0
    t = O.mc_area_tensor(m)
    cavity = np.full(m.shape, 255, np.uint8)
    cavity[14:17, 6:10, 4:8] = 0
    assert t[5] == pytest.approx(1.0, abs=2e-6)
    assert t[4] > 1.5
    assert O.mc_area(m) == pytest.approx(t[4] + t[5] + O.mc_area(cavity), rel=1e-6)


@pytest.mark.parametrize("r", [8, 14, 20])
def test_sphere_area_near_4_pi_r2(r):
    n = 2 * r + 8
    z, y, x = np.indices((n, n, n)) + 0.5
    c = n / 2
    m = np.where((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2 < r * r, 255, 0).astype(np.uint8)
    ratio = O.mc_area(m) * n * n / (4 * math.pi * r * r)
    assert 1.0 < ratio < 1.02, ratio


def test_cube_area_sums_to_total():
    rng = np.random.default_rng(5)
    m = np.where(rng.random((5, 6, 7)) < 0.5, 255, 0).astype(np.uint8)
    per = [O.mc_cube_area(m, x, y, z) for z in range(5) for y in range(6) for x in range(7)]
    assert math.fsum(per) == pytest.approx(O.mc_area(m), rel=1e-15)


def test_segment_filter_and_ratio_host_logic():
    """filterSegmentInfo keeps, with the inverted mask, the labels >= 1 that are not on the
    given face row (row 6: not on every face); label 0 never; without inversion, the labels on
    the row.  getCCLSegmentRatio is float(sum of voxels) / float(N)."""
    import fast_amd as F
    # pore regions: a slab open to x- only, a cavity sealed off, and a channel through x
    m = np.full((6, 7, 9), 255, np.uint8)
    m[1:3, 1:3, 0:4] = 0          # touches x- (and no other face)
    m[4, 4, 4] = 0                # sealed
    m[4, 1, :] = 0                # x- to x+
    lab, nl = O.ccl(m)
    table = O.boundary_table(lab, nl)
    counts = np.bincount(lab.ravel().astype(np.int64), minlength=nl)
    sinfo = [{"labelID": l, "voxelNum": int(counts[l])} for l in range(nl)]
    for d in range(7):
        kept = [s["labelID"] for s in F.filterSegmentInfo(sinfo, table, d)]
        assert kept == [l for l in range(1, nl) if not table[d][l]]
        assert [s["labelID"] for s in F.filterSegmentInfo(sinfo, table, d, invertMask=False)] == \
            [l for l in range(nl) if table[d][l]]
    sealed = lab[4, 4, 4]
    none = F.filterSegmentInfo(sinfo, table, 6)
    assert sealed in [s["labelID"] for s in none] and lab[4, 1, 0] in [s["labelID"] for s in none]
    assert F.getCCLSegmentRatio(none, m.size) == float(np.float32(sum(s["voxelNum"] for s in none)) /
                                                       np.float32(m.size))
