"""The oracle's reachable-pore labels against an independent library: scipy.ndimage.label
with 6-connectivity on the pore voxels of {0, 255} masks.  For such masks the reference's
span set is the pore space (spans toggle on every value change along x, VolumeCCL.cu:104-115),
and its labels are numbered 1..K in increasing order of each component's minimum linear voxel
index (roots scanned in (z, y, span) order, VolumeCCL.cu:326-375, 547).  scipy numbers
components in raster order of their first voxel, which for a [z, y, x] array is the same
order, so the two label arrays must be equal element for element."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

SIX = ndimage.generate_binary_structure(3, 1)


def _scipy_labels(mask):
    lab, k = ndimage.label(mask == 0, structure=SIX)
    return lab.astype(np.uint32), k


@pytest.mark.parametrize("seed", range(40))
def test_oracle_ccl_equals_scipy_on_random_masks(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in rng.integers(1, 24, size=3))
    frac = float(rng.uniform(0.2, 0.8))
    m = ((rng.random(shape) < frac) * 255).astype(np.uint8)
    lab, nl = O.ccl(m)
    ref, k = _scipy_labels(m)
    assert nl == k + 1
    assert np.array_equal(lab, ref)


def test_oracle_reachable_equals_scipy_face_sets():
    """Reachable mask along each axis: the pore components touching both faces of the axis."""
    rng = np.random.default_rng(99)
    m = ((rng.random((30, 26, 34)) < 0.35) * 255).astype(np.uint8)
    ref, _ = _scipy_labels(m)
    for d, ax in ((0, 2), (2, 1), (4, 0)):
        lo = set(np.unique(np.take(ref, 0, axis=ax))) - {0}
        hi = set(np.unique(np.take(ref, -1, axis=ax))) - {0}
        want = np.isin(ref, sorted(lo & hi)) & (ref > 0)
        out, _, _ = O.reachable(m, d)
        assert np.array_equal(out == 255, want), d
