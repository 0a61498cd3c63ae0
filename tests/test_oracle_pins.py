"""Pins of the CPU oracle (oracle/tau_oracle.cpp) -- CPU only.

Every oracle function is checked against something that does not restate it: exact-rational
solutions of tiny systems (tests/golden/tiny_tau.json), closed forms of the discrete problem,
symmetry invariants, and a literal restatement of the reference span CCL
(tests/golden/ccl_cases.json).
"""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
D_SOLID = float(np.float32(0.001))


def _tiny():
    return json.loads((GOLD / "tiny_tau.json").read_text())


@pytest.mark.parametrize("solver", [0, 1])
def test_tiny_exact_rational(solver):
    for r in _tiny():
        m = np.array(r["mask"], dtype=np.uint8).reshape(r["shape_zyx"])
        res, c = O.tau(m, r["dir"], solver=solver, tol=1e-13, conc=True)
        if np.isinf(r["tau"]):
            assert np.isinf(res.tau), r["name"]
            continue
        assert res.status == 0, (r["name"], r["dir"])
        assert abs(res.tau - r["tau"]) <= 1e-10 * r["tau"], (r["name"], r["dir"], res.tau, r["tau"])
        np.testing.assert_allclose(c.ravel(), np.array(r["conc"]), rtol=0, atol=1e-11)


@pytest.mark.parametrize("d", range(6))
@pytest.mark.parametrize("solver", [0, 1])
def test_all_pore_tau_is_one(d, solver):
    m = np.zeros((7, 6, 5), np.uint8)
    res, c = O.tau(m, d, solver=solver, tol=1e-12, conc=True)
    assert res.status == 0
    assert abs(res.tau - 1.0) < 1e-10
    # c = 1 - (k + 1/2)/N along the flow direction
    k = d // 2
    n = m.shape[2 - k]
    kk = np.arange(n) + 0.5
    prof = 1 - kk / n if d % 2 == 0 else kk / n
    sl = [np.newaxis] * 3
    sl[2 - k] = slice(None)
    np.testing.assert_allclose(c, np.broadcast_to(prof[tuple(sl)], m.shape), atol=1e-10)


def test_all_solid_tau_zero():
    m = np.full((4, 4, 4), 255, np.uint8)
    res, _ = O.tau(m, 0, solver=0, tol=1e-10)
    assert res.porosity == 0.0 and res.tau == 0.0


@pytest.mark.parametrize("d", [0, 1, 2, 4])
def test_slabs_parallel_to_flow(d):
    # solid planes normal to an axis that is not the flow axis: c is linear in every column
    m = np.zeros((6, 9, 8), np.uint8)
    lateral = (d // 2 + 1) % 3          # an axis other than the flow axis
    idx = [slice(None)] * 3
    idx[2 - lateral] = slice(0, None, 3)
    m[tuple(idx)] = 255
    phi = (m == 0).mean()
    phi_f = float(np.float32(phi))
    expected = phi_f / (phi + D_SOLID * (1 - phi))
    res, _ = O.tau(m, d, solver=0, tol=1e-13)
    assert abs(res.tau - expected) < 1e-9 * expected


@pytest.mark.parametrize("d", [0, 1])
def test_slabs_across_flow_series_resistance(d):
    # layers normal to x: a 1-D problem.  Face conductance (D_k + D_k+1)/2, half cells of 2 D at
    # the Dirichlet faces: constant flux J = 1 / sum of resistances.  The reference flux formula
    # then uses each cell's own D: plane k contributes D_k (c_k - c_next), the last plane
    # D_last c_last / 0.5.
    layers = np.array([0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 0], np.uint8)
    n = len(layers)
    m = np.zeros((3, 4, n), np.uint8)
    m[:, :, layers == 1] = 255
    Dv = [Fraction(1) if v == 0 else Fraction(D_SOLID) for v in layers]
    if d == 1:
        Dv = Dv[::-1]   # flow towards -x: walk the layers from the high-concentration face
    h = Fraction(float(np.float32(1.0 / n)))
    R = 1 / (2 * Dv[0]) + sum(1 / ((Dv[k] + Dv[k + 1]) / 2) for k in range(n - 1)) + 1 / (2 * Dv[-1])
    J = 1 / R
    c = [1 - J / (2 * Dv[0])]
    for k in range(n - 1):
        c.append(c[-1] - J / ((Dv[k] + Dv[k + 1]) / 2))
    flux = [Dv[k] * (c[k] - c[k + 1]) for k in range(n - 1)] + [Dv[-1] * c[-1] * 2]
    dc = sum(flux) / n
    phi = Fraction(float(np.float32((m == 0).mean())))
    tau = float(phi / (dc * n))
    res, _ = O.tau(m, d, solver=1, tol=1e-14)
    assert abs(res.tau - tau) < 1e-9 * tau
    assert h > 0


def _vol(seed, shape=(9, 10, 11), frac=0.45):
    rng = np.random.default_rng(seed)
    return ((rng.random(shape) < frac) * 255).astype(np.uint8)


def test_flip_invariance():
    m = _vol(1)
    t_pos, _ = O.tau(m, 0, tol=1e-12)
    t_neg, _ = O.tau(m[:, :, ::-1], 1, tol=1e-12)
    assert abs(t_pos.tau - t_neg.tau) < 1e-9 * t_pos.tau
    # the own-cell diffusivity in the flux makes X_POS and X_NEG differ in general
    t_neg_same, _ = O.tau(m, 1, tol=1e-12)
    assert abs(t_neg_same.tau - t_pos.tau) > 1e-6 * t_pos.tau


def test_axis_permutation_and_lateral_mirror():
    m = _vol(2, shape=(8, 12, 12))
    tx, _ = O.tau(m, 0, tol=1e-12)
    ty, _ = O.tau(np.ascontiguousarray(m.transpose(0, 2, 1)), 2, tol=1e-12)
    assert abs(tx.tau - ty.tau) < 1e-9 * tx.tau
    tm, _ = O.tau(m[:, ::-1, :], 0, tol=1e-12)
    assert abs(tx.tau - tm.tau) < 1e-9 * tx.tau


def test_cg_and_bicgstab_agree_at_tight_tolerance():
    m = _vol(3, shape=(12, 12, 12), frac=0.5)
    a, _ = O.tau(m, 2, solver=0, tol=1e-12)
    b, _ = O.tau(m, 2, solver=1, tol=1e-12)
    assert abs(a.tau - b.tau) < 1e-9 * a.tau


def test_operator_symmetric_and_row_sums():
    m = _vol(4, shape=(5, 6, 7))
    coef, rhs, x0 = O.system(m, 0)
    # A[i, i+1] (X_POS of i) equals A[i+1, i] (X_NEG of i+1), likewise for y and z, bitwise
    assert np.array_equal(coef[:, :, :-1, 0], coef[:, :, 1:, 1])
    assert np.array_equal(coef[:, :-1, :, 2], coef[:, 1:, :, 3])
    assert np.array_equal(coef[:-1, :, :, 4], coef[1:, :, :, 5])
    # interior rows: diagonal = -(sum of off-diagonals); lateral box faces carry no coefficient
    off = coef[..., :6].sum(axis=-1)
    np.testing.assert_allclose(coef[1:-1, 1:-1, 1:-1, 6], -off[1:-1, 1:-1, 1:-1], rtol=1e-14)
    assert (coef[:, 0, :, 3] == 0).all() and (coef[0, :, :, 5] == 0).all()
    # Dirichlet rows carry the half-cell coefficient in the rhs: b = -2 D h at x = 0
    h = float(np.float32(1.0 / 7))
    Dx = np.where(m[:, :, 0] == 0, 1.0, D_SOLID)
    np.testing.assert_allclose(rhs[:, :, 0], -2 * Dx * h, rtol=1e-12)
    assert (rhs[:, :, 1:] == 0).all()
    # x0 = 1 - k/(N+1)
    np.testing.assert_array_equal(x0[0, 0, :], 1.0 - np.arange(7) / 8.0)


def test_ccl_matches_span_restatement():
    for r in json.loads((GOLD / "ccl_cases.json").read_text()):
        m = np.array(r["mask"], dtype=np.uint8).reshape(r["shape_zyx"])
        lab, nl = O.ccl(m)
        assert nl == r["num_labels"], r["name"]
        assert np.array_equal(lab.ravel(), np.array(r["labels"], dtype=np.uint32)), r["name"]
        t = O.boundary_table(lab, nl)
        assert np.array_equal(t.ravel(), np.array(r["table"], dtype=np.uint8)), r["name"]


def test_reachable_special_cases():
    tunnel = np.full((5, 5, 7), 255, np.uint8)
    tunnel[2, 2, :] = 0
    for d, expect in ((0, True), (2, False), (4, False)):
        out, lab, nl = O.reachable(tunnel, d)
        assert nl == 2
        assert bool(out[2, 2, 3] == 255) == expect
        assert (out[tunnel != 0] == 0).all()
    cb = ((np.indices((4, 4, 4)).sum(axis=0) % 2) * 255).astype(np.uint8)
    lab, nl = O.ccl(cb)
    assert nl == 1 + int((cb == 0).sum())   # every pore voxel isolated


def test_porosity_counts_all_pores_as_float():
    m = _vol(5, shape=(13, 17, 19))
    assert O.count_zero(m) == int((m == 0).sum())
    assert O.porosity(m) == float(np.float32((m == 0).sum() / m.size))


@pytest.mark.parametrize("solver", [0, 1])
def test_nonuniform_flux_exact_rational(solver):
    """useNonUniform (VolumeMeasures.cpp:228-264) on the exact solutions of the tiny systems:
    tau_nu from the rational generator, nan where a plane has no pore cell, inf for zero flux."""
    for r in _tiny():
        m = np.array(r["mask"], dtype=np.uint8).reshape(r["shape_zyx"])
        res, _ = O.tau(m, r["dir"], solver=solver, tol=1e-13, nonuniform=True)
        want = r["tau_nu"]
        if np.isnan(want):
            assert np.isnan(res.tau), (r["name"], r["dir"], res.tau)
        elif np.isinf(want):
            assert np.isinf(res.tau), (r["name"], r["dir"], res.tau)
        else:
            assert abs(res.tau - want) <= 1e-10 * want, (r["name"], r["dir"], res.tau, want)


def test_nonuniform_flux_closed_forms():
    """Parallel slabs: every pore cell carries the same flux 1/N, so the pore-count mean is 1/N
    and tau_nu = float(phi).  A fully solid plane across the flow has no pore cell: 0/0.
    All pore: no solid cells, so tau_nu = tau = 1."""
    m = np.zeros((30, 27, 45), np.uint8)
    m[:, ::3, :] = 255
    phi_f = float(np.float32((m == 0).mean()))
    for solver in (0, 1):
        res, _ = O.tau(m, 0, solver=solver, tol=1e-12, nonuniform=True)
        assert abs(res.tau - phi_f) < 1e-9 * phi_f
    blocked = np.zeros((8, 6, 10), np.uint8)
    blocked[:, :, 4] = 255
    res, _ = O.tau(blocked, 0, solver=1, tol=1e-12, nonuniform=True)
    assert np.isnan(res.tau)
    res, _ = O.tau(np.zeros((8, 6, 10), np.uint8), 2, solver=1, tol=1e-12, nonuniform=True)
    assert abs(res.tau - 1.0) < 1e-12


def test_nonuniform_flip_invariance():
    m = _vol(4)
    t_pos, _ = O.tau(m, 0, tol=1e-12, nonuniform=True)
    t_neg, _ = O.tau(m[:, :, ::-1], 1, tol=1e-12, nonuniform=True)
    assert abs(t_pos.tau - t_neg.tau) < 1e-9 * t_pos.tau
    t_uni, _ = O.tau(m, 0, tol=1e-12)
    assert abs(t_uni.tau - t_pos.tau) > 1e-3 * t_pos.tau   # the two flux modes differ
