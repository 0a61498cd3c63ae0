"""Exact-arithmetic golden fixtures for the oracle and the GPU path.

tiny_tau.json  -- for small masks (<= 36 voxels) and every direction: the exact solution of
                  the reference's discrete diffusion system (rational arithmetic, dense
                  Gauss-Jordan elimination) and the exact boundary-flux tau.  The system is
                  built from the reference's definitions (fastlib/src/cuda/Diffusion.cuh:54-148:
                  face D = mean of the two cells, half-cell Dirichlet faces on the primary axis,
                  no flux through lateral box faces; VolumeMeasures.cpp:188-280 flux with the
                  cell's own D; h = float(1/maxDim); D = float(1.0), float(0.001)) with exact
                  rationals -- no Krylov iteration, no rounding.  tau_nu is the same solution's
                  tau with useNonUniform (VolumeMeasures.cpp:228-264: pore cells only, half
                  spacing towards a solid neighbour, plane mean over the plane's pore count;
                  "nan" when a plane has no pore cell, as 0/0 in the reference).
ccl_cases.json -- reachable-pore labels from a literal restatement of the reference's span
                  algorithm (VolumeCCL.cu:86-375: toggle spans per x-row, overlap links between
                  neighbouring rows, minimum span-start label, renumbering in row-major root
                  order) and the per-face tables (VolumeCCL.cu:735-818).

Run: python tests/golden/make_golden.py   (pure Python + numpy; a few seconds)
"""
from __future__ import annotations

import json
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def f32(x: float) -> Fraction:
    return Fraction(float(np.float32(x)))


def exact_tau(mask: np.ndarray, d: int):
    nz, ny, nx = mask.shape
    dims = (nx, ny, nz)
    N = nx * ny * nz
    h = f32(1.0 / max(dims))
    D = {0: f32(1.0), 1: f32(0.001)}
    k = d // 2
    sgn = 1 if d % 2 == 0 else -1
    cbeg, cend = (Fraction(1), Fraction(0)) if sgn > 0 else (Fraction(0), Fraction(1))

    def idx(x, y, z):
        return x + nx * (y + ny * z)

    A = [[Fraction(0)] * N for _ in range(N)]
    b = [Fraction(0)] * N
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = idx(x, y, z)
                p = (x, y, z)
                own = 1 if mask[z, y, x] else 0
                for ax in range(3):
                    for s in (1, -1):
                        q = list(p)
                        q[ax] += s
                        if 0 <= q[ax] < dims[ax]:
                            nb = 1 if mask[q[2], q[1], q[0]] else 0
                            a = (D[own] + D[nb]) / 2 * h * h / h
                            A[i][idx(*q)] += a
                            A[i][i] -= a
                        elif ax == k:
                            a = D[own] * h * h / (h / 2)
                            A[i][i] -= a
                            # the rhs takes the begin face, else (`else if`) the end face: with a
                            # single plane along the flow axis only the begin face contributes
                            # (Diffusion.cuh:113-121)
                            if s < 0 or dims[k] > 1:
                                b[i] -= a * (cbeg if s < 0 else cend)
    # Gauss-Jordan on the dense rational system
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(N):
        piv = next(r for r in range(c, N) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        pv = M[c][c]
        M[c] = [v / pv for v in M[c]]
        for r in range(N):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [vr - f * vc for vr, vc in zip(M[r], M[c])]
    sol = [M[i][N] for i in range(N)]
    # flux integral with the cell's own diffusivity
    Np = dims[k]
    sec = [(k + 1) % 3, (k + 2) % 3]
    n = dims[sec[0]] * dims[sec[1]]
    dc_sum = Fraction(0)
    nu_sum = Fraction(0)
    nu_nan = False
    for kk in range(Np):
        s_plane = Fraction(0)
        s_nu = Fraction(0)
        pores = 0
        for j in range(dims[sec[1]]):
            for i in range(dims[sec[0]]):
                p = [0, 0, 0]
                p[k] = kk
                p[sec[0]] = i
                p[sec[1]] = j
                own = 1 if mask[p[2], p[1], p[0]] else 0
                pores += 1 - own
                c = sol[idx(*p)]
                q = list(p)
                q[k] += sgn
                if 0 <= q[k] < dims[k]:
                    s_plane += D[own] * (c - sol[idx(*q)])
                    nb = 1 if mask[q[2], q[1], q[0]] else 0
                    if not own:
                        s_nu += D[own] * (c - sol[idx(*q)]) / (Fraction(1, 2) if nb else 1)
                else:
                    s_plane += D[own] * c / Fraction(1, 2)
                    if not own:
                        s_nu += D[own] * c / Fraction(1, 2)
        dc_sum += s_plane / n
        if pores:
            nu_sum += s_nu / pores
        else:
            nu_nan = True
    dc = dc_sum / Np
    zeros = int((mask == 0).sum())
    phi = Fraction(float(np.float32(zeros / N)))
    # a plane without pore cells gives 0/0; zero flux gives phi/0 (inf, or 0/0 with no pores)
    if nu_nan or (nu_sum == 0 and phi == 0):
        tau_nu = float("nan")
    elif nu_sum == 0:
        tau_nu = float("inf")
    else:
        tau_nu = phi / nu_sum
    if dc == 0:
        return float("inf"), tau_nu, sol   # zero rhs: c = 0, no flux (the reference divides by zero too)
    tau = phi / (dc * Np)
    return tau, tau_nu, sol


def tiny_masks():
    rng = np.random.default_rng(12345)
    cases = []
    cases.append(("all_pore_3x3x3", np.zeros((3, 3, 3), np.uint8)))
    cases.append(("single_solid_centre", np.pad(np.full((1, 1, 1), 255, np.uint8), 1)))
    for t in range(6):
        shape = [(2, 3, 4), (3, 3, 3), (3, 2, 4), (4, 3, 2), (2, 4, 4), (1, 3, 5)][t]
        m = (rng.random(shape) < 0.4).astype(np.uint8) * 255
        cases.append((f"random_{t}_{shape[2]}x{shape[1]}x{shape[0]}", m))
    # values other than 255 count as solid too (Diffusion.cpp:94-96: mask > 0)
    m = np.zeros((2, 3, 3), np.uint8)
    m[0, 1, 1] = 1
    m[1, 2, 0] = 7
    cases.append(("nonbinary_solid_values", m))
    return cases


# ---- literal span CCL (the reference algorithm, sequential) -------------------------------
def span_ccl(mask: np.ndarray, background: int = 255):
    nz, ny, nx = mask.shape
    spans = {}  # (y, z) -> list of [begin, end]
    for z in range(nz):
        for y in range(ny):
            row = mask[z, y]
            prev = background
            is_open = False
            begin = 0
            lst = []
            for x in range(nx):
                v = int(row[x])
                if v != prev:
                    if not is_open:
                        begin = x
                    else:
                        lst.append([begin, x - 1])
                    is_open = not is_open
                prev = v
            if is_open:
                lst.append([begin, nx - 1])
            spans[(y, z)] = lst
    # labels: span-start linear index, union by overlap with rows y-1 and z-1
    parent = {}

    def key(y, z, s):
        return (y, z, s)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    start = {}
    for (y, z), lst in spans.items():
        for s, (bgn, _) in enumerate(lst):
            kk = key(y, z, s)
            parent[kk] = kk
            start[kk] = bgn + nx * (y + ny * z)

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra == rb:
            return
        if start[ra] < start[rb]:
            parent[rb] = ra
        else:
            parent[ra] = rb

    for (y, z), lst in spans.items():
        for s, (b0, e0) in enumerate(lst):
            for (yy, zz) in ((y - 1, z), (y, z - 1)):
                if yy < 0 or zz < 0:
                    continue
                for t, (b1, e1) in enumerate(spans[(yy, zz)]):
                    if b0 <= e1 and b1 <= e0:
                        union(key(y, z, s), key(yy, zz, t))
    # renumber: roots in (z, y, span) order get 1..K
    roots = sorted({find(kk) for kk in parent}, key=lambda r: (r[1], r[0], r[2]))
    newlab = {r: i + 1 for i, r in enumerate(roots)}
    labels = np.zeros(mask.shape, dtype=np.uint32)
    for (y, z), lst in spans.items():
        for s, (b0, e0) in enumerate(lst):
            labels[z, y, b0:e0 + 1] = newlab[find(key(y, z, s))]
    num = len(roots) + 1
    table = np.zeros((7, num), dtype=np.uint8)
    faces = [labels[:, :, -1], labels[:, :, 0], labels[:, -1, :], labels[:, 0, :], labels[-1, :, :], labels[0, :, :]]
    for f, face in enumerate(faces):
        for lab in np.unique(face):
            if lab > 0:
                table[f, lab] = 1
    table[6] = table[:6].min(axis=0)
    return labels, num, table


def ccl_masks():
    rng = np.random.default_rng(777)
    cases = [
        ("all_pore", np.zeros((3, 4, 5), np.uint8)),
        ("all_solid", np.full((3, 4, 5), 255, np.uint8)),
        ("checkerboard", ((np.indices((4, 4, 4)).sum(axis=0) % 2) * 255).astype(np.uint8)),
    ]
    tunnel = np.full((5, 5, 7), 255, np.uint8)
    tunnel[2, 2, :] = 0
    cases.append(("x_tunnel", tunnel))
    for t in range(5):
        shape = [(4, 5, 6), (6, 6, 6), (3, 7, 9), (8, 4, 5), (5, 9, 3)][t]
        frac = [0.3, 0.45, 0.55, 0.5, 0.35][t]
        cases.append((f"random_{t}", ((rng.random(shape) < frac) * 255).astype(np.uint8)))
    # multi-valued rows: spans toggle on every value change (VolumeCCL.cu:104-115)
    mv = rng.choice(np.array([0, 1, 255], np.uint8), size=(3, 4, 8), p=[0.5, 0.2, 0.3])
    cases.append(("multivalued", mv))
    return cases


def main():
    recs = []
    for name, m in tiny_masks():
        for d in range(6):
            tau, tau_nu, sol = exact_tau(m, d)
            exact = f"{tau.numerator}/{tau.denominator}" if isinstance(tau, Fraction) else "inf"
            exact_nu = f"{tau_nu.numerator}/{tau_nu.denominator}" if isinstance(tau_nu, Fraction) else str(tau_nu)
            recs.append({"name": name, "shape_zyx": list(m.shape), "mask": m.ravel().tolist(), "dir": d,
                         "tau": float(tau), "tau_exact": exact, "tau_nu": float(tau_nu), "tau_nu_exact": exact_nu,
                         "conc": [float(v) for v in sol]})
    (HERE / "tiny_tau.json").write_text(json.dumps(recs))
    crecs = []
    for name, m in ccl_masks():
        labels, num, table = span_ccl(m)
        crecs.append({"name": name, "shape_zyx": list(m.shape), "mask": m.ravel().tolist(), "num_labels": num,
                      "labels": labels.ravel().tolist(), "table": table.ravel().tolist()})
    (HERE / "ccl_cases.json").write_text(json.dumps(crecs))
    print(f"{len(recs)} tau fixtures, {len(crecs)} ccl fixtures")


if __name__ == "__main__":
    main()
