"""Write FP64 CPU-oracle results for the full-size workloads to tests/golden/*.json.

The oracle (oracle/tau_oracle.cpp, a restatement of the reference `--cpu` solver) needs
minutes per solve at 256^3, too slow to run inside the test suite, so its results are
committed here and the GPU parity tests compare against them.  Each record stores the
SHA-256 of the generated mask so a test can prove it solved the same volume.

  python tests/golden/make_oracle_fixtures.py config2             # 256^3 ellipsoids, CG + BiCGSTAB, X/Y/Z
  python tests/golden/make_oracle_fixtures.py config2_converged   # same volume, CG to 1e-11
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import fast_amd as F  # noqa: E402  (generator only; no GPU use)
from oracle import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def config2_mask():
    # 256^3, semi-axes (12, 4.2, 4.2), uniform orientations, solid fraction 0.65, seed 0
    return F.generate_ellipsoids(256, 12.0, 0.35, 0.65, 0)


def run(name: str, mask, cases, tol=1e-6, threads=0):
    out_path = HERE / f"{name}_oracle.json"
    recs = json.loads(out_path.read_text()) if out_path.exists() else {}
    digest = hashlib.sha256(mask.tobytes()).hexdigest()
    for solver, d in cases:
        key = f"{['BICGSTAB', 'CG'][solver]}_{['X_POS', 'X_NEG', 'Y_POS', 'Y_NEG', 'Z_POS', 'Z_NEG'][d]}"
        if key in recs and recs[key]["sha256"] == digest:
            continue
        t = time.time()
        r, _ = O.tau(mask, d, solver=solver, precision=0, tol=tol, max_iter=10000, threads=threads)
        recs[key] = {
            "sha256": digest, "dims": list(mask.shape[::-1]), "solver": solver, "dir": d, "tol": tol,
            "tau": r.tau, "porosity": r.porosity, "iterations": r.iterations, "error": r.error, "dc": r.dc,
            "status": r.status, "restarts": r.restarts, "seconds": time.time() - t,
        }
        out_path.write_text(json.dumps(recs, indent=1, sort_keys=True))
        print(key, recs[key], flush=True)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "config2"
    if which == "config2":
        run("config2", config2_mask(), [(0, 0), (1, 0), (0, 2), (1, 2), (0, 4), (1, 4)])
    elif which == "config2_converged":
        # the converged tau (CG to 1e-11): the yardstick for tau accuracy at 256^3, where
        # BiCGSTAB stopped at 1e-6 is still up to ~1e-4 away from it
        run("config2_converged", config2_mask(), [(1, 0), (1, 2), (1, 4)], tol=1e-11)
