"""Write the FP64 reference tau of the serving workload's request pool to
tests/golden/serving_pool_fp64.json (runs on the GPU).

Every (edge, volume, direction) of fast_amd.workloads.request_pool(per_edge) is solved with
FP64 CG to 1e-11 through the library: the FP64 path runs the oracle's recurrences and
reproduces the oracle's tau bit for bit (tests/test_gpu_config4.py checks all 96 volumes of
the batched workload), and the CPU oracle itself would need days for the 384^3 volumes.
Each record carries the SHA-256 of its mask.  bench.py's serving leg reads the file to
report every served request's tau error.

  python tests/golden/make_serving_reference.py [per_edge] [out.json]   (default 16, this file)

On a remote GPU give an output path that is copied back; each line printed also carries the
key, tau (repr), iterations and time of one solve.
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fast_amd import tau as T  # noqa: E402
from fast_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent / "serving_pool_fp64.json"


def main() -> int:
    per_edge = int(sys.argv[1]) if len(sys.argv) > 1 else 16
    out = Path(sys.argv[2]) if len(sys.argv) > 2 else OUT
    recs = json.loads(out.read_text()) if out.exists() else {}
    ctx = T.Context(0)
    t0 = time.time()
    for e in W.REQUEST_EDGES:
        for j in range(per_edge):
            v = W.request_volume(e, j)
            digest = hashlib.sha256(v.mask.tobytes()).hexdigest()
            vol = None
            for d in (0, 2, 4):
                key = f"{e}_{j}_{d}"
                if key in recs and recs[key]["sha256"] == digest:
                    continue
                vol = vol or ctx.volume(v.mask)
                res, _ = vol.solve(T.make_params(d, 1, "double", 1e-11, 60000))
                recs[key] = {"sha256": digest, "edge": e, "volume": j, "dir": d, "r": v.r, "solid": v.solid,
                             "tau": res.tau, "iterations": res.iterations, "status": res.status,
                             "solve_ms": res.solve_ms}
                print(f"{key}: tau {res.tau!r} ({res.iterations} it, {res.solve_ms:.0f} ms) "
                      f"[{time.time() - t0:.0f} s]", flush=True)
            if vol is not None:
                vol.close()
            out.write_text(json.dumps(recs, indent=1, sort_keys=True))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
