"""Benchmark of the tortuosity hot path on one MI355X.

Workload (the headline configuration): one 256^3 voxel volume of overlapping ellipsoids,
semi-axes (12, 4.2, 4.2) (a = 12, r = 0.35), uniformly random orientations and centres,
filled to solid fraction 0.65 (seed 0).  tau is solved in the x, y and z directions with
BiCGSTAB at tolerance 1e-6.

  * plain FP64 GPU solve (the drop-in): one direction per solve, timed per solve;
  * mixed + batched (the headline form): FP32 stencil/vectors, FP64 reductions and FP64
    iterative refinement, the three directions solved as one batched solve.
A "step" is one batched mixed solve of the three directions.  value = mixed-batched
solves/s.  Inputs are resident in HBM (fast_volume) before the timed region.

Also reported: the FP64 solve rate, tau errors of the mixed form against the FP64 GPU
solve and against the committed CPU-oracle results, the roofline of the dominant kernel
(BiCGSTAB fused pass 1) timed with HIP events on the solve stream, and the CPU oracle
timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "tau solves/s & ms/solve @256^3 (FP64 vs mixed+batched); p50/p99 latency"
HBM_PEAK_GBS = 8000.0
DIRS = (0, 2, 4)          # X_POS, Y_POS, Z_POS
TOL = 1e-6


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def pass1_bytes_per_voxel(vec_bytes: int) -> int:
    # BiCGSTAB fused pass 1 reads r, p_old, v_old, r0 and the stencil code byte, writes p, v
    return 6 * vec_bytes + 1


def cpu_baseline(mask: np.ndarray, gpu_iters: float, threads: int, sample_iters: int) -> dict:
    """Time the CPU oracle (restated reference `--cpu` BiCGSTAB, FP64) on a bounded sample:
    assembly alone, then assembly + `sample_iters` iterations; a full solve is extrapolated
    with the FP64 GPU iteration count (same algorithm)."""
    from oracle import oracle as O
    t = time.perf_counter()
    O.tau(mask, 0, solver=0, tol=TOL, max_iter=0, threads=threads)
    t_setup = time.perf_counter() - t
    t = time.perf_counter()
    O.tau(mask, 0, solver=0, tol=TOL, max_iter=sample_iters, threads=threads)
    t_run = time.perf_counter() - t
    per_iter = max(t_run - t_setup, 1e-9) / sample_iters
    solve_s = t_setup + per_iter * gpu_iters
    return {
        "value": 1.0 / solve_s, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": (f"256^3 config volume, X_POS BiCGSTAB FP64: assembly ({t_setup:.2f} s) + {sample_iters} "
                   f"iterations ({per_iter * 1e3:.0f} ms/iteration) timed, extrapolated to {gpu_iters:.0f} "
                   f"iterations = {solve_s:.1f} s/solve"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-iters", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent volumes shard one queue per GPU with no communication)", file=sys.stderr)
        return 2

    import fast_amd as F
    from fast_amd import tau as T
    from fast_amd._lib import lib

    n = args.size
    log(f"[bench] generating {n}^3 ellipsoid volume")
    mask = F.generate_ellipsoids(n, 12.0, 0.35, 0.65, 0)
    ctx = T.Context(0)
    vol = ctx.volume(mask)
    nvox = mask.size

    # ---- plain FP64 drop-in, one direction per solve ----
    fp64 = {}
    log("[bench] FP64 BiCGSTAB warm-up")
    vol.solve(T.make_params(0, 0, "double", TOL, 10000))
    for d in DIRS:
        t = time.perf_counter()
        res, _ = vol.solve(T.make_params(d, 0, "double", TOL, 10000))
        fp64[d] = (res, time.perf_counter() - t)
        log(f"[bench] FP64 d={d} tau={res.tau:.9f} it={res.iterations} {fp64[d][1] * 1e3:.1f} ms")
    fp64_ms = 1e3 * float(np.mean([fp64[d][1] for d in DIRS]))

    # ---- mixed + batched: the three directions in one batched solve ----
    params = [T.make_params(d, 0, "mixed", TOL, 10000) for d in DIRS]
    vols = [vol, vol, vol]
    for _ in range(args.warmup):
        T.solve_batch(ctx, vols, params)
    log(f"[bench] timing {args.steps} mixed batched steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = T.solve_batch(ctx, vols, params)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = len(DIRS) * args.steps / elapsed
    fp64_rate = 1e3 / fp64_ms

    errs_gpu = [abs(o.tau - fp64[d][0].tau) / fp64[d][0].tau for o, d in zip(out, DIRS)]
    fixture = ROOT / "tests" / "golden" / "config2_oracle.json"
    errs_oracle = None
    if n == 256 and fixture.exists():
        fx = json.loads(fixture.read_text())
        errs_oracle = {}
        for o, d in zip(out, DIRS):
            key = f"BICGSTAB_{['X_POS', 'X_NEG', 'Y_POS', 'Y_NEG', 'Z_POS', 'Z_NEG'][d]}"
            if key in fx:
                errs_oracle[key] = abs(o.tau - fx[key]["tau"]) / fx[key]["tau"]
                errs_oracle["fp64_" + key] = abs(fp64[d][0].tau - fx[key]["tau"]) / fx[key]["tau"]

    # ---- roofline of the dominant kernel (fused pass 1), HIP events on the solve stream ----
    import ctypes as C
    ms = (C.c_double * 4)()
    lib().fast_debug_time_passes(vol.handle, 0, 2, 20, ms)
    mixed_pass = list(ms)
    lib().fast_debug_time_passes(vol.handle, 0, 0, 20, ms)
    fp64_pass = list(ms)
    b32 = pass1_bytes_per_voxel(4) * nvox
    b64 = pass1_bytes_per_voxel(8) * nvox
    ach32 = b32 / (mixed_pass[0] * 1e-3) / 1e9
    ach64 = b64 / (fp64_pass[0] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] CPU oracle sample on {threads} threads")
        cpu = cpu_baseline(mask, float(fp64[0][0].iterations), threads, args.cpu_iters)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vectors/stencil + f64 reductions and refinement",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 ellipsoids a=12 r=0.35 solid 0.65 seed 0; BiCGSTAB tol 1e-6; "
                        "directions x,y,z as one batched mixed-precision solve per step",
            "volume": [n, n, n], "solves_per_step": len(DIRS), "parallelism": "single GPU",
        },
        "fp64_solves_per_s": fp64_rate,
        "fp64_ms_per_solve": fp64_ms,
        "mixed_batched_ms_per_solve": ms_per_step / len(DIRS),
        "speedup_mixed_batched_vs_fp64": value / fp64_rate,
        "tau_fp64": {str(d): fp64[d][0].tau for d in DIRS},
        "iters_fp64": {str(d): int(fp64[d][0].iterations) for d in DIRS},
        "tau_mixed": {str(d): o.tau for o, d in zip(out, DIRS)},
        "iters_mixed": {str(d): int(o.iterations) for o, d in zip(out, DIRS)},
        "refinements_mixed": {str(d): int(o.refinements) for o, d in zip(out, DIRS)},
        "tau_rel_err_mixed_vs_fp64_gpu": max(errs_gpu),
        "tau_rel_err_vs_cpu_oracle": errs_oracle,
        "roofline": {
            "bound": "hbm", "kernel": "k_stencil<float,double,...,BICG1> (fused BiCGSTAB pass 1, mixed inner)",
            "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach32 / HBM_PEAK_GBS,
            "traffic": None, "bytes_per_launch": b32, "avg_launch_ms": mixed_pass[0],
        },
        "roofline_fp64_pass1": {
            "achieved": ach64, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach64 / HBM_PEAK_GBS,
            "bytes_per_launch": b64, "avg_launch_ms": fp64_pass[0],
        },
        "pass_ms_mixed": mixed_pass,
        "pass_ms_fp64": fp64_pass,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    vol.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
