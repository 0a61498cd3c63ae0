"""Benchmark of the tortuosity hot path on one MI355X.

Workload (the headline configuration): one 256^3 voxel volume of overlapping ellipsoids,
semi-axes (12, 4.2, 4.2) (a = 12, r = 0.35), uniformly random orientations and centres,
filled to solid fraction 0.65 (seed 0); tau in the x, y and z directions at tolerance 1e-6.

  * plain FP64 GPU solve (the drop-in): BiCGSTAB and CG, one direction per solve;
  * mixed + batched (the headline form): Jacobi-preconditioned CG on the red-black Schur
    complement of the same linear system, FP32 vectors in HBM, FP64 arithmetic and dot
    products, the solution accumulated in FP64 with reliable updates against the FP64 true
    residual of the full system; the three directions solved as one batched solve;
  * the same mixed CG on the full grid (for comparison).
A "step" is one batched mixed solve of the three directions; value = mixed-batched solves/s.
Inputs are resident in HBM (fast_volume) before the timed region.

tau accuracy is stated against the converged FP64 tau (CG to 1e-11): at 256^3 the FP64
BiCGSTAB tau at tolerance 1e-6 is itself up to ~1e-4 away from it.  Also reported: the
roofline of the dominant kernel (timed with HIP events on the solve stream; the rocprofv3
summary under profiles/ agrees) and the CPU oracle timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
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
DIRNAMES = ["X_POS", "X_NEG", "Y_POS", "Y_NEG", "Z_POS", "Z_NEG"]
TOL = 1e-6
BICG, CG = 0, 1


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def converged_taus(vol, T) -> tuple:
    fx = ROOT / "tests" / "golden" / "config2_converged_oracle.json"
    out, src = {}, "GPU FP64 CG tol 1e-11 (bit-identical to the CPU oracle's recurrences)"
    if fx.exists():
        recs = json.loads(fx.read_text())
        if all(f"CG_{DIRNAMES[d]}" in recs for d in DIRS):
            return {d: recs[f"CG_{DIRNAMES[d]}"]["tau"] for d in DIRS}, "CPU oracle CG tol 1e-11"
    for d in DIRS:
        res, _ = vol.solve(T.make_params(d, CG, "double", 1e-11, 20000))
        out[d] = res.tau
    return out, src


def cpu_baseline(mask: np.ndarray, gpu_iters: float, threads: int, sample_iters: int) -> dict:
    """Time the CPU oracle (the restated reference `--cpu` CG, FP64, the CLI's default solver)
    on a bounded sample: assembly alone, then assembly + `sample_iters` iterations; a full
    solve is extrapolated with the FP64 GPU CG iteration count (same recurrences)."""
    from oracle import oracle as O
    t = time.perf_counter()
    O.tau(mask, 0, solver=CG, tol=TOL, max_iter=0, threads=threads)
    t_setup = time.perf_counter() - t
    t = time.perf_counter()
    O.tau(mask, 0, solver=CG, tol=TOL, max_iter=sample_iters, threads=threads)
    t_run = time.perf_counter() - t
    per_iter = max(t_run - t_setup, 1e-9) / sample_iters
    solve_s = t_setup + per_iter * gpu_iters
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    t = time.perf_counter()
    O.mc_area(mask)
    t_area = time.perf_counter() - t
    return {
        "value": 1.0 / solve_s, "unit": "solves/s", "cores": threads, "kind": "port",
        "caveat": ("overstates the reference's CPU time: the oracle sums every dot product in double-double "
                   "(slower than Eigen's plain sums), and a full solve is extrapolated from assembly and a "
                   "bounded number of iterations"),
        "alpha_area_s": t_area,
        "cpu_model": model, "nproc": os.cpu_count(),
        "sample": (f"256^3 volume, X_POS, CG FP64 (oracle restating the reference --cpu path): assembly "
                   f"{t_setup:.2f} s + {sample_iters} iterations at {per_iter * 1e3:.0f} ms each, extrapolated to "
                   f"{gpu_iters:.0f} iterations = {solve_s:.1f} s per solve"),
    }


SERVE_TOL = 1e-9   # FP64 true relative residual of served requests (tau within 1e-5 of converged)


def serving(ctx, T, W, n_req: int, loads=(0.5, 0.8, 1.1), per_edge: int = 16, rate: float = 0.0,
            check_tau: bool = True) -> dict:
    """Continuous-batching server under Poisson arrivals.  Requests: cubic edge from
    {64,128,192,256,384} (weights .3/.3/.2/.15/.05), direction uniform in x/y/z, ellipsoid
    volumes (r log-uniform on [0.1, 10], solid 0.5..0.75) from a pool of `per_edge` per edge;
    mixed CG (red-black) to an FP64 true residual of SERVE_TOL.  The saturated rate is measured
    with a burst (or given as `rate`); then `n_req` requests arrive at each offered load
    (fraction of that rate).  Latency = completion - submission.  Afterwards every (volume,
    direction) pair served is solved in FP64 CG to 1e-11 and each request's tau error against
    it is reported."""
    import threading
    pool = W.request_pool(per_edge, threads=16)
    mk = lambda d: T.make_params(d, CG, "mixed", SERVE_TOL, 40000)
    out = {"requests_per_load": n_req, "volume_pool": f"{per_edge} volumes per edge, {len(pool)} in all",
           "tolerance": SERVE_TOL, "solver": "mixed CG (red-black Schur complement)"}
    served = []   # (load index, edge, volume, dir, latency ms, tau, status)
    with T.Server(ctx, max_active=64, iters_per_round=16) as srv:
        if rate > 0:
            sat = rate
            out["saturated_solves_per_s"] = sat
            out["saturated_source"] = "given"
        else:
            burst = W.request_stream(min(max(n_req // 2, 24), 200), 1e9, per_edge=per_edge, seed=11)
            t0 = time.perf_counter()
            for q in burst:
                srv.submit(pool[(q.edge, q.volume)], mk(q.dir))
            for _ in burst:
                srv.wait(-1)
            sat = len(burst) / (time.perf_counter() - t0)
            out["saturated_solves_per_s"] = sat
            out["saturated_source"] = f"burst of {len(burst)} requests"
            log(f"[bench] serving: saturated {sat:.1f} solves/s over a burst of {len(burst)}")
        runs = []
        for li, f in enumerate(loads):
            reqs = W.request_stream(n_req, f * sat, per_edge=per_edge, seed=100 + int(round(10 * f)))
            tickets = {}
            lock = threading.Lock()
            start = time.perf_counter()

            def feed():
                for q in reqs:
                    dt = start + q.arrival - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    t = srv.submit(pool[(q.edge, q.volume)], mk(q.dir))
                    with lock:
                        tickets[t] = q

            th = threading.Thread(target=feed)
            th.start()
            done = []
            for k in range(len(reqs)):
                done.append(srv.wait(-1))
                if (k + 1) % 250 == 0:
                    log(f"[bench] serving load {f:.2f}: {k + 1}/{len(reqs)} done")
            th.join()
            wall = time.perf_counter() - start
            lat = []
            for t, res, ms in done:
                q = tickets[t]
                served.append((li, q.edge, q.volume, q.dir, ms, res.tau, res.status))
                lat.append(ms)
            byedge = {}
            for e in W.REQUEST_EDGES:
                le = [x[4] for x in served if x[0] == li and x[1] == e]
                if le:
                    byedge[str(e)] = {"n": len(le), "p50_ms": float(np.percentile(le, 50)),
                                      "p99_ms": float(np.percentile(le, 99)), "max_ms": float(np.max(le))}
            runs.append({"offered_fraction": f, "offered_solves_per_s": f * sat, "achieved_solves_per_s": n_req / wall,
                         "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                         "mean_ms": float(np.mean(lat)), "by_edge": byedge,
                         "failed": int(sum(1 for x in served if x[0] == li and x[6] != 0))})
            log(f"[bench] serving load {f:.2f}: p50 {runs[-1]['p50_ms']:.0f} ms p99 {runs[-1]['p99_ms']:.0f} ms "
                f"{runs[-1]['achieved_solves_per_s']:.1f} solves/s")
        out["loads"] = runs
    log("[bench] serving: requests " + json.dumps([list(x) for x in served]))
    if check_tau:
        # the FP64 reference of every (volume, direction) pair served, then each request's error:
        # from tests/golden/serving_pool_fp64.json (FP64 CG to 1e-11 through this library, with
        # the masks' SHA-256) where it covers the pair, else solved here
        import hashlib
        fx = ROOT / "tests" / "golden" / "serving_pool_fp64.json"
        recs = json.loads(fx.read_text()) if fx.exists() else {}
        conv, solved = {}, 0
        keys = sorted({(x[1], x[2], x[3]) for x in served})
        log(f"[bench] serving: FP64 reference tau of the {len(keys)} pairs served")
        for key in keys:
            rec = recs.get(f"{key[0]}_{key[1]}_{key[2]}")
            if rec and rec["sha256"] == hashlib.sha256(pool[key[:2]].tobytes()).hexdigest():
                conv[key] = rec["tau"]
                continue
            vol = ctx.volume(pool[key[:2]])
            res, _ = vol.solve(T.make_params(key[2], CG, "double", 1e-11, 60000))
            vol.close()
            conv[key] = res.tau
            solved += 1
            log(f"[bench] serving: FP64 reference {key}: tau {res.tau:.9f} ({solved} solved here)")
        errs = np.array([abs(x[5] - conv[x[1:4]]) / conv[x[1:4]] for x in served])
        for li, run in enumerate(out["loads"]):
            e = np.array([errs[k] for k, x in enumerate(served) if x[0] == li])
            run["tau_rel_err_max"] = float(e.max())
            run["tau_rel_err_p99"] = float(np.percentile(e, 99))
            run["tau_within_1e-5"] = int((e <= 1e-5).sum())
        out["tau_reference"] = (f"FP64 CG to 1e-11 of each of the {len(conv)} (volume, direction) pairs served "
                                f"({len(conv) - solved} from tests/golden/serving_pool_fp64.json)")
        out["tau_rel_err_max"] = float(errs.max())
        log(f"[bench] serving: max tau rel err {errs.max():.2e} over {len(served)} requests")
    return out


def hbm_ceiling(ctx, lib, gib: float = 2.0, iters: int = 10) -> dict:
    """The measured HBM ceiling: hand-written streaming kernels (16-byte accesses) over three
    buffers of `gib` GiB each, far larger than the 256 MB Infinity Cache, timed with HIP
    events on the solve stream: triad c = a + s b (2 reads, 1 write), the same on a striding
    grid of 8192 blocks, copy b = a, fill a = s."""
    ms = (C.c_double * 4)()
    nbytes = int(gib * 2 ** 30) // 16 * 16
    rc = lib().fast_debug_stream(ctx.handle, nbytes, iters, ms)
    if rc != 0:
        return {"error": f"fast_debug_stream returned {rc}"}
    moved = {"triad": 3 * nbytes, "triad_8192_blocks": 3 * nbytes, "copy": 2 * nbytes, "fill": nbytes}
    out = {"buffer_bytes": nbytes, "timing": f"HIP events around {iters} back-to-back launches"}
    for i, (k, b) in enumerate(moved.items()):
        out[k] = {"ms": ms[i], "GBs": b / (ms[i] * 1e-3) / 1e9, "frac_of_peak": b / (ms[i] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def measured_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` on the n^3 workload of the bench (256: ellipsoids, 512:
    needles), from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE counters scaled by what the
    same counters read for the triad streaming kernel; None when not measured."""
    path = ROOT / "profiles" / f"rb_passes_{n}_traffic.txt"
    if not path.exists():
        return None
    kv = dict(line.split("=", 1) for line in path.read_text().splitlines() if "=" in line and not line.startswith("#"))
    key = f"traffic_bytes[{kernel}]"
    return float(kv[key]) if key in kv else None


def profile_program(kind: str, n: int) -> int:
    """Short, traceable runs for rocprofv3 (kernel trace or counters)."""
    os.environ["FAST_TAU_GRAPHS"] = "0"   # every kernel launched directly
    from fast_amd import tau as T
    from fast_amd import workloads as W
    from fast_amd._lib import lib
    mask = W.ellipsoids_256(0) if n == 256 else (W.needles_512(0) if n == 512 else
                                                  T.generate_ellipsoids(n, 12.0, 0.35, 0.65, 0))
    ctx = T.Context(0)
    vol = ctx.volume(mask)
    if kind == "rb":
        rb = (C.c_double * 5)()
        lib().fast_debug_time_rb_passes(vol.handle, 0, 20, rb)
        print(f"mixed cg red-black: P, Y, Q, U ms {list(rb)[:4]} iteration ms {rb[4]}", flush=True)
        # calibration kernels of known traffic: triad (read 2 x 1 GiB, write 1 GiB), copy, fill
        st = (C.c_double * 4)()
        lib().fast_debug_stream(ctx.handle, 2 ** 30, 3, st)
        print(f"stream 1 GiB buffers: triad, triad (8192 blocks), copy, fill ms {list(st)}", flush=True)
    elif kind == "solves":
        for name, solver, prec, full in (("fp64 bicgstab", BICG, "double", False), ("fp64 cg", CG, "double", False),
                                         ("mixed cg red-black", CG, "mixed", False),
                                         ("mixed cg full grid", CG, "mixed", True)):
            res, _ = vol.solve(T.make_params(0, solver, prec, TOL, 10000, full_grid=full))
            print(f"{name}: tau {res.tau!r} iterations {res.iterations} {res.solve_ms:.1f} ms", flush=True)
    else:
        ms = (C.c_double * 4)()
        for name, prec, solver in (("fp64 bicgstab", 0, BICG), ("fp64 cg", 0, CG), ("mixed cg full grid", 2, CG)):
            lib().fast_debug_time_passes(vol.handle, 0, prec, solver, 20, ms)
            print(f"{name}: pass ms {list(ms)[:3]} iteration ms {ms[3]}", flush=True)
        rb = (C.c_double * 5)()
        lib().fast_debug_time_rb_passes(vol.handle, 0, 20, rb)
        print(f"mixed cg red-black: P, Y, Q, U ms {list(rb)[:4]} iteration ms {rb[4]}", flush=True)
        al = (C.c_double * 3)()
        lib().fast_debug_time_alpha(vol.handle, 5, al)
        print(f"alpha: area ms {al[0]} tensor (cached labels) ms {al[1]} tensor with CCL ms {al[2]}", flush=True)
        import torch   # calibration kernel of known traffic: read 256 MiB, write 256 MiB
        a = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copy 256 MiB x 3 done", flush=True)
    vol.close()
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-iters", type=int, default=25)
    ap.add_argument("--serve", type=int, default=200, help="requests per offered load (0: skip serving)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving workload and print its JSON (with --serve, --serve-loads, --serve-rate)")
    ap.add_argument("--serve-loads", default="0.5,0.8,1.1", help="offered loads, fractions of the saturated rate")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="saturated rate (solves/s) the loads refer to; 0 measures it with a burst")
    ap.add_argument("--serve-pool", type=int, default=8, help="volumes per edge length in the request pool")
    ap.add_argument("--no-dram-check", dest="dram_check", action="store_false",
                    help="skip the red-black pass timings on the 512^3 needle volume")
    ap.add_argument("--profile", choices=["solves", "passes", "rb"],
                    help="short program for rocprofv3 instead of the benchmark: 'solves' runs the X_POS solves "
                         "(FP64 BiCGSTAB, FP64 CG, mixed CG red-black and full grid) with direct launches; 'passes' "
                         "times 20 launches of each solver's passes and runs a 256 MiB device copy for counter "
                         "calibration; 'rb' times 20 launches of each red-black pass and 3 of each streaming kernel over 1 GiB "
                         "buffers (triad, copy, fill: the counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent volumes shard as one queue per GPU with no communication)", file=sys.stderr)
        return 2
    if args.profile:
        return profile_program(args.profile, args.size)
    loads = tuple(float(f) for f in args.serve_loads.split(","))
    if args.serve_only:
        from fast_amd import tau as T
        from fast_amd import workloads as W
        ctx = T.Context(0)
        serve = serving(ctx, T, W, args.serve, loads, args.serve_pool, args.serve_rate)
        print(json.dumps({"serving": serve}), flush=True)
        ctx.close()
        return 0

    import fast_amd as F
    from fast_amd import tau as T
    from fast_amd import workloads as W
    from fast_amd._lib import lib

    n = args.size
    log(f"[bench] generating {n}^3 ellipsoid volume")
    mask = F.generate_ellipsoids(n, 12.0, 0.35, 0.65, 0)
    ctx = T.Context(0)
    vol = ctx.volume(mask)
    nvox = mask.size

    conv, conv_src = converged_taus(vol, T)

    # ---- plain FP64 drop-in, one direction per solve, both solvers ----
    log("[bench] FP64 warm-up")
    vol.solve(T.make_params(0, BICG, "double", TOL, 10000))
    fp64 = {}
    for solver in (BICG, CG):
        for d in DIRS:
            t = time.perf_counter()
            res, _ = vol.solve(T.make_params(d, solver, "double", TOL, 10000))
            fp64[(solver, d)] = (res, time.perf_counter() - t)
            log(f"[bench] FP64 solver={solver} d={d} tau={res.tau:.9f} it={res.iterations} "
                f"{fp64[(solver, d)][1] * 1e3:.1f} ms err={abs(res.tau - conv[d]) / conv[d]:.1e}")
    fp64_ms = {s: 1e3 * float(np.mean([fp64[(s, d)][1] for d in DIRS])) for s in (BICG, CG)}
    fp64_rate = {s: 1e3 / fp64_ms[s] for s in (BICG, CG)}

    # ---- mixed + batched: the three directions in one batched mixed CG solve ----
    params = [T.make_params(d, CG, "mixed", TOL, 10000) for d in DIRS]
    vols = [vol] * len(DIRS)
    for _ in range(args.warmup):
        T.solve_batch(ctx, vols, params)
    log(f"[bench] timing {args.steps} mixed batched steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = T.solve_batch(ctx, vols, params)
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = len(DIRS) * args.steps / elapsed
    err_mixed = {d: abs(o.tau - conv[d]) / conv[d] for o, d in zip(out, DIRS)}

    # ---- the same batch with mixed CG on the full grid ----
    params_full = [T.make_params(d, CG, "mixed", TOL, 10000, full_grid=True) for d in DIRS]
    T.solve_batch(ctx, vols, params_full)
    t0 = time.perf_counter()
    out_full = T.solve_batch(ctx, vols, params_full)
    full_rate = len(DIRS) / (time.perf_counter() - t0)
    err_full = {d: abs(o.tau - conv[d]) / conv[d] for o, d in zip(out_full, DIRS)}

    # ---- roofline of the dominant kernel, HIP events on the solve stream ----
    ms = (C.c_double * 4)()
    lib().fast_debug_time_passes(vol.handle, 0, 2, CG, 20, ms)
    mixed_full_pass = list(ms)
    lib().fast_debug_time_passes(vol.handle, 0, 0, CG, 20, ms)
    fp64_cg_pass = list(ms)
    lib().fast_debug_time_passes(vol.handle, 0, 0, BICG, 20, ms)
    fp64_bicg_pass = list(ms)
    rb = (C.c_double * 5)()
    lib().fast_debug_time_rb_passes(vol.handle, 0, 20, rb)
    rb_pass = list(rb)
    # `fast alpha` on the same volume: whole-volume area, six per-face areas from cached
    # labels, and the six from scratch (CCL included), HIP events on the same stream
    al = (C.c_double * 3)()
    lib().fast_debug_time_alpha(vol.handle, 10, al)
    alpha = {"area_ms": al[0], "tensor_cached_labels_ms": al[1], "tensor_with_ccl_ms": al[2],
             "area_voxels_per_s": nvox / (al[0] * 1e-3),
             "area_mask_read_GBs": nvox / (al[0] * 1e-3) / 1e9,
             "timing": "HIP events around 10 back-to-back runs of each"}
    # algorithmic bytes per launch (each vector element and code byte once): FP32 r, p, x, q
    # and FP64 Y, 1-byte codes, per red (n_r) and black (n_b) voxel
    n_r = (nvox + 1) // 2 if all(e % 2 for e in mask.shape) else nvox // 2
    n_b = nvox - n_r
    fused = os.environ.get("FAST_TAU_RB_STAGED", "2") == "2"
    if fused:   # Y computed inside the Q pass: p, red and black codes in, q out
        kernels = {
            "k_rb_it<0> (P, red: p = D^-1 r + beta p, x += a0 p_prev + a1 p_old)": (rb_pass[0], (4 + 4 + 1 + 4 + 4 + 2 * 4) * n_r),
            "k_rb_yq (Y+Q fused, red: q = S p = D_r p - A_rb D_b^-1 A_br p, p.q)": (rb_pass[2], (4 + 1 + 4) * n_r + 1 * n_b),
            "k_rb_it<3> (U, red: r -= alpha q, r.r, r.D^-1 r)": (rb_pass[3], (2 * 4 + 1 + 4) * n_r),
        }
    else:
        kernels = {
            "k_rb_it<0> (P, red: p = D^-1 r + beta p, x += a0 p_prev + a1 p_old)": (rb_pass[0], (4 + 4 + 1 + 4 + 4 + 2 * 4) * n_r),
            "k_rb_it<1> (Y, black: Y = D_b^-1 A_br p)": (rb_pass[1], (1 + 8) * n_b + 4 * n_r),
            "k_rb_q (Q, red: q = S p = D_r p - A_rb Y, p.q)": (rb_pass[2], (1 + 4 + 4) * n_r + 8 * n_b),
            "k_rb_it<3> (U, red: r -= alpha q, r.r, r.D^-1 r)": (rb_pass[3], (2 * 4 + 1 + 4) * n_r),
        }
    name, (kms, kbytes) = max(kernels.items(), key=lambda kv: kv[1][0])
    ach = kbytes / (kms * 1e-3) / 1e9
    traffic = measured_traffic(name.split(" ")[0], n)
    others = {k: {"avg_launch_ms": v[0], "bytes_per_launch": v[1], "achieved_GBs": v[1] / (v[0] * 1e-3) / 1e9,
                  "traffic": measured_traffic(k.split(" ")[0], n)}
              for k, v in kernels.items()}
    # P is timed in the iterations that add the two pending x steps (25 B per red voxel); the
    # other half of the iterations' P passes move 13 B, so an iteration averages 6 B less
    iter_bytes = sum(v[1] for v in kernels.values()) - 6 * n_r
    # the red-black state of one 256^3 volume (~260 MB) is about the size of the 256 MB
    # Infinity Cache, so part of the 256^3 traffic may be served from it; the same passes on
    # the 512^3 needle volume (~3 GB of state) are DRAM-bound
    dram = None
    if args.dram_check:
        log("[bench] red-black passes on the 512^3 needle volume")
        vol.close()
        big = ctx.volume(W.needles_512(0))
        lib().fast_debug_time_rb_passes(big.handle, 0, 10, rb)
        big.close()
        vol = ctx.volume(mask)
        nb = 512 ** 3
        nr5 = nb // 2
        if fused:
            b5 = {0: ("k_rb_it<0> (P)", 25 * nr5), 2: ("k_rb_yq (Y+Q)", 10 * nr5), 3: ("k_rb_it<3> (U)", 13 * nr5)}
        else:
            b5 = {0: ("k_rb_it<0> (P)", 25 * nr5), 1: ("k_rb_it<1> (Y)", 13 * nr5), 2: ("k_rb_q (Q)", 17 * nr5),
                  3: ("k_rb_it<3> (U)", 13 * nr5)}
        dram = {"volume": "512^3 needles a=12 r=0.1 solid 0.55 seed 0, X_POS",
                "red_black_tile": dict(zip(("rows", "planes", "prefetch_planes"), ctx.rb_shape((512, 512, 512)))),
                "kernels": {k: {"avg_launch_ms": rb[i], "bytes_per_launch": b,
                                "achieved_GBs": b / (rb[i] * 1e-3) / 1e9, "frac": b / (rb[i] * 1e-3) / 8e12,
                                "traffic": measured_traffic(k.split(" ")[0], 512)}
                            for i, (k, b) in b5.items()},
                "iteration_ms": rb[4], "iteration_GBs": (sum(b for _, b in b5.values()) - 6 * nr5) / (rb[4] * 1e-3) / 1e9}
        log("[bench] streaming kernels over 3 x 2 GiB (measured HBM ceiling)")
        dram["hbm_ceiling"] = hbm_ceiling(ctx, lib)
        if fused:   # the separate Y and Q passes the fused one replaces would move 30 B per red voxel
            dram["fused_yq_unfused_equivalent_GBs"] = 30 * nr5 / (rb[2] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] CPU oracle sample on {threads} threads")
        cpu = cpu_baseline(mask, float(fp64[(CG, 0)][0].iterations), threads, args.cpu_iters)

    serve = None
    if args.serve > 0:
        vol.close()
        serve = serving(ctx, T, W, args.serve, loads, args.serve_pool, args.serve_rate)

    best_fp64 = max(fp64_rate.values())
    at_n = {"volume": f"{n}^3 ellipsoids (the headline workload)", "kernel": name, "achieved": ach,
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": kbytes, "avg_launch_ms": kms}
    if dram:
        # the dominant kernel of the same passes on the 512^3 needle volume (~3 GB of state, DRAM-bound)
        dk, dv = max(dram["kernels"].items(), key=lambda kv: kv[1]["avg_launch_ms"])
        ceil = dram["hbm_ceiling"].get("triad", {}).get("GBs")
        roofline = {"bound": "hbm", "kernel": dk, "achieved": dv["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": dv["achieved_GBs"] / HBM_PEAK_GBS, "traffic": dv["traffic"],
                    "bytes_per_launch": dv["bytes_per_launch"], "avg_launch_ms": dv["avg_launch_ms"],
                    "volume": dram["volume"], "measured_ceiling_GBs": ceil,
                    "frac_of_measured_ceiling": dv["achieved_GBs"] / ceil if ceil else None,
                    "timing": "HIP events on the solve stream around 10 back-to-back launches of the kernel",
                    "at_256": at_n}
        if "fused_yq_unfused_equivalent_GBs" in dram and dk.startswith("k_rb_yq"):
            roofline["note"] = ("the fused Y+Q pass is bound by instruction issue, not HBM: it moves 10 B per red "
                                "voxel where the separate Y and Q passes it replaces move 30 B; rate on those 30 B: "
                                f"{dram['fused_yq_unfused_equivalent_GBs']:.0f} GB/s")
            eq = dram["fused_yq_unfused_equivalent_GBs"]
            roofline["unfused_equivalent"] = {"GBs": eq, "frac": eq / HBM_PEAK_GBS,
                                              "frac_of_measured_ceiling": eq / ceil if ceil else None}
    else:
        roofline = dict(at_n, bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
                        timing="HIP events on the solve stream around 20 back-to-back launches of the kernel")
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
        "dtype": "f32 storage, f64 arithmetic/reductions/solution",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 ellipsoids a=12 r=0.35 solid 0.65 seed 0; tol 1e-6; x,y,z submitted as one "
                        "batch of mixed-precision CG solves (red-black Schur complement) per step; the library "
                        "runs a batch of volumes this large as consecutive one-volume launches (each volume's "
                        "state is about the size of the 256 MB Infinity Cache)",
            "volume": [n, n, n], "solves_per_step": len(DIRS), "parallelism": "single GPU",
            "red_black_tile": dict(zip(("rows", "planes", "prefetch_planes"), ctx.rb_shape((n, n, n)))),
            "red_black_tile_source": "autotuned table (fast_amd/tuning_gfx950.txt, rb_ keys) or the size rule",
        },
        "mixed_batched_ms_per_solve": ms_per_step / len(DIRS),
        "mixed_full_grid_batched_solves_per_s": full_rate,
        "fp64_bicgstab_solves_per_s": fp64_rate[BICG],
        "fp64_cg_solves_per_s": fp64_rate[CG],
        "speedup_vs_fastest_fp64": value / best_fp64,
        "speedup_vs_fp64_cg": value / fp64_rate[CG],
        "speedup_attribution": {
            "precision_only": {"what": "mixed CG on the full grid vs FP64 CG (same recurrences, FP32 storage)",
                               "ratio": full_rate / fp64_rate[CG]},
            "red_black": {"what": "mixed CG on the red-black Schur complement vs on the full grid (half the "
                                  "iterations, half the unknowns)", "ratio": value / full_rate},
        },
        "converged_tau": {DIRNAMES[d]: conv[d] for d in DIRS},
        "converged_tau_source": conv_src,
        "tau_rel_err_vs_converged": {
            "mixed_batched_cg": {DIRNAMES[d]: err_mixed[d] for d in DIRS},
            "mixed_full_grid_cg": {DIRNAMES[d]: err_full[d] for d in DIRS},
            "fp64_bicgstab": {DIRNAMES[d]: abs(fp64[(BICG, d)][0].tau - conv[d]) / conv[d] for d in DIRS},
            "fp64_cg": {DIRNAMES[d]: abs(fp64[(CG, d)][0].tau - conv[d]) / conv[d] for d in DIRS},
        },
        "iterations": {
            "mixed_batched_cg": {DIRNAMES[d]: int(o.iterations) for o, d in zip(out, DIRS)},
            "mixed_full_grid_cg": {DIRNAMES[d]: int(o.iterations) for o, d in zip(out_full, DIRS)},
            "fp64_bicgstab": {DIRNAMES[d]: int(fp64[(BICG, d)][0].iterations) for d in DIRS},
            "fp64_cg": {DIRNAMES[d]: int(fp64[(CG, d)][0].iterations) for d in DIRS},
        },
        "roofline": roofline,
        "kernels_mixed_cg": others,
        "mixed_cg_iteration": {"ms": rb_pass[4], "algorithmic_bytes": iter_bytes,
                               "achieved_GBs": iter_bytes / (rb_pass[4] * 1e-3) / 1e9},
        "mixed_cg_passes_512_dram": dram,
        "pass_ms": {"mixed_cg_red_black": rb_pass, "mixed_cg_full_grid": mixed_full_pass, "fp64_cg": fp64_cg_pass,
                    "fp64_bicgstab": fp64_bicg_pass},
        "alpha_reactive_area_density": alpha,
        "cpu_baseline": cpu,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)
    vol.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
