"""Autotuner of the tile shape (block tx x ty, tz planes per tile) per volume size class.

For every size class (per-axis edge class 32/64/128/256/512, precision) it times one
Krylov iteration of the solver that precision mostly runs (FP64: BiCGSTAB, mixed: CG) with
HIP events for every compiled block shape and a range of tz, and writes the fastest as
lines "<key> <tx> <ty> <tz>" to tuning_gfx950.txt next to libfasttau.so, where the library
picks them up (or wherever $FAST_TAU_TUNING points).  Tile shapes never change results:
every dot product is rounded once from a double-double sum.

  python -m fast_amd.autotune [--classes 32,64,128,256] [--out PATH]
"""
from __future__ import annotations

import argparse
import ctypes as C
import itertools
import sys
import time
from pathlib import Path

import numpy as np

from . import tau as T
from ._lib import lib, LIB_PATH

SHAPES = [(64, 8), (64, 4), (64, 16), (32, 8), (32, 16), (32, 32), (16, 16)]
TZS = [4, 8, 16, 32, 64]
PRECISIONS = {0: 0, 2: 1}   # precision -> solver timed (FP64: BiCGSTAB, mixed: CG)


def cls(e: int) -> int:
    for c in (16, 32, 64, 128, 256):
        if e <= c:
            return c
    return 512


def key(nx: int, ny: int, nz: int, precision: int) -> str:
    return f"{cls(nx)}_{cls(ny)}_{cls(nz)}_{precision}"


def time_iteration(ctx, vol, precision: int, solver: int, iters: int) -> float:
    ms = (C.c_double * 4)()
    rc = lib().fast_debug_time_passes(vol.handle, 0, precision, solver, iters, ms)
    if rc != 0:
        raise RuntimeError(f"timing failed ({rc}): {ctx.last_error()}")
    return ms[3]


def tune(classes, out: Path, iters: int = 12, verbose: bool = True) -> dict:
    ctx = T.Context(0)
    best = {}
    for dims in itertools.product(classes, repeat=3):
        if max(dims) >= 512 and min(dims) < 512:
            continue
        nx, ny, nz = dims
        m = T.generate_ellipsoids((nx, ny, nz), 6.0, 1.0, 0.6, 1)
        vol = ctx.volume(m)
        for prec, solver in PRECISIONS.items():
            k = key(nx, ny, nz, prec)
            results = []
            for (tx, ty), tz in itertools.product(SHAPES, TZS):
                if tx > 2 * nx and (tx, ty) != (16, 16):
                    continue
                if tz > nz:
                    continue
                lib().fast_ctx_set_tile_shape(ctx.handle, tx, ty, tz)
                t = time_iteration(ctx, vol, prec, solver, iters)
                results.append((t, tx, ty, tz))
            lib().fast_ctx_set_tile_shape(ctx.handle, 0, 0, 0)
            results.sort()
            t, tx, ty, tz = results[0]
            best[k] = (tx, ty, tz, t)
            if verbose:
                print(f"{k}: best {tx}x{ty} tz={tz} {t * 1e3:.1f} us/iteration "
                      f"(worst {results[-1][0] * 1e3:.1f})", flush=True)
        vol.close()
    ctx.close()
    with open(out, "w") as f:
        for k, (tx, ty, tz, t) in sorted(best.items()):
            f.write(f"{k} {tx} {ty} {tz}\n")
    return best


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="32,64,128,256")
    ap.add_argument("--out", default=str(LIB_PATH.parent / "tuning_gfx950.txt"))
    ap.add_argument("--iters", type=int, default=12)
    a = ap.parse_args()
    t = time.time()
    tune([int(c) for c in a.classes.split(",")], Path(a.out), a.iters)
    print(f"tuned in {time.time() - t:.0f} s -> {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
