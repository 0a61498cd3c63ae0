"""Autotuner of the tile shapes per volume size class.

Full grid (FP64 and FP32/mixed BiCGSTAB, mixed CG with mixed_full_grid): for every size
class (per-axis edge class 32/64/128/256/512, precision) it times one Krylov iteration of
the solver that precision mostly runs (FP64: BiCGSTAB, mixed: CG) with HIP events for every
compiled block shape (tx x ty) and a range of tz, and writes the fastest as lines
"<key> <tx> <ty> <tz>".

Red-black mixed CG (the default mixed CG): for every size class it times one red-black
iteration (the P, fused Y+Q and U passes with both finalizes) for every rows-per-tile (2, 4,
8, 9, 10, 12, 13), planes per tile and fused-pass prefetch distance (2, 3), and writes the fastest as lines
"rb_<x>_<y>_<z> <ty> <tz> <ahead>".

The lines go to tuning_gfx950.txt next to libfasttau.so, where the library picks them up
(or wherever $FAST_TAU_TUNING points).  Tile shapes never change results: every dot product
is rounded once from a double-double sum.

  python -m fast_amd.autotune [--classes 32,64,128,256] [--rb-classes 32,...,512] [--out PATH]
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


RB_TY = (2, 4, 8, 9, 10, 12, 13)
RB_TZ = (4, 8, 12, 16, 24, 32, 48, 64)
RB_AHEAD = (2, 3)


def rb_key(nx: int, ny: int, nz: int) -> str:
    return f"rb_{cls(nx)}_{cls(ny)}_{cls(nz)}"


def time_rb_iteration(ctx, vol, iters: int) -> float:
    ms = (C.c_double * 5)()
    rc = lib().fast_debug_time_rb_passes(vol.handle, 0, iters, ms)
    if rc != 0:
        raise RuntimeError(f"timing failed ({rc}): {ctx.last_error()}")
    return ms[4]


def tune_rb(classes, iters: int = 10, verbose: bool = True) -> dict:
    """Red-black shapes: every (ty, tz, ahead) per size class; cubes of class 512 only."""
    ctx = T.Context(0)
    best = {}
    for dims in itertools.product(classes, repeat=3):
        if max(dims) >= 512 and min(dims) < 512:
            continue
        nx, ny, nz = dims
        m = T.generate_ellipsoids((nx, ny, nz), 12.0, 0.35, 0.6, 1)
        vol = ctx.volume(m)
        results = []
        for ty, tz, ahead in itertools.product(RB_TY, RB_TZ, RB_AHEAD):
            if tz > nz or ty > 2 * ny:
                continue
            ctx.set_rb_shape(ty, tz, ahead)
            results.append((time_rb_iteration(ctx, vol, iters), ty, tz, ahead))
        ctx.set_rb_shape(0)
        results.sort()
        t, ty, tz, ahead = results[0]
        best[rb_key(nx, ny, nz)] = (ty, tz, ahead, t)
        if verbose:
            print(f"{rb_key(nx, ny, nz)}: best ty={ty} tz={tz} ahead={ahead} {t * 1e3:.1f} us/iteration "
                  f"(worst {results[-1][0] * 1e3:.1f})", flush=True)
        vol.close()
    ctx.close()
    return best


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
    return best


def write_table(out: Path, grid: dict, rb: dict) -> None:
    """Merge new entries into the table at `out` (entries of other keys are kept)."""
    lines = {}
    if out.exists():
        for line in out.read_text().splitlines():
            parts = line.split()
            if len(parts) == 4:
                lines[parts[0]] = " ".join(parts[1:])
    for k, (a, b, c, _) in grid.items():
        lines[k] = f"{a} {b} {c}"
    for k, (a, b, c, _) in rb.items():
        lines[k] = f"{a} {b} {c}"
    out.write_text("".join(f"{k} {v}\n" for k, v in sorted(lines.items())))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="32,64,128,256", help="full-grid size classes ('' to skip)")
    ap.add_argument("--rb-classes", default="32,64,128,256,512", help="red-black size classes ('' to skip)")
    ap.add_argument("--out", default=str(LIB_PATH.parent / "tuning_gfx950.txt"))
    ap.add_argument("--iters", type=int, default=12)
    a = ap.parse_args()
    t = time.time()
    grid = tune([int(c) for c in a.classes.split(",")], Path(a.out), a.iters) if a.classes else {}
    rb = tune_rb([int(c) for c in a.rb_classes.split(",")], a.iters) if a.rb_classes else {}
    write_table(Path(a.out), grid, rb)
    print(f"tuned in {time.time() - t:.0f} s -> {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
