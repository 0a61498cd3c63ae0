"""Synthetic volumes of the benchmark and test workloads (seeded, reproducible).

All volumes are [z, y, x] uint8 arrays, 0 = pore, 255 = solid, made by the library's own
generators (xoshiro256** seeded; ellipsoids with semi-axes (a, a r, a r), orientations
uniform on SO(3), centres uniform, added until the solid fraction is reached).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .tau import generate_ellipsoids, generate_spheres

A_SEMI = 12.0   # long semi-axis of every ellipsoid workload


def spheres_64(seed: int = 0) -> np.ndarray:
    """64^3, spheres of radius 6, solid fraction 0.60."""
    return generate_spheres(64, 6.0, 0.60, seed)


def ellipsoids_256(seed: int = 0) -> np.ndarray:
    """256^3, ellipsoids a = 12, r = 0.35, solid fraction 0.65."""
    return generate_ellipsoids(256, A_SEMI, 0.35, 0.65, seed)


def needles_512(seed: int = 0) -> np.ndarray:
    """512^3, needle-like ellipsoids a = 12, r = 0.1, solid fraction 0.55."""
    return generate_ellipsoids(512, A_SEMI, 0.1, 0.55, seed)


@dataclass
class BatchVolume:
    mask: np.ndarray
    r: float
    solid: float
    seed: int


def _draw_r(rng: np.random.Generator) -> float:
    return float(math.exp(rng.uniform(math.log(0.1), math.log(10.0))))


def uneven_batch(count: int = 96, base_seed: int = 1000) -> list:
    """`count` volumes, each edge drawn from {32, 48, ..., 192} independently per axis,
    r log-uniform on [0.1, 10], solid fraction uniform on [0.5, 0.75], seed base + i."""
    edges = np.arange(32, 193, 16)
    out = []
    for i in range(count):
        rng = np.random.default_rng(base_seed + i)
        nx, ny, nz = (int(e) for e in rng.choice(edges, size=3))
        r = _draw_r(rng)
        solid = float(rng.uniform(0.5, 0.75))
        out.append(BatchVolume(generate_ellipsoids((nx, ny, nz), A_SEMI, r, solid, base_seed + i), r, solid,
                               base_seed + i))
    return out


@dataclass
class Request:
    arrival: float      # seconds after the stream starts
    edge: int
    dir: int            # 0 = X_POS, 2 = Y_POS, 4 = Z_POS
    volume: int         # index into the volume pool


REQUEST_EDGES = (64, 128, 192, 256, 384)
REQUEST_WEIGHTS = (0.3, 0.3, 0.2, 0.15, 0.05)


def request_volume(edge: int, j: int, base_seed: int = 5000) -> BatchVolume:
    """Volume j of edge length `edge` in the request pool: an edge^3 ellipsoid volume with r
    log-uniform on [0.1, 10] and solid fraction uniform on [0.5, 0.75], seeded by (edge, j)."""
    seed = base_seed + 100 * edge + j
    rng = np.random.default_rng(seed)
    r = _draw_r(rng)
    solid = float(rng.uniform(0.5, 0.75))
    return BatchVolume(generate_ellipsoids(edge, A_SEMI, r, solid, seed), r, solid, seed)


def request_pool(per_edge: int = 4, base_seed: int = 5000, threads: int = 8) -> dict:
    """Volumes the request stream draws from: `per_edge` per edge length (request_volume),
    generated on `threads` host threads (the generator releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    keys = [(e, j) for e in REQUEST_EDGES for j in range(per_edge)]
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        vols = list(ex.map(lambda k: request_volume(k[0], k[1], base_seed).mask, keys))
    return dict(zip(keys, vols))


def request_stream(count: int, rate: float, per_edge: int = 4, seed: int = 7) -> list:
    """Poisson arrivals at `rate` requests/s; cubic edge from REQUEST_EDGES with
    REQUEST_WEIGHTS; direction uniform in x, y, z; volume uniform within its edge's pool."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        e = int(rng.choice(REQUEST_EDGES, p=REQUEST_WEIGHTS))
        out.append(Request(t, e, int(rng.choice([0, 2, 4])), int(rng.integers(0, per_edge))))
    return out
