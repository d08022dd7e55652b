"""Seeded synthetic inputs of the hot path's workloads (numpy, host side).

light_leaves       emissive brick leaves (LightBvhLeaf) for the light-tree build: a share uniform in a
                   cube, the rest in Gaussian clusters; emission log-uniform. The 4,194,304-leaf
                   rebuild stress workload is light_leaves(1 << 22, seed, extent=8192, clusters=64,
                   sigma=32, clustered=0.5, emission=(0.01, 100)).
reference_sort_inputs
                   the keys of the reference's radix sort test (tests/light_tree_test.cpp:1468-1486):
                   one std::mt19937(3) stream over n = 1, 2, 1000, 100000; a third of the keys
                   (rng() % 64) * 0xFFFFF, the rest rng() & 0x3FFFFFFF. numpy's legacy RandomState(3)
                   with 32-bit draws is the same generator (init_genrand seeding).
"""
from __future__ import annotations

import numpy as np

from . import layout


def light_leaves(n: int, seed: int, extent: float = 8192.0, clusters: int = 64, sigma: float = 32.0,
                 clustered: float = 0.5, emission: tuple[float, float] = (0.01, 100.0)) -> np.ndarray:
    """n LightBvhLeaf records (numpy LEAF). Each leaf is an 8^3 brick's light box in world voxels: a
    tight integer box inside the brick that contains the sampled point, and six face sums whose
    total is a log-uniform emission."""
    rng = np.random.default_rng(seed)
    n_clustered = int(round(n * clustered))
    n_uniform = n - n_clustered
    points = np.empty((n, 3), np.float64)
    points[:n_uniform] = rng.uniform(0.0, extent, size=(n_uniform, 3))
    if n_clustered > 0:
        centres = rng.uniform(4 * sigma, extent - 4 * sigma, size=(clusters, 3))
        which = rng.integers(0, clusters, size=n_clustered)
        points[n_uniform:] = centres[which] + rng.normal(0.0, sigma, size=(n_clustered, 3))
    points = np.clip(points, 0.0, extent - 1.0)
    origin = np.floor(points / 8.0) * 8.0
    inside = np.floor(points - origin)  # 0..7, the voxel of the point
    a = rng.integers(0, 8, size=(n, 3)).astype(np.float64)
    b = rng.integers(1, 9, size=(n, 3)).astype(np.float64)
    lo = origin + np.minimum(np.minimum(a, b - 1), inside)
    hi = origin + np.maximum(np.maximum(a + 1, b), inside + 1)
    lo_e, hi_e = np.log(emission[0]), np.log(emission[1])
    power = np.exp(rng.uniform(lo_e, hi_e, size=n))
    weights = rng.uniform(0.0, 1.0, size=(n, 6)) * (rng.uniform(size=(n, 6)) > 0.3)
    weights[:, 0] += 1e-3  # every brick has some light
    faces = power[:, None] * weights / weights.sum(axis=1, keepdims=True)
    leaves = np.zeros(n, layout.LEAF)
    leaves["summary"]["face"] = faces.astype(np.float32)
    leaves["summary"]["lo"] = lo.astype(np.float32)
    leaves["summary"]["hi"] = hi.astype(np.float32)
    leaves["group"] = (np.arange(n) // 64).astype(np.uint32)
    leaves["entry"] = (np.arange(n) % 64).astype(np.uint32)
    leaves["parent"] = layout.LT_NONE
    return leaves


def reference_sort_inputs() -> dict[int, np.ndarray]:
    """{n: keys} for n = 1, 2, 1000, 100000, drawn in that order from one mt19937(3) stream."""
    rs = np.random.RandomState(3)
    out = {}
    for n in (1, 2, 1000, 100000):
        draws = rs.randint(0, 2**32, size=2 * n, dtype=np.uint32).astype(np.uint64)
        first, second = draws[0::2], draws[1::2]
        keys = np.where(first % 3 == 0, ((second % 64) * 0x00FFFFF) & 0xFFFFFFFF, second & 0x3FFFFFFF)
        out[n] = keys.astype(np.uint32)
    return out
