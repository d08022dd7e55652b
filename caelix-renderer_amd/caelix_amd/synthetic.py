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


def _morton_cells(codes: np.ndarray) -> np.ndarray:
    """The (x, y, z) cells of 30-bit Morton codes (bit 3b -> x bit b, 3b + 1 -> y, 3b + 2 -> z)."""
    codes = codes.astype(np.uint64)
    cells = np.zeros((len(codes), 3), np.int64)
    for b in range(10):
        for axis in range(3):
            cells[:, axis] |= ((codes >> np.uint64(3 * b + axis)) & np.uint64(1)).astype(np.int64) << b
    return cells


def leaves_at_codes(codes: np.ndarray, seed: int) -> np.ndarray:
    """Leaves whose box centres quantise to the given Morton codes, provided the codes hold cell 0 and
    cell 1023 on every axis (then the centre bounds are 0.25 .. 1023.25 and cell X is centre X + 0.25:
    X * 1024 / 1023 rounds down to X). Boxes are [X, X + 0.5]; face sums log-uniform."""
    cells = _morton_cells(codes).astype(np.float32)
    n = len(codes)
    rng = np.random.default_rng(seed)
    leaves = np.zeros(n, layout.LEAF)
    leaves["summary"]["lo"] = cells
    leaves["summary"]["hi"] = cells + np.float32(0.5)
    power = np.exp(rng.uniform(np.log(0.01), np.log(100.0), size=n))
    weights = rng.uniform(0.1, 1.0, size=(n, 6))
    leaves["summary"]["face"] = (power[:, None] * weights / weights.sum(axis=1, keepdims=True)).astype(np.float32)
    leaves["group"] = (np.arange(n) // 64).astype(np.uint32)
    leaves["entry"] = (np.arange(n) % 64).astype(np.uint32)
    leaves["parent"] = layout.LT_NONE
    return leaves


def degenerate_leaves(kind: str, n: int, seed: int) -> np.ndarray:
    """Light-tree inputs at the edges of the key distribution (n leaves, in shuffled order):
    equal      every box centre the same point: all codes 0, the tree is the positions' radix tree
    caterpillar codes 0 and 2^k (k = 0..29) plus the far corner, most leaves on code 1: a 30-deep spine
               of single-code groups whose deepest group adds log2(n) levels of equal-key splits
    cluster    one Gaussian cluster of sigma 0.5 voxels: a few distinct centres, huge equal-key groups
    """
    rng = np.random.default_rng(seed)
    if kind == "equal":
        codes = np.zeros(n, np.uint32)
        leaves = leaves_at_codes(codes, seed)
    elif kind == "caterpillar":
        spine = np.array([0] + [1 << k for k in range(30)] + [(1 << 30) - 1], np.uint64)
        sizes = np.maximum(1, (n >> np.arange(2, 2 + len(spine))).astype(np.int64))
        sizes[1] += n - sizes.sum()
        codes = np.repeat(spine, sizes).astype(np.uint32)
        leaves = leaves_at_codes(codes, seed)
    elif kind == "cluster":
        return light_leaves(n, seed, extent=8192.0, clusters=1, sigma=0.5, clustered=1.0)
    else:
        raise ValueError(kind)
    return leaves[rng.permutation(n)]


def pool_stress_leaves(seed: int) -> np.ndarray:
    """16,385 leaves whose first 16,384 sorted positions (one upper-level block of 32 sub-blocks of 512)
    hold the most nodes crossing sub-block boundaries that 30-bit codes allow: the block is cut into 33
    parts by a 6-bit code prefix, part k (1..31) spanning positions [512k - 256, 512k + 256), so each
    sub-block boundary lies in the middle of a part; inside a part, 24 suffix bits peel a group of 20
    leaves off alternately the left and the right end, around a middle group of 32 leaves that
    straddles the boundary. Every boundary is then crossed by a chain of 25 nested nodes (807 nodes
    built in the block, 1,614 arrivals). The last leaf, at the far corner, fixes the centre bounds."""
    codes = []
    for part in range(33):
        prefix = part << 24
        if part in (0, 32):
            codes.append(np.full(256, prefix, np.uint32))  # a flat part: one equal-key group
            continue
        path = 0
        left_groups, right_groups = [], []
        for j in range(23, -1, -1):
            if (23 - j) % 2 == 0:  # left peel: the group has bit j clear, the middle path sets it
                left_groups.append(prefix | path)
                path |= 1 << j
            else:  # right peel: the group sets bit j, the middle path keeps it clear
                right_groups.append(prefix | path | (1 << j))
        middle = prefix | path
        part_codes = [np.full(20, c, np.uint32) for c in left_groups]
        part_codes.append(np.full(32, middle, np.uint32))
        part_codes += [np.full(20, c, np.uint32) for c in reversed(right_groups)]  # outermost last
        codes.append(np.concatenate(part_codes))
    codes.append(np.array([(1 << 30) - 1], np.uint32))
    codes = np.concatenate(codes)
    assert len(codes) == 16385 and np.all(np.diff(codes.astype(np.int64)) >= 0)
    rng = np.random.default_rng(seed)
    return leaves_at_codes(codes, seed)[rng.permutation(len(codes))]
