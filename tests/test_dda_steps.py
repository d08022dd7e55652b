"""Pins the brick DDA's step counts (CaelixTraceBrickRay, shaders/render/caelix_brick_trace.hlsl:299-368)
independently of its code: in a brick whose eight coarse cells are all occupied there is no coarse jump,
so a ray that starts inside the brick takes one step per voxel it visits, found by a float64 voxel walk:
the voxels up to and including the first occupied one, or, when the ray leaves the brick, the voxels
crossed plus the step that finds it outside; at most 22 steps. The hit voxel, entry face and distance
come with it. Rays whose walk crosses two planes within 5e-4 of each other (where the DDA's 1e-4 push
may cut a corner) or start within 1e-3 of a voxel face are left out. On the oracle, and (gpu) the HIP
path against the oracle and the walk on the same rays."""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd.scene import SceneData

LOAD = 0.7
MAX_STEPS = 22


def brick_scene(kind: str, seed: int) -> SceneData:
    rng = np.random.default_rng(seed)
    v = np.arange(512)
    x, y, z = v & 7, (v >> 3) & 7, v >> 6
    if kind == "random":
        solid = rng.uniform(size=512) < 0.3
    else:  # "corners": one voxel in each corner, every walk long
        solid = ((x == 0) | (x == 7)) & ((y == 0) | (y == 7)) & ((z == 0) | (z == 7))
    payload = np.where(solid, 0x8000 | rng.integers(1, 256, size=512), 0).astype(np.uint16)[None, :]
    table = np.zeros((65536, 8), np.float32)
    table[0x8000:0x8100, 0:3] = 0.5
    camera = {"position": np.zeros(3, np.float32), "forward": np.array([0, 0, 1], np.float32),
              "right": np.array([1, 0, 0], np.float32), "up": np.array([0, 1, 0], np.float32),
              "tan_half_fov_y": np.float32(0.5), "aspect": np.float32(1.0), "width": 8, "height": 8}
    return SceneData(keys=np.zeros((1, 3), np.int32), payload=payload, materials=table,
                     light_codes=np.zeros(2048, np.uint32), grid_min=(0, 0, 0), grid_dim=8, group_shift=4,
                     camera=camera, emissive_bricks=np.zeros(0, np.int64))


def rays_inside(data, rng, count):
    """Origins in unoccupied voxels (opaque neighbours all around leave no occupied voxel in these
    bricks except the solid ones), at least 1e-3 from every voxel face; random unit directions, a
    sixth of them along an axis."""
    solid = data.payload[0] != 0
    o = []
    while len(o) < count:
        p = rng.uniform(0.0, 8.0, size=3)
        f = p - np.floor(p)
        c = np.floor(p).astype(int)
        if f.min() > 1e-3 and f.max() < 1 - 1e-3 and not solid[c[0] + 8 * c[1] + 64 * c[2]]:
            o.append(p)
    d = rng.normal(size=(count, 3))
    k = count // 6
    d[:k] = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1, 1], size=(k, 1))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([np.array(o), d, np.full((count, 1), 1e30)], axis=1).astype(np.float32)


def walk_steps(occupied: np.ndarray, o, d):
    """(steps, hit voxel or None, entry axis, t, smallest gap between two plane crossings)."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    cell = np.floor(o).astype(np.int64)
    step = np.where(d >= 0, 1, -1)
    with np.errstate(divide="ignore"):
        inv = np.where(d != 0, 1.0 / np.where(d == 0, 1.0, d), np.inf)
    t_next = np.where(d != 0, (cell + (step > 0) - o) * inv, np.inf)
    t_delta = np.abs(inv)
    t, axis, gap = 0.0, -1, np.inf
    for steps in range(1, MAX_STEPS + 1):
        if np.any(cell < 0) or np.any(cell > 7):
            return steps, None, axis, t, gap
        if occupied[cell[0] + 8 * cell[1] + 64 * cell[2]]:
            return steps, cell.copy(), axis, t, gap
        order = np.argsort(t_next)
        gap = min(gap, t_next[order[1]] - t_next[order[0]])
        axis = int(order[0])
        t = t_next[axis]
        cell[axis] += step[axis]
        t_next[axis] += t_delta[axis]
    return MAX_STEPS, None, axis, t, gap


def check_steps(data, rays, hits) -> int:
    occupied = data.payload[0] != 0
    checked = 0
    for r, h in zip(rays, hits):
        steps, voxel, axis, t, gap = walk_steps(occupied, r[0:3], r[3:6])
        if gap < 5e-4:
            continue
        checked += 1
        assert int(h[8]) == steps, (r, h, steps)
        if voxel is None:
            assert h[1] == 0
            continue
        cell = (int(h[1]) >> 6) & 0x1FF
        assert [cell & 7, (cell >> 3) & 7, cell >> 6] == voxel.tolist()
        if axis >= 0:
            assert int(h[1]) & 0x3F == 1 << (2 * axis + (1 if r[3 + axis] > 0 else 0))
            assert abs(float(h[0:1].view(np.float32)[0]) - t) < 1e-4 * max(1.0, t)
    return checked


@pytest.mark.parametrize("kind", ["random", "corners"])
def test_brick_dda_steps_oracle(oracle, kind):
    data = brick_scene(kind, 3)
    o = oracle.OracleScene.build(data, 2)
    assert (int(o.a["info"][0]) >> 16) & 0xFF == 0xFF  # every coarse cell occupied: no jumps
    rays = rays_inside(data, np.random.default_rng(4), 3000)
    hits = o.trace_rays(rays)
    assert check_steps(data, rays, hits) > 2500
    steps = hits[:, 8]
    if kind == "corners":
        assert steps.max() >= 15 and (hits[:, 1] == 0).mean() > 0.5
    assert len(np.unique(steps)) > 8


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["random", "corners"])
def test_brick_dda_steps_gpu(cx, oracle, kind):
    data = brick_scene(kind, 3)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    rays = rays_inside(data, np.random.default_rng(4), 3000)
    got = gpu.trace_rays(rays)
    assert np.array_equal(got, oracle.OracleScene.build(data, int(np.ceil(1 / LOAD))).trace_rays(rays))
    assert check_steps(data, rays, got) > 2500
