"""GPU parity of the trace side (libcaelix_amd through its C ABI) against the oracle on the same
scenes: record words, bitmaps, light records and the light tree bit-exact; world DDA hits, light
draws and face pdfs bit-exact; the path-traced frame's hit attributes, depths and ray/step counters
exact and its radiance within 2e-5 relative of the oracle's float64 accumulation (the kernels add the
same float32 terms in float32); the furnace closed form on the GPU path."""
from __future__ import annotations

import numpy as np
import pytest

import lt_structure as lts
import trace_checks as tc
from caelix_amd import scene

pytestmark = pytest.mark.gpu

SMALL = scene.small_scene()
LOAD = 0.7


def slot_count(data) -> int:
    return int(np.ceil(len(data.keys) / LOAD))


@pytest.fixture(scope="module")
def pair(cx, oracle):
    gpu = cx.TraceScene(SMALL, load_factor=LOAD)
    assert gpu.hash.slot_count == slot_count(SMALL)
    arrays = gpu.host_arrays()
    ref = oracle.OracleScene.build(SMALL, slot_count(SMALL))
    return gpu, arrays, ref


def test_scene_records_bit_exact(pair):
    gpu, got, ref = pair
    want = ref.a
    assert np.array_equal(got["slots"], want["slots"].reshape(-1))
    assert got["hash_info"][0] == want["hash_info"][0]
    for name in ("info", "occupancy", "brick_bits", "super_bits", "group_of_brick"):
        assert np.array_equal(got[name], want[name]), name


def test_light_tables_bit_exact(pair):
    """The group pass, gathers and builds of the light tree adapter (after its heap and leaf growth)
    equal the oracle's; and the GPU's tables pass CheckStructure."""
    gpu, got, ref = pair
    tables = lts.from_gpu({k[3:]: v for k, v in got.items() if k.startswith("lt_")})
    lts.compare_tables(tables, lts.from_oracle(ref.lt))
    totals = lts.check_structure(SMALL, got["group_of_brick"], gpu.group_coords, tables)
    assert totals["bricks"] == len(SMALL.emissive_bricks)
    st = gpu.lights.stats()
    assert st["light_bricks"] == totals["bricks"] and st["light_faces"] == totals["faces"]
    assert st["tree_leaves"] == [totals["bricks"], 0] and st["tree_leaves_dropped"] == [0, 0]
    assert st["passes"] >= 1 and st["passes"] == st["failed_passes"] + 1


def _rays(solid, rng, count):
    pts = []
    while len(pts) < count:
        p = rng.uniform(0.5, 127.5, size=3)
        c = np.floor(p).astype(int)
        if not solid[c[0], c[1], c[2]]:
            pts.append(p)
    d = rng.normal(size=(count, 3))
    d[: count // 8] = np.eye(3)[rng.integers(0, 3, count // 8)] * rng.choice([-1, 1], size=(count // 8, 1))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t_max = rng.choice([1e30, 5.0, 40.0], size=(count, 1))
    return np.concatenate([np.array(pts), d, t_max], axis=1).astype(np.float32)


def test_trace_rays_bit_exact(pair):
    gpu, _, ref = pair
    solid = tc.dense_blocks(SMALL) != 0
    rays = _rays(solid, np.random.default_rng(3), 20000)
    # rays that start outside the grid too
    rays[:500, 0:3] = np.random.default_rng(4).uniform(-50, 180, size=(500, 3))
    got = gpu.trace_rays(rays)
    want = ref.trace_rays(rays)
    assert (want[:, 1] != 0).mean() > 0.1
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (bad[:5], got[bad[:3]], want[bad[:3]])


def _receivers(solid, rng, count):
    pts = []
    while len(pts) < count:
        p = rng.uniform(1.0, 127.0, size=3)
        c = np.floor(p).astype(int)
        if not solid[c[0], c[1], c[2]]:
            pts.append(p)
    n = rng.normal(size=(count, 3))
    n[: count // 4] = 0.0
    n[count // 4: count // 2] = [0.0, 1.0, 0.0]
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return np.concatenate([np.array(pts), n], axis=1).astype(np.float32)


def test_light_sample_bit_exact(pair):
    gpu, _, ref = pair
    rng = np.random.default_rng(8)
    recv = _receivers(tc.dense_blocks(SMALL) != 0, rng, 20000)
    seeds = rng.integers(0, 2 ** 32, size=len(recv), dtype=np.uint64).astype(np.uint32)
    got, ok = gpu.light_sample(recv, seeds)
    want, ok_want = ref.light_sample(recv, seeds)
    assert np.array_equal(ok, ok_want.astype(ok.dtype)) and ok.all()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_light_face_pdf_bit_exact(pair):
    gpu, _, ref = pair
    rng = np.random.default_rng(10)
    recv = _receivers(tc.dense_blocks(SMALL) != 0, rng, 4)
    open_ = tc.open_light_faces(SMALL)
    b, v, f = np.nonzero(open_)
    vox = SMALL.keys[b].astype(np.int64) * 8 + np.stack([v & 7, (v >> 3) & 7, v >> 6], axis=1)
    for r in recv:
        q = np.zeros((len(b), 12), np.uint32)
        q[:, 0:6] = r.view(np.uint32)
        q[:, 6] = b
        q[:, 7:10] = vox.astype(np.int32).view(np.uint32)
        q[:, 10] = f
        q[:, 11] = SMALL.payload[b, v]
        # plus queries that are not light faces (wrong face, non-light block): pdf 0
        q[:50, 10] = (q[:50, 10] + 1) % 6
        got = gpu.light_face_pdf(q)
        want = ref.light_face_pdf(q)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert abs(want.astype(np.float64).sum() - 1.0) < 0.05


@pytest.mark.parametrize("bounces,spp,mis", [(0, 1, True), (1, 1, True), (2, 1, True), (2, 2, False), (4, 1, True)])
def test_trace_frame_matches_oracle(cx, pair, bounces, spp, mis):
    gpu, _, ref = pair
    gpu.view.mis = 1 if mis else 0
    ref.view.mis = 1 if mis else 0
    try:
        color, radiance, attrib, counters = gpu.trace_frame(SMALL.camera, bounces, 1234, spp)
        w_color, w_rad, w_attrib, w_counters = ref.trace_frame(SMALL.camera, bounces, 1234, spp)
    finally:
        gpu.view.mis = 1
        ref.view.mis = 1
    color = color.cpu().numpy().reshape(-1, 4).astype(np.float64)
    radiance = radiance.cpu().numpy().reshape(-1, 4).astype(np.float64)
    attrib = attrib.cpu().numpy().view(np.uint32)
    counters = counters.cpu().numpy().view(np.uint64)[:5]
    assert np.array_equal(attrib, w_attrib)
    assert np.array_equal(color[:, 3], w_color[:, 3])  # primary depth, float32 in both
    assert np.array_equal(counters, w_counters[:5]), (counters, w_counters)
    tol = 2e-5 * (1.0 + np.abs(w_rad[:, :3]))
    assert np.all(np.abs(radiance[:, :3] - w_rad[:, :3]) <= tol)
    tol = 2e-5 * (1.0 + np.abs(w_color[:, :3]))
    assert np.all(np.abs(color[:, :3] - w_color[:, :3]) <= tol)


@pytest.mark.parametrize("bounces", [1, 3])
def test_trace_frame_furnace_closed_form(cx, oracle, bounces):
    data = scene.furnace_box(width=128, height=128)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    spp = 8
    _, radiance, attrib, _ = gpu.trace_frame(data.camera, bounces, 99, spp)
    r = radiance.cpu().numpy().reshape(-1, 4)[:, 0].astype(np.float64) / spp
    assert np.all(attrib.cpu().numpy() != 0)
    expected = (1 - 0.5 ** bounces) / (1 - 0.5)
    stderr = r.std() / np.sqrt(len(r))
    assert abs(r.mean() - expected) < 5 * stderr + 1e-4, (r.mean(), expected, stderr)


def test_trace_frame_deterministic_and_seeded(pair):
    gpu, _, _ = pair
    a = gpu.trace_frame(SMALL.camera, 2, 5, 1)[1].cpu().numpy()
    b = gpu.trace_frame(SMALL.camera, 2, 5, 1)[1].cpu().numpy()
    c = gpu.trace_frame(SMALL.camera, 2, 6, 1)[1].cpu().numpy()
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_trace_rejects_bad_arguments(cx, pair):
    gpu, _, _ = pair
    cam = dict(SMALL.camera)
    cam["width"] = 0
    with pytest.raises(RuntimeError):
        gpu.trace_frame(cam, 1, 0, 1)
    with pytest.raises(RuntimeError):
        gpu.trace_frame(SMALL.camera, 1, 0, 0)


def test_occupied_box_clip_changes_nothing(cx, oracle):
    """The same bricks in a grid four times wider: cx_scene_finalize's occupied super-cell box is the
    box of the bricks, and ending rays where they leave it (widened by one super cell) changes no hit,
    no counter and no pixel: the same results as walking the whole grid, and as the oracle."""
    import dataclasses

    big = dataclasses.replace(SMALL, grid_min=(-24, -16, -20), grid_dim=64)
    gpu = cx.TraceScene(big, load_factor=LOAD)
    keys = np.asarray(big.keys)
    has_voxel = (np.asarray(big.payload).reshape(len(keys), -1) != 0).any(axis=1)
    sc = (keys[has_voxel] - np.asarray(big.grid_min)) // 8
    assert np.array_equal(gpu.occupied_box.cpu().numpy(), np.concatenate([sc.min(axis=0), sc.max(axis=0)]))
    solid = tc.dense_blocks(SMALL) != 0
    rays = _rays(solid, np.random.default_rng(13), 20000)
    rays[:2000, 0:3] = np.random.default_rng(14).uniform(-200, 330, size=(2000, 3))
    clipped_rays = gpu.trace_rays(rays)
    clipped_frame = [a.cpu().numpy() for a in gpu.trace_frame(SMALL.camera, 2, 9, 1)]
    gpu.scene.occupied_box = None
    whole_rays = gpu.trace_rays(rays)
    whole_frame = [a.cpu().numpy() for a in gpu.trace_frame(SMALL.camera, 2, 9, 1)]
    assert np.array_equal(clipped_rays, whole_rays)
    for a, b in zip(clipped_frame, whole_frame):
        assert a.tobytes() == b.tobytes()
    ref = oracle.OracleScene.build(big, slot_count(big))
    assert np.array_equal(clipped_rays, ref.trace_rays(rays))


def test_trace_frame_odd_image_size(pair):
    """An image that is not a multiple of the 8x8 path tiles: padding paths trace nothing."""
    gpu, _, ref = pair
    cam = dict(SMALL.camera)
    cam["width"], cam["height"] = 100, 37
    color, radiance, attrib, counters = gpu.trace_frame(cam, 2, 77, 1)
    w_color, w_rad, w_attrib, w_counters = ref.trace_frame(cam, 2, 77, 1)
    assert np.array_equal(attrib.cpu().numpy().view(np.uint32), w_attrib)
    assert np.array_equal(counters.cpu().numpy().view(np.uint64)[:5], w_counters[:5])
    rad = radiance.cpu().numpy().reshape(-1, 4)[:, :3].astype(np.float64)
    assert np.all(np.abs(rad - w_rad[:, :3]) <= 2e-5 * (1.0 + np.abs(w_rad[:, :3])))


@pytest.mark.parametrize("encoding", [0, 1])
def test_glass_world_finalize_and_rays(cx, oracle, encoding):
    """The reference finalize test's glass world in both block encodings: the finalized payload (face
    bits of glass and of the air around it), record words, occupancy and bitmaps equal the oracle's, and
    rays through it (terminating at the transparent faces, caelix_brick_trace.hlsl:275-284) hit the same
    voxels with the same t bits, faces and step counts."""
    data = scene.glass_world(encoding)
    print(f"glass world encoding={encoding}", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    arrays = gpu.host_arrays()
    ref = oracle.OracleScene.build(data, slot_count(data))
    for name in ("payload", "info", "occupancy", "brick_bits", "super_bits"):
        assert np.array_equal(arrays[name][: ref.a[name].size], ref.a[name]), name
    assert (arrays["payload"] >> (8 if encoding == 0 else 9) & 0x3F).any()  # face bits were written
    rng = np.random.default_rng(50 + encoding)
    count = 20000
    o = rng.uniform([-4.0, 6.0, -4.0], [28.0, 16.0, 12.0], size=(count, 3))
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d, np.full((count, 1), 1e30)], axis=1).astype(np.float32)
    got = gpu.trace_rays(rays)
    want = ref.trace_rays(rays)
    glass_hits = ((want[:, 1] >> 16) & 0xFFFF) & (0xC000 if encoding == 0 else 0x8000) == 0
    assert ((want[:, 1] != 0) & glass_hits).sum() > 1000  # many rays end on a glass or air interface
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (bad[:5], got[bad[:3]], want[bad[:3]])


def test_bounce_nee_budget_gpu(cx, oracle):
    """bounce_nee_probability 0.5 on the GPU: the oracle's frame (counters exact, radiance 2e-5) and the
    furnace's closed form."""
    data = scene.furnace_box(width=96, height=96)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = oracle.OracleScene.build(data, slot_count(data))
    color, radiance, attrib, counters = gpu.trace_frame(data.camera, 3, 17, 4, bounce_nee_probability=0.5)
    w_color, w_rad, w_attrib, w_counters = ref.trace_frame(data.camera, 3, 17, 4, bounce_nee_probability=0.5)
    rad = radiance.cpu().numpy().reshape(-1, 4).astype(np.float64)
    assert np.array_equal(attrib.cpu().numpy().view(np.uint32), w_attrib)
    assert np.array_equal(counters.cpu().numpy().view(np.uint64)[:5], w_counters[:5])
    assert np.all(np.abs(rad[:, :3] - w_rad[:, :3]) <= 2e-5 * (1.0 + np.abs(w_rad[:, :3])))
    per_pixel = rad[:, 0] / 4
    assert abs(per_pixel.mean() - 1.75) < 5 * per_pixel.std() / np.sqrt(len(per_pixel)) + 1e-4
    with pytest.raises(RuntimeError):
        gpu.trace_frame(data.camera, 3, 17, 1, bounce_nee_probability=1.5)
