"""GPU tests at the workloads' own sizes, through the C ABI:

1080p DDA + NEE   hd_scene(): 1,048,576 bricks in 4096^3, hash load 0.7, 65,536 emissive bricks,
                  1920 x 1080, 1 spp, 2 bounces. The scene built on the GPU equals the oracle's built on the
                  CPU (hash slots, record words, occupancy, bitmaps, light records, light tree); the whole
                  frame equals the oracle's (hit attributes, depths and ray / shadow / cell / probe / step
                  counters exact, color and radiance within 2e-5 relative), every value finite
batched           chunk_batch(): 256 chunk scenes of 64..65,536 bricks (log-uniform), 5 % emissive, each a
                  light-tree build and a 256 x 256, 1 spp, 2-bounce frame: the batched build and wavefront
                  equal every scene built and traced alone, and 8 scenes equal the oracle
differentiable    diff_scene(): 262,144 bricks, 16,384 emitters, 512 x 512, 4 spp, 3 bounces, L2 loss to a
                  target rendered from parameters perturbed by seeded N(0, 0.05): the adjoint against
                  central finite differences (h = 1e-3, roulette off) on its 8 largest gradients, and
                  against the oracle's float64 gradient over a band of rows
"""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd import scene
from test_glass_gpu import compare
import lt_structure as lts

pytestmark = pytest.mark.gpu

LOAD = 0.7


def slot_count(data) -> int:
    return int(np.ceil(len(data.keys) / LOAD))


# ------------------------------------------------------------------------------------ 1080p frame


@pytest.fixture(scope="module")
def hd(cx, oracle):
    data = scene.hd_scene()
    print(f"1080p scene: {len(data.keys)} bricks, {len(data.emissive_bricks)} emissive", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = oracle.OracleScene.build(data, slot_count(data))
    print("1080p scene built on the GPU and by the oracle", flush=True)
    return data, gpu, ref


def test_hd_scene_records_match_oracle(hd):
    data, gpu, ref = hd
    got, want = gpu.host_arrays(), ref.a
    assert len(data.keys) == 1 << 20 and gpu.hash.slot_count == 1497966
    assert np.array_equal(got["slots"], want["slots"].reshape(-1))
    for name in ("info", "occupancy", "brick_bits", "super_bits", "payload", "group_of_brick"):
        assert np.array_equal(got[name], want[name]), name
    tables = lts.from_gpu({k[3:]: v for k, v in got.items() if k.startswith("lt_")})
    lts.compare_tables(tables, lts.from_oracle(ref.lt))
    assert int(tables["control"]["light_bricks"]) == 65536
    assert int(tables["headers"][0]["leaf_count"]) == 65536


def test_hd_frame_matches_oracle(hd):
    data, gpu, ref = hd
    cam = data.camera
    got = gpu.trace_frame(cam, 2, 1, 1)
    want = ref.trace_frame(cam, 2, 1, 1)
    color = compare(got, want)
    radiance = got[1].cpu().numpy()
    assert np.all(np.isfinite(color)) and np.all(np.isfinite(radiance))
    counters = got[3].cpu().numpy().view(np.uint64)
    pixels = cam["width"] * cam["height"]
    hits = int((got[2].cpu().numpy() != 0).sum())
    assert pixels == 1920 * 1080 and hits > 0.5 * pixels
    # at most 1 + 2 rays and 2 shadow rays per path
    assert pixels <= counters[0] <= 3 * pixels and 0 < counters[1] <= 2 * pixels
    depth = color[:, 3]
    assert np.all(depth[got[2].cpu().numpy() != 0] > 0)


# ---------------------------------------------------------------------------------------- batched


@pytest.fixture(scope="module")
def chunks(cx):
    from caelix_amd.batch import PlainScenes, TraceBatch

    datas = scene.chunk_batch()
    print(f"chunk batch: {len(datas)} scenes, {sum(len(d.keys) for d in datas)} bricks", flush=True)
    return datas, TraceBatch(datas), PlainScenes(datas)


def test_chunk_batch_equals_each_scene_alone(chunks):
    datas, batch, plain = chunks
    assert len(datas) == 256 and all(d.camera["width"] == 256 for d in datas)
    batch.counters.zero_()
    plain.counters.zero_()
    batch.build_trees()
    batch.trace(2, 17, 1)
    plain.step(2, 17, 1)
    color, radiance, attrib = batch.images()
    p = plain.pixels
    lights = 0
    for i, (b, ts) in enumerate(zip(batch.builders, plain.scenes)):
        g, w = b.read(), ts.tree_read()
        assert g["header"].tobytes() == w["header"].tobytes(), i
        count = int(w["header"]["node_count"])
        assert g["nodes"][:count].tobytes() == w["nodes"][:count].tobytes(), i
        lights += ts.light_bricks
        assert np.array_equal(attrib[i].cpu().numpy(), plain.attrib[i * p:(i + 1) * p].cpu().numpy()), i
        assert np.array_equal(color[i].cpu().numpy().reshape(-1), plain.color[i * p * 4:(i + 1) * p * 4].cpu().numpy()), i
        assert np.array_equal(radiance[i].cpu().numpy().reshape(-1),
                              plain.radiance[i * p * 4:(i + 1) * p * 4].cpu().numpy()), i
    bricks = sum(len(d.keys) for d in datas)
    assert lights > 0.04 * bricks  # 5 % of the bricks are emitters with an open face
    assert np.array_equal(batch.counters.cpu().numpy()[[0, 1, 2, 4]], plain.counters.cpu().numpy()[[0, 1, 2, 4]])


def test_chunk_batch_scenes_match_oracle(chunks, oracle):
    datas, batch, _ = chunks
    batch.build_trees()
    batch.trace(2, 17, 1)
    color, radiance, attrib = batch.images()
    sizes = np.array([len(d.keys) for d in datas])
    picks = list(np.argsort(sizes)[[0, 1, 64, 128, 192, 254, 255]]) + [100]
    for i in picks:
        d = datas[int(i)]
        print(f"chunk scene {int(i)}: {len(d.keys)} bricks vs oracle", flush=True)
        w_color, w_rad, w_attrib, _ = oracle.OracleScene.build(d, slot_count(d)).trace_frame(d.camera, 2, 17, 1)
        assert np.array_equal(attrib[i].cpu().numpy().view(np.uint32), w_attrib)
        c = color[i].cpu().numpy().astype(np.float64)
        r = radiance[i].cpu().numpy().astype(np.float64)
        assert np.array_equal(c[:, 3], w_color[:, 3])
        assert np.all(np.abs(r[:, :3] - w_rad[:, :3]) <= 2e-5 * (1.0 + np.abs(w_rad[:, :3])))
        assert np.all(np.abs(c[:, :3] - w_color[:, :3]) <= 2e-5 * (1.0 + np.abs(w_color[:, :3])))


# --------------------------------------------------------------------------------- differentiable

NO_RR = 1
SPP, BOUNCES, SEED = 4, 3, 21


@pytest.fixture(scope="module")
def diff(cx):
    import torch

    data = scene.diff_scene()
    print(f"differentiable scene: {len(data.keys)} bricks, {len(data.emissive_bricks)} emitters", flush=True)
    ts = cx.TraceScene(data, load_factor=LOAD)
    rng = np.random.default_rng(77)
    albedo = data.materials[0x8000:0x8100, 0:3].copy()
    scale = np.ones(len(data.keys), np.float32)
    e = data.emissive_bricks
    # the target: parameters perturbed by N(0, 0.05)
    t_albedo = np.clip(albedo + rng.normal(0, 0.05, albedo.shape), 0.0, 1.0).astype(np.float32)
    t_scale = scale.copy()
    t_scale[e] = (1.0 + rng.normal(0, 0.05, len(e))).astype(np.float32)
    mats = ts.materials.view(65536, 8)
    mats[0x8000:0x8100, 0:3] = torch.from_numpy(t_albedo).cuda()
    scale_dev = torch.from_numpy(t_scale).cuda()
    ts.set_brick_scale(scale_dev)
    target = image(ts, data.camera)
    mats[0x8000:0x8100, 0:3] = torch.from_numpy(albedo).cuda()
    scale_dev.copy_(torch.from_numpy(scale))
    return data, ts, scale_dev, target


def image(ts, cam):
    color = ts.trace_frame(cam, BOUNCES, SEED, SPP, flags=NO_RR)[0]
    return color.view(-1, 4)[:, :3].double().cpu().numpy() / SPP


def adjoint(ts, d_color):
    import torch

    ga = torch.zeros(256, 4, dtype=torch.float32, device="cuda")
    gs = torch.zeros(ts.bricks, dtype=torch.float32, device="cuda")
    d = torch.from_numpy(np.ascontiguousarray(d_color, np.float32).reshape(-1)).cuda()
    ts.trace_frame_adjoint(ts.data.camera, BOUNCES, SEED, SPP, d, ga, gs, flags=NO_RR)
    return ga[:, :3].double().cpu().numpy(), gs.double().cpu().numpy()


def test_diff_adjoint_matches_finite_differences(diff):
    data, ts, scale_dev, target = diff
    cam = data.camera
    assert cam["width"] == 512 and len(data.keys) == 262144 and len(data.emissive_bricks) == 16384
    img = image(ts, cam)
    d_color = np.zeros((img.shape[0], 4), np.float32)
    d_color[:, :3] = 2.0 * (img - target)
    ga, gs = adjoint(ts, d_color)

    def loss():
        return float(((image(ts, cam) - target) ** 2).sum())

    mats = ts.materials.view(65536, 8)
    h = 1e-3
    for flat in np.argsort(-np.abs(ga).ravel())[:4]:
        mid, ch = divmod(int(flat), 3)
        base = float(mats[0x8000 + mid, ch])
        mats[0x8000 + mid, ch] = base + h
        lu = loss()
        mats[0x8000 + mid, ch] = base - h
        ld = loss()
        mats[0x8000 + mid, ch] = base
        fd = (lu - ld) / (2 * h)
        print(f"albedo[{mid}][{ch}]: fd {fd:.6g} adjoint {ga[mid, ch]:.6g}", flush=True)
        assert abs(fd - ga[mid, ch]) <= 1e-3 * abs(fd), (mid, ch, fd, ga[mid, ch])
    for b in np.argsort(-np.abs(gs))[:4]:
        base = float(scale_dev[b])
        scale_dev[b] = base + h
        lu = loss()
        scale_dev[b] = base - h
        ld = loss()
        scale_dev[b] = base
        fd = (lu - ld) / (2 * h)
        print(f"scale[{b}]: fd {fd:.6g} adjoint {gs[b]:.6g}", flush=True)
        assert abs(fd - gs[b]) <= 1e-3 * abs(fd), (b, fd, gs[b])


def test_diff_adjoint_matches_oracle_band(diff, oracle):
    data, ts, scale_dev, target = diff
    cam = data.camera
    img = image(ts, cam)
    rows = (240, 272)
    d_color = np.zeros((img.shape[0], 4), np.float32)
    band = slice(rows[0] * cam["width"], rows[1] * cam["width"])
    d_color[band, :3] = 2.0 * (img[band] - target[band])
    ga, gs = adjoint(ts, d_color)
    o = oracle.OracleScene.from_gpu_arrays(ts.host_arrays(), data)
    o.set_parameters(data.materials[0x8000:0x8100, 0:3], scale_dev.cpu().numpy())
    wa, ws = o.trace_frame_adjoint(cam, BOUNCES, SEED, SPP, d_color, flags=NO_RR, rows=rows)
    assert np.abs(wa).max() > 0 and np.abs(ws).max() > 0
    assert np.abs(ga - wa).max() <= 1e-4 * np.abs(wa).max()
    assert np.abs(gs - ws).max() <= 1e-4 * np.abs(ws).max()
