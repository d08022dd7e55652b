"""GPU parity of the subsurface walk (CxScene::subsurface, CX_TRACE_SUBSURFACE) through the C ABI, on the
scenes of test_sss_oracle.py: the furnace cube (steps 64; one step with p 0.5 and sss_albedo 0.25), walks at
bounce vertices beside the camera's floor under a sky with sun NEE, the dead-walk slab: against the oracle
(hit attributes, depths and counters exact, colour and radiance within 2e-5 relative); the one-step closed
form on the GPU's own pixels; an SSS entry on an unused code changes nothing on the GPU either (the full
model's kernel instances against the plain ones, bit for bit); the adjoint refuses the walk."""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd import TRACE_SUBSURFACE
from test_sss_oracle import (PLAIN_CODE, SSS_CODE, WHITE_SKY, build, cube_world, floor_block_world, look_down_from,
                             one_step_survival, walk_scene)

pytestmark = pytest.mark.gpu

LOAD = 0.7


def compare(got, want):
    color, radiance, attrib, counters = got
    w_color, w_rad, w_attrib, w_counters = want
    color = color.cpu().numpy().reshape(-1, 4).astype(np.float64)
    radiance = radiance.cpu().numpy().reshape(-1, 4).astype(np.float64)
    assert np.array_equal(attrib.cpu().numpy().view(np.uint32), w_attrib)
    assert np.array_equal(color[:, 3], w_color[:, 3])
    assert np.array_equal(counters.cpu().numpy().view(np.uint64)[:5], w_counters[:5])
    assert np.all(np.abs(radiance[:, :3] - w_rad[:, :3]) <= 2e-5 * (1.0 + np.abs(w_rad[:, :3])))
    assert np.all(np.abs(color[:, :3] - w_color[:, :3]) <= 2e-5 * (1.0 + np.abs(w_color[:, :3])))
    return color


@pytest.mark.parametrize("entry,bounces", [(((1.0, 1.0, 1.0), 0.0, 1.0, 64), 4),
                                           (((0.25, 0.25, 0.25), 0.0, 0.5, 1), 2),
                                           (((0.8, 0.6, 0.4), -0.5, 0.7, 16), 3)])
def test_cube_matches_oracle(cx, oracle, entry, bounces):
    cam = look_down_from(2.0, 10.0, 2.0, 64)
    data = walk_scene(cube_world(), cam, {SSS_CODE: entry})
    print(f"sss cube {entry} bounces={bounces}", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = build(oracle, data)
    for seed in range(100, 104):
        compare(gpu.trace_frame(cam, bounces, seed, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY),
                ref.trace_frame(cam, bounces, seed, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY))


def test_one_step_closed_form_on_the_gpu(cx):
    cam = look_down_from(2.0, 10.0, 2.0, 64)
    data = walk_scene(cube_world(), cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 1)})
    print("sss one-step closed form", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    n_frames = 48
    c = np.stack([gpu.trace_frame(cam, 2, 500 + k, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY)[0].cpu().numpy()
                  .reshape(-1, 4)[:, 0] for k in range(n_frames)])
    assert np.all((c == 0.0) | (c == 1.0))
    p = one_step_survival(cam)
    sigma = np.sqrt((p * (1.0 - p)).sum() * n_frames) / (len(p) * n_frames)
    assert abs(c.mean() - p.mean()) < 5 * sigma, (c.mean(), p.mean(), sigma)


def test_bounce_walks_under_sun_match_oracle(cx, oracle):
    cam = look_down_from(48.0, 6.0, 32.0, 48)
    data = walk_scene(floor_block_world(), cam, {PLAIN_CODE: ((0.9, 0.9, 0.9), 0.3, 0.8, 32)})
    sky = cx.sky_params(sun_dir=(-0.6, 0.7, 0.2), sun_radiance=(30.0, 28.0, 25.0), sun_radius_deg=2.0,
                        zenith=(0.3, 0.4, 0.8), horizon=(0.7, 0.7, 0.7), ground=(0.2, 0.2, 0.2))
    print("sss bounce walks under the sun", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = build(oracle, data)
    for seed, bounces in ((7, 1), (8, 3), (9, 4)):
        compare(gpu.trace_frame(cam, bounces, seed, 1, flags=TRACE_SUBSURFACE, sky=sky),
                ref.trace_frame(cam, bounces, seed, 1, flags=TRACE_SUBSURFACE, sky=sky))


def test_dead_slab_matches_oracle(cx, oracle):
    from ref_fixtures import fill_box

    world: dict = {}
    fill_box(world, (0, 0, 0), (31, 15, 31), SSS_CODE)
    cam = look_down_from(16.0, 24.0, 16.0, 48)
    data = walk_scene(world, cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.9, 1.0, 1)})
    print("sss dead slab", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = build(oracle, data)
    colors = [compare(gpu.trace_frame(cam, 4, 700 + k, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY),
                      ref.trace_frame(cam, 4, 700 + k, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY)) for k in range(4)]
    assert np.stack(colors)[:, :, :3].mean() < 0.05


def test_unused_entry_changes_nothing_on_the_gpu(cx):
    cam = look_down_from(48.0, 6.0, 32.0, 48)
    print("sss unused entry", flush=True)
    plain = cx.TraceScene(walk_scene(floor_block_world(), cam, None), load_factor=LOAD)
    unused = cx.TraceScene(walk_scene(floor_block_world(), cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 8)}),
                           load_factor=LOAD)
    for seed in range(40, 43):
        want = [t.cpu().numpy() for t in plain.trace_frame(cam, 3, seed, 1, sky=WHITE_SKY)]
        got = [t.cpu().numpy() for t in unused.trace_frame(cam, 3, seed, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY)]
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_adjoint_refuses_the_walk(cx):
    import torch

    cam = look_down_from(2.0, 10.0, 2.0, 16)
    data = walk_scene(cube_world(), cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 8)})
    gpu = cx.TraceScene(data, load_factor=LOAD)
    d_color = torch.ones(16 * 16 * 4, dtype=torch.float32, device="cuda")
    ga = torch.zeros(256 * 4, dtype=torch.float32, device="cuda")
    gs = torch.zeros(len(data.keys), dtype=torch.float32, device="cuda")
    with pytest.raises(cx.CxError):
        gpu.trace_frame_adjoint(cam, 2, 1, 1, d_color, ga, gs, flags=TRACE_SUBSURFACE)
