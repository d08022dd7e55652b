"""GPU on the reference tests' own scenes (tests/ref_fixtures.py), through the C ABI: the same checks as
tests/test_ref_scenes_oracle.py on the HIP path, and the HIP path against the oracle on the same scenes
(hit attributes, depths and counters exact; color and radiance within 2e-5 relative). The lamp box runs
at the reference test's own sample counts (path tracer 4096 samples, light NEE 1024)."""
from __future__ import annotations

import numpy as np
import pytest

import ref_fixtures as rf
from test_glass_gpu import compare
from test_ref_scenes_oracle import (LAMP_VARIANTS, build, check_ceiling, check_red_brick, compare_light,
                                    lamp_box_sky, primary_rays, random_world)

pytestmark = pytest.mark.gpu

LOAD = 0.7


def host(frame):
    color, radiance, attrib, counters = frame
    return (color.cpu().numpy().astype(np.float64), radiance.cpu().numpy().astype(np.float64),
            attrib.cpu().numpy().view(np.uint32), counters.cpu().numpy().view(np.uint64))


def test_random_world_gpu(cx, oracle):
    cam = rf.look_plus_z(0.0, 0.0, -900.0)
    data = rf.scene_from_world(random_world(), cam, lights=False)
    print(f"random opaque world: {len(data.keys)} bricks, grid {data.grid_dim}", flush=True)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = build(oracle, data)
    compare(gpu.trace_frame(cam, 0, 1, 1), ref.trace_frame(cam, 0, 1, 1))
    rays = primary_rays(cam)
    got = gpu.trace_rays(rays)
    assert np.array_equal(got, ref.trace_rays(rays))
    hits = got[:, 1] != 0
    assert hits.sum() > 500 and np.all((got[hits, 1] >> 16) & 0xC000)


def test_red_brick_gpu(cx, oracle):
    cam = rf.look_plus_z(4.0, 4.0, -20.0)
    data = rf.scene_from_world(rf.red_brick(), cam, lights=False)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    frame = gpu.trace_frame(cam, 0, 1, 1)
    compare(frame, build(oracle, data).trace_frame(cam, 0, 1, 1))
    color, _, attrib, _ = host(frame)
    check_red_brick(color, attrib, data, cam)


def test_emissive_ceiling_gpu(cx, oracle):
    cam = rf.look_down()
    data = rf.scene_from_world(rf.ceiling_room(), cam)
    gpu = cx.TraceScene(data, load_factor=LOAD)
    ref = build(oracle, data)
    compare(gpu.trace_frame(cam, 1, 5, 4), ref.trace_frame(cam, 1, 5, 4))
    albedo = float(data.materials[rf.GREY_FLOOR, 0])
    c1, r1, _, _ = host(gpu.trace_frame(cam, 1, 77, 1))
    spp = 1024
    _, r, _, _ = host(gpu.trace_frame(cam, 1, 1000, spp))
    check_ceiling(r1, c1, r, spp, cam, albedo)


@pytest.fixture(scope="module")
def lamp_box(cx):
    cam = rf.lamp_box_camera()
    plain = rf.scene_from_world(rf.lamp_box(), cam, glass=True, lights=False)
    lit = rf.scene_from_world(rf.lamp_box(), cam, glass=True)
    return cam, plain, lit, cx.TraceScene(plain, load_factor=LOAD), cx.TraceScene(lit, load_factor=LOAD)


@pytest.mark.parametrize("sun", [0.0, 1000.0])
def test_lamp_box_gpu_matches_oracle(oracle, lamp_box, sun):
    cam, plain, lit, gpu_plain, gpu_lit = lamp_box
    sky = lamp_box_sky(sun)
    compare(gpu_plain.trace_frame(cam, 4, 7, 2, bounces_transparent=4, sky=sky),
            build(oracle, plain).trace_frame(cam, 4, 7, 2, bounces_transparent=4, sky=sky))
    for name, mis, q in LAMP_VARIANTS:
        print(f"lamp box sun={sun} {name} vs oracle", flush=True)
        gpu_lit.view.mis = 1 if mis else 0
        try:
            got = gpu_lit.trace_frame(cam, 4, 9, 2, bounces_transparent=4, sky=sky, bounce_nee_probability=q)
        finally:
            gpu_lit.view.mis = 1
        compare(got, build(oracle, lit, mis=mis).trace_frame(cam, 4, 9, 2, bounces_transparent=4, sky=sky,
                                                              bounce_nee_probability=q))


@pytest.mark.parametrize("sun", [0.0, 1000.0])
def test_lamp_box_gpu_light_nee_matches_path_tracer(lamp_box, sun):
    """The reference's acceptance test at its own sample counts: the path tracer over 4096 samples, each
    light NEE variant over 1024."""
    cam, _, _, gpu_plain, gpu_lit = lamp_box
    sky = lamp_box_sky(sun)
    pt = gpu_plain.trace_frame(cam, 4, 100000, 4096, bounces_transparent=4, sky=sky)[0].cpu().numpy() / 4096
    for name, mis, q in LAMP_VARIANTS:
        gpu_lit.view.mis = 1 if mis else 0
        try:
            nee = gpu_lit.trace_frame(cam, 4, 200000, 1024, bounces_transparent=4, sky=sky,
                                      bounce_nee_probability=q)[0].cpu().numpy() / 1024
        finally:
            gpu_lit.view.mis = 1
        ref_mean, ratio, psnr = compare_light(pt, nee)
        print(f"lamp box sun={sun} {name}: mean {ref_mean:.5f}, ratio {ratio:.4f}, block PSNR {psnr:.2f} dB", flush=True)
        assert ref_mean > 0.01, name
        assert 0.97 < ratio < 1.03, (name, ratio)
        assert psnr > 30.0, (name, psnr)
