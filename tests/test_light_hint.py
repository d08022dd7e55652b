"""The light NEE's visibility hint (caelix_path_trace.hlsl:229-236, 338-413, 453-580), as the reference's
trace test checks it ("light NEE: visibility hint converges and reprojects", tests/trace_test.cpp:
3387-3580, and the hint variants of "light NEE: unbiased against the path tracer", :3100-3160), on the
lamp box (tests/ref_fixtures.py), on the oracle (CPU) and on the HIP path through the C ABI (gpu):

converges        with a small sun (0.25 degrees, the energy of the 2-degree test sun) the primary
                 vertices' direct sun is measured without light NEE (sun on minus sun off, same seeds,
                 one bounce); a lit and a shadowed pixel on upward floors with 5 x 5 neighbourhoods lit
                 or shadowed alike and no lamp in view; after 64 frames with the hint (patch 4) the sun
                 hint the lit pixel's primary vertex read is above 0.8, the shadowed one's below 0.2
reprojects       the camera turns by a pixel per frame for 8 frames (yaw about +Y, no parallax): the
                 points moved by more than 4 pixels and the hint read where they now lie stays within 0.2
patches          patches 1, 2 and 4 run with finite images; an invalid patch is rejected
update rule      under an open sun with no lamps every primary shadow ray reaches the sun: the sun hint of
                 every floor patch follows k <- round(255 (k / 255 + 0.2 (1 - k / 255))) from 128 exactly
                 (float64 closed form of the documented rule), the lamp hint stays 128
unbiased         the hint changes only the sun / lamp choice: the image over many frames with the hint
                 (patch 4 and 1) matches the path tracer (ratio of means 0.97-1.03, 16 x 16 block PSNR
                 above 30 dB, as the reference's light NEE variants)
GPU = oracle     every frame's hint image and read words bit for bit, hit attributes and counters
                 exact, colours within 2e-5 relative
"""
from __future__ import annotations

import numpy as np
import pytest

import ref_fixtures as rf
from caelix_amd import sky_params
from test_ref_scenes_oracle import SUN_DIR, build, compare_light, lamp_box_sky

SIZE = rf.LAMP_BOX_IMAGE
SMALL_SUN_DEG = 0.25
SMALL_SUN = 1000.0 * (2.0 / SMALL_SUN_DEG) ** 2
DIRECT_LAMP = 32.0
MARGIN = 2


def lowest_face(attrib: np.ndarray) -> np.ndarray:
    """The face bit of a hit attribute (its lowest set bit of 0x3F): 4 = +Y."""
    a = (attrib & 0x3F).astype(np.int64)
    return a & -a


def small_sun_sky(radiance: float):
    return sky_params(sun_dir=tuple(SUN_DIR), sun_radiance=(radiance,) * 3, sun_radius_deg=SMALL_SUN_DEG)


class OracleRenderer:
    def __init__(self, oracle, data, mis: bool = True):
        self.oracle = oracle
        self.s = build(oracle, data, mis=mis)

    def hint(self, patch: int = 4, width: int = SIZE, height: int = SIZE):
        return self.oracle.HintState(width, height, patch)

    def frame(self, cam, bounces, seed, spp, sky, hint=None):
        color, _, attrib, counters = self.s.trace_frame(cam, bounces, seed, spp, sky=sky, bounces_transparent=4,
                                                        hint=hint)
        return color.astype(np.float64), attrib.astype(np.uint32), counters


class GpuRenderer:
    def __init__(self, cx, data):
        self.cx = cx
        self.s = cx.TraceScene(data, load_factor=0.7)

    def hint(self, patch: int = 4, width: int = SIZE, height: int = SIZE):
        return self.cx.LightHint(width, height, patch, keep_read=True)

    def frame(self, cam, bounces, seed, spp, sky, hint=None):
        color, _, attrib, counters = self.s.trace_frame(cam, bounces, seed, spp, sky=sky, bounces_transparent=4,
                                                        hint=hint)
        return (color.cpu().numpy().reshape(-1, 4).astype(np.float64), attrib.cpu().numpy().view(np.uint32),
                counters.cpu().numpy().view(np.uint64))


def pixel_of_direction(cam, d):
    r = np.dot(d, cam["right"])
    u = np.dot(d, cam["up"])
    f = np.dot(d, cam["forward"])
    t = float(cam["tan_half_fov_y"])
    return (r / (f * t) + 1.0) * 0.5 * SIZE, (1.0 - u / (f * t)) * 0.5 * SIZE


def direction_of_pixel(cam, x, y):
    nx = (x + 0.5) / SIZE * 2.0 - 1.0
    ny = 1.0 - (y + 0.5) / SIZE * 2.0
    t = float(cam["tan_half_fov_y"])
    return (np.asarray(cam["right"], np.float64) * nx * t + np.asarray(cam["up"], np.float64) * ny * t +
            np.asarray(cam["forward"], np.float64))


def turned(cam, angle):
    f = np.asarray(cam["forward"], np.float64)
    g = (f[0] * np.cos(angle) + f[2] * np.sin(angle), f[1], -f[0] * np.sin(angle) + f[2] * np.cos(angle))
    return rf.look_along(cam["position"], g, float(cam["tan_half_fov_y"]), SIZE)


def lit_and_shadowed(plain):
    """Step 1 of the reference test: the direct sun of the primary vertices (light NEE off, one bounce,
    the sun on and off with the same seeds), a lit and a shadowed pixel on an upward floor whose 5 x 5
    neighbourhood is alike and sees no lamp, closest to the image centre."""
    cam = rf.lamp_box_camera()
    lamps, _, _ = plain.frame(cam, 1, 7000, 64, small_sun_sky(0.0))
    sun, attrib, _ = plain.frame(cam, 1, 7000, 64, small_sun_sky(SMALL_SUN))
    direct = ((sun - lamps)[:, 0] / 64.0).reshape(SIZE, SIZE)
    sees_lamp = (lamps[:, 0] / 64.0 > DIRECT_LAMP).reshape(SIZE, SIZE)
    up = (lowest_face(attrib) == 4).reshape(SIZE, SIZE)  # face +Y
    max_sun = direct[~sees_lamp].max()
    assert max_sun > 0.1
    best = {"lit": (None, 1 << 30), "dark": (None, 1 << 30)}
    for y in range(MARGIN, SIZE - MARGIN):
        for x in range(MARGIN, SIZE - MARGIN):
            win = np.s_[y - MARGIN:y + MARGIN + 1, x - MARGIN:x + MARGIN + 1]
            ok = up[win] & ~sees_lamp[win]
            d = abs(x - SIZE // 2) + abs(y - SIZE // 2)
            if np.all(ok & (direct[win] > 0.9 * max_sun)) and d < best["lit"][1]:
                best["lit"] = ((x, y), d)
            if np.all(ok & (direct[win] == 0.0)) and d < best["dark"][1]:
                best["dark"] = ((x, y), d)
    lit, dark = best["lit"][0], best["dark"][0]
    assert lit is not None and dark is not None
    return cam, lit, dark


def read_sun(hint, xy) -> float:
    return float(hint.read_words()[xy[1], xy[0]] & 0xFF) / 255.0


def converge_and_reproject(plain, lit_renderer, on_frame=None):
    cam, lit, dark = lit_and_shadowed(plain)
    sky = small_sun_sky(SMALL_SUN)
    hint = lit_renderer.hint(4)
    seed = 9000
    for _ in range(64):
        lit_renderer.frame(cam, 4, seed, 1, sky, hint)
        if on_frame is not None:
            on_frame(cam, seed, sky, hint)
        seed += 1
    lit_r, dark_r = read_sun(hint, lit), read_sun(hint, dark)
    print(f"hint after 64 frames: lit {lit} sun {lit_r:.3f}, shadowed {dark} sun {dark_r:.3f}", flush=True)
    assert lit_r > 0.8 and dark_r < 0.2
    lit_dir, dark_dir = direction_of_pixel(cam, *lit), direction_of_pixel(cam, *dark)
    pixel_angle = np.arctan(2.0 * float(cam["tan_half_fov_y"]) / SIZE)
    view = cam
    for k in range(1, 9):
        view = turned(cam, pixel_angle * k)
        lit_renderer.frame(view, 4, seed, 1, sky, hint)
        if on_frame is not None:
            on_frame(view, seed, sky, hint)
        seed += 1
    lx, ly = pixel_of_direction(view, lit_dir)
    sx, sy = pixel_of_direction(view, dark_dir)
    moved_lit, moved_dark = read_sun(hint, (int(lx), int(ly))), read_sun(hint, (int(sx), int(sy)))
    print(f"after 8 turned frames: lit at ({lx:.2f}, {ly:.2f}) sun {moved_lit:.3f}, shadowed at ({sx:.2f}, {sy:.2f}) "
          f"sun {moved_dark:.3f}", flush=True)
    assert abs(lx - lit[0]) > 4.0
    assert abs(moved_lit - lit_r) < 0.2 and abs(moved_dark - dark_r) < 0.2
    for patch in (1, 2, 4):
        h = lit_renderer.hint(patch)
        for _ in range(4):
            color, _, _ = lit_renderer.frame(cam, 4, seed, 1, sky, h)
            if on_frame is not None:
                on_frame(cam, seed, sky, h)
            seed += 1
            assert np.all(np.isfinite(color))


def sun_floor_update(renderer, frames: int = 12) -> int:
    """The update rule under an open sun: floor patches (every pixel's primary vertex sees the floor and
    the sun) follow the closed form; the lamp channel stays 128 (no lamp: no lamp shadow ray)."""
    data = renderer_data_sun_floor()
    cam = data.camera
    sky = sky_params(sun_dir=(0.3, 1.0, 0.2), sun_radiance=(50.0,) * 3, sun_radius_deg=1.0)
    r = renderer(data)
    hint = r.hint(2, cam["width"], cam["height"])
    expected = 128
    for f in range(frames):
        _, attrib, _ = r.frame(cam, 1, 300 + f, 1, sky, hint)
        expected = int(np.floor(min(max(expected / 255.0 + 0.2 * (1.0 - expected / 255.0), 0.0), 1.0) * 255.0 + 0.5))
        img = hint.current()
        floor = (lowest_face(attrib) == 4).reshape(cam["height"], cam["width"])
        patches = floor.reshape(cam["height"] // 2, 2, cam["width"] // 2, 2).all(axis=(1, 3))
        assert patches.sum() > 100
        assert np.all(img[patches, 0] == expected), (f, np.unique(img[patches, 0]), expected)
        assert np.all(img[..., 1] == 128)
    return expected


def renderer_data_sun_floor():
    from caelix_amd import scene

    return scene.sun_floor(size_bricks=8, lamp="below")


def unbiased(plain, lit_renderer, frames: int, patch: int):
    cam = rf.lamp_box_camera()
    sky = lamp_box_sky(1000.0)
    spp_pt = 4 * frames
    pt = plain.frame(cam, 4, 100000, spp_pt, sky)[0] / spp_pt
    hint = lit_renderer.hint(patch)
    total = np.zeros_like(pt)
    for i in range(frames):
        total += lit_renderer.frame(cam, 4, 200000 + i, 1, sky, hint)[0]
    ref_mean, ratio, psnr = compare_light(pt, total / frames)
    print(f"hint patch {patch}, {frames} frames: mean {ref_mean:.5f}, ratio {ratio:.4f}, block PSNR {psnr:.2f} dB",
          flush=True)
    assert ref_mean > 0.01 and 0.97 < ratio < 1.03 and psnr > 30.0


# ------------------------------------------------------------------------------------------- oracle


@pytest.fixture(scope="module")
def lamp_box_data():
    cam = rf.lamp_box_camera()
    return (rf.scene_from_world(rf.lamp_box(), cam, glass=True, lights=False),
            rf.scene_from_world(rf.lamp_box(), cam, glass=True))


def test_hint_converges_and_reprojects_oracle(oracle, lamp_box_data):
    plain, lit = lamp_box_data
    converge_and_reproject(OracleRenderer(oracle, plain), OracleRenderer(oracle, lit))


def test_hint_update_rule_oracle(oracle):
    assert sun_floor_update(lambda d: OracleRenderer(oracle, d)) > 200


def test_hint_unbiased_oracle(oracle, lamp_box_data):
    plain, lit = lamp_box_data
    unbiased(OracleRenderer(oracle, plain), OracleRenderer(oracle, lit), 256, 4)


# ------------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_hint_converges_and_reprojects_gpu(cx, oracle, lamp_box_data):
    """The reference test on the GPU, every frame against the oracle run alongside (the same hint
    history): hint images and read words bit for bit, the frames as the other frame tests compare."""
    plain, lit = lamp_box_data
    ref = OracleRenderer(oracle, lit)
    ref_hints = {}

    def check(cam, seed, sky, hint):
        key = (id(hint), hint.patch)
        if key not in ref_hints:
            ref_hints[key] = ref.hint(hint.patch)
        h = ref_hints[key]
        _, attrib, _ = ref.frame(cam, 4, seed, 1, sky, h)
        assert np.array_equal(hint.current(), h.current())
        assert np.array_equal(hint.read_words(), h.read_words())

    converge_and_reproject(GpuRenderer(cx, plain), GpuRenderer(cx, lit), check)


@pytest.mark.gpu
def test_hint_frames_equal_oracle_gpu(cx, oracle, lamp_box_data):
    from test_glass_gpu import compare

    _, lit = lamp_box_data
    cam = rf.lamp_box_camera()
    gpu = GpuRenderer(cx, lit)
    ref = OracleRenderer(oracle, lit)
    sky = lamp_box_sky(1000.0)
    for patch, spp in ((4, 1), (2, 3), (1, 1)):
        hg, ho = gpu.hint(patch), ref.hint(patch)
        for f in range(3):
            got = gpu.s.trace_frame(cam, 4, 500 + f, spp, sky=sky, bounces_transparent=4, hint=hg)
            want = ref.s.trace_frame(cam, 4, 500 + f, spp, sky=sky, bounces_transparent=4, hint=ho)
            compare(got, want)
            assert np.array_equal(hg.current(), ho.current()) and np.array_equal(hg.read_words(), ho.read_words())


@pytest.mark.gpu
def test_hint_update_rule_gpu(cx):
    assert sun_floor_update(lambda d: GpuRenderer(cx, d), frames=24) > 240


@pytest.mark.gpu
@pytest.mark.parametrize("patch", [4, 1])
def test_hint_unbiased_gpu(cx, lamp_box_data, patch):
    plain, lit = lamp_box_data
    unbiased(GpuRenderer(cx, plain), GpuRenderer(cx, lit), 1024, patch)


@pytest.mark.gpu
def test_hint_rejects_bad_arguments(cx, lamp_box_data):
    import ctypes

    _, lit = lamp_box_data
    gpu = GpuRenderer(cx, lit)
    cam = rf.lamp_box_camera()
    hint = gpu.hint(4)
    struct = hint.frame_struct(cam)
    struct.patch_shift = 3
    params = cx.CxTraceParams(4, 1, 1, 0)
    params.hint = ctypes.cast(ctypes.pointer(struct), ctypes.c_void_p)
    torch = cx._torch()
    color = torch.empty(SIZE * SIZE * 4, dtype=torch.float32, device="cuda")
    counters = torch.zeros(8, dtype=torch.int64, device="cuda")
    out = gpu.s._output(cam, color, None, None, counters)
    camera = cx.camera_struct(cam)
    rc = gpu.s.lib.cx_trace_frame(ctypes.byref(gpu.s.scene), ctypes.byref(gpu.s.view), ctypes.byref(camera),
                                  ctypes.byref(params), ctypes.byref(out), None)
    assert rc == 1  # CX_ERROR_INVALID_ARGUMENT
