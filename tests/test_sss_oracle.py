"""Pins the oracle's subsurface walk (shaders/render/caelix_subsurface.hlsl, docs/SUBSURFACE.md; the
reference's TracePath integration, caelix_path_trace.hlsl:729-978) without the GPU:

HG phase function   SampleHenyeyGreenstein: unit directions whose cosine about w has the Henyey-Greenstein
                    distribution (mean g within 5 sigma, Kolmogorov-Smirnov distance against the HG cdf)
                    for g in {0, 0.3, -0.6, 0.9}; a wrong sign or a dropped term of the inversion fails it
furnace             the reference's "subsurface: furnace" (tests/trace_test.cpp:4602-4653): a 4^3 SSS cube
                    (sss_albedo 1, p 1, g 0) under a white sky, the camera 6 above its top face: a pixel is
                    exactly 1 (the walk left: guide albedo lerp(0.5, 1, 1) = 1, an exit vertex sees only
                    the sky) or 0 (dead); steps 64: mean >= 0.95; steps 2: lower by more than 0.05 and
                    above 0.1; the flag off and sss_model 0: exactly the plain albedo 0.5
one-step walk       steps 1, g 0, a closed form: the walk survives when its one uniform direction's dominant
                    signed axis leads out of the cube from the start voxel (probability 1/6 per face of the
                    start voxel on the cube's surface, the top one always), the start voxel being the hit
                    voxel moved by the face jitter when that stays in the cube; the mean over frames within
                    5 sigma of the float64 expectation of every pixel, also on the pixels the jitter moves
branch weights      p 0.5, sss_albedo 0.25 on albedo 0.5: the guide albedo lerp(0.5, 0.25, 0.5) = 0.375 is the
                    albedo output, the branches carry 0.5 / 0.375 (surface) and 0.25 / 0.375 (a walk that
                    left) as their demodulated radiance; the mean colour by the one-step closed form
lifecycle           trace_test.cpp:4655-4731: an SSS entry on a code no voxel uses changes nothing (bit for
                    bit, the full model's kernels against the plain ones); the block beside the camera's
                    floor becomes SSS (p 0.8, sss_albedo 0.9, g 0.3, steps 32): the image changes (walks at
                    bounce vertices), stays finite, mean above 0.1; dead walks in a 16-thick slab (steps 1,
                    g 0.9): the top face's mean below 0.05
"""
from __future__ import annotations

import numpy as np
import pytest

import ref_fixtures as rf
import trace_checks as tc
from caelix_amd import TRACE_SUBSURFACE, sky_params
from caelix_amd.scene import SSS_MODEL_HG, subsurface_table

LOAD = 0.7
SSS_CODE = 0x8000 | (10 << 10) | (20 << 5) | 5      # kSssCode (trace_test.cpp:4168)
PLAIN_CODE = 0x8000 | (20 << 10) | (10 << 5) | 5    # kPlainCode
WHITE_FLOOR = 0x8000 | (31 << 10) | (31 << 5) | 31  # kWhiteFloor
WHITE_SKY = sky_params(zenith=(1.0, 1.0, 1.0), horizon=(1.0, 1.0, 1.0), ground=(1.0, 1.0, 1.0))


def look_down_from(x: float, height: float, z: float, size: int) -> dict:
    """LookDownFrom (trace_test.cpp:1568-1575): the look-down camera moved to (x, height, z)."""
    cam = rf.look_down(size)
    cam["position"] = np.array([x, height, z], np.float32)
    return cam


def walk_scene(world: dict, camera: dict, entries: dict | None):
    """The world with the packed materials, the SSS and plain codes at albedo 0.5 (WalkEntry,
    trace_test.cpp:4553-4566), and the subsurface table of `entries` (None: no table)."""
    data = rf.scene_from_world(world, camera, lights=False)
    for code in (SSS_CODE, PLAIN_CODE):
        data.materials[code, 0:3] = 0.5
    if entries is not None:
        data.subsurface = subsurface_table(entries)
    return data


def cube_world() -> dict:
    world: dict = {}
    rf.fill_box(world, (0, 0, 0), (3, 3, 3), SSS_CODE)
    return world


def build(oracle, data):
    return oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / LOAD)))


def frames(ref, cam, bounces: int, seeds, flags: int) -> np.ndarray:
    """[frames, pixels, 3] colours."""
    out = []
    for seed in seeds:
        color, radiance, attrib, _ = ref.trace_frame(cam, bounces, seed, 1, flags=flags, sky=WHITE_SKY)
        assert np.all(np.isfinite(color)) and np.all(np.isfinite(radiance))
        out.append(color.reshape(-1, 4)[:, :3])
    return np.stack(out)


# ------------------------------------------------------------------------------------------ HG


def hg_cdf(mu: np.ndarray, g: float) -> np.ndarray:
    """P(cos theta <= mu) of the Henyey-Greenstein phase function."""
    if abs(g) < 1e-3:
        return (mu + 1.0) / 2.0
    return (1.0 - g * g) / (2.0 * g) * (1.0 / np.sqrt(1.0 + g * g - 2.0 * g * mu) - 1.0 / (1.0 + g))


@pytest.mark.parametrize("g", [0.0, 0.3, -0.6, 0.9])
def test_hg_phase_function(oracle, g):
    rng = np.random.default_rng(7)
    n = 200_000
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    q = np.concatenate([w, np.full((n, 1), g)], axis=1).astype(np.float32)
    d = oracle.sss_hg(q, 12345).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    mu = np.sum(d * q[:, :3].astype(np.float64), axis=1)
    # E[mu] = g, Var[mu] = (1 + 2 g^2) / 3 - g^2 for HG
    sigma = np.sqrt(((1.0 + 2.0 * g * g) / 3.0 - g * g) / n)
    assert abs(mu.mean() - g) < 5 * sigma, (mu.mean(), g)
    s = np.sort(mu)
    ks = np.max(np.abs(hg_cdf(s, g) - np.arange(1, n + 1) / n))
    assert ks < 1.63 / np.sqrt(n) * 1.5, ks
    # the azimuth is uniform: the component along a fixed axis orthogonal to w averages 0
    t = np.cross(q[:, :3].astype(np.float64), [0.0, 0.0, 1.0])
    ok = np.linalg.norm(t, axis=1) > 0.5
    t = t[ok] / np.linalg.norm(t[ok], axis=1, keepdims=True)
    assert abs(np.sum(d[ok] * t, axis=1).mean()) < 5 * np.sqrt(1.0 / 3.0 / ok.sum())


# ------------------------------------------------------------------------------------------ furnace


@pytest.fixture(scope="module")
def cube_camera():
    return look_down_from(2.0, 10.0, 2.0, 64)


def test_furnace_walks_leave_or_die(oracle, cube_camera):
    data = walk_scene(cube_world(), cube_camera, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 64)})
    ref = build(oracle, data)
    c = frames(ref, cube_camera, 4, range(100, 132), TRACE_SUBSURFACE)
    assert np.all((c == 0.0) | (c == 1.0)), np.unique(c)
    sss = c.mean()
    assert 0.95 <= sss <= 1.0 + 1e-4, sss
    # the step cap works: steps 2 leave a visible dead fraction
    data.subsurface = subsurface_table({SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 2)})
    capped = frames(build(oracle, data), cube_camera, 4, range(200, 232), TRACE_SUBSURFACE)
    assert np.all((capped == 0.0) | (capped == 1.0))
    assert 0.1 < capped.mean() < sss - 0.05, (capped.mean(), sss)


def test_furnace_without_the_walk_is_the_plain_surface(oracle, cube_camera):
    data = walk_scene(cube_world(), cube_camera, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 64)})
    off = frames(build(oracle, data), cube_camera, 4, range(300, 304), 0)
    assert np.all(off == 0.5)
    data.subsurface = subsurface_table({SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 64, 0)})  # sss_model 0
    none = frames(build(oracle, data), cube_camera, 4, range(300, 304), TRACE_SUBSURFACE)
    assert np.all(none == 0.5)


def one_step_survival(cam, jitter_on: bool = True) -> np.ndarray:
    """Per pixel, the float64 probability that a one-step, g = 0 walk entered at the cube's top face
    survives (the start voxel and the uniform direction's dominant signed axis); jitter_on False: the
    walk started in the hit voxel."""
    d = tc.primary_dirs(cam).astype(np.float64)
    o = np.asarray(cam["position"], np.float64)
    t = (o[1] - 4.0) / -d[:, 1]
    hx, hz = o[0] + d[:, 0] * t, o[2] + d[:, 2] * t
    assert np.all((hx > 0.0) & (hx < 4.0) & (hz > 0.0) & (hz < 4.0))

    def jitter(h):
        v = np.floor(h)
        f = h - v
        lo, hi = np.maximum(0.0, 0.5 - f), np.maximum(0.0, f - 0.5)
        if not jitter_on:
            lo, hi = 0.0 * f, 0.0 * f
        return v, {-1: lo, 0: 1.0 - lo - hi, 1: hi}

    vx, px = jitter(hx)
    vz, pz = jitter(hz)
    p = np.zeros(len(d))
    for ex, wx in px.items():
        for ez, wz in pz.items():
            sx, sz = vx + ex, vz + ez
            inside = (sx >= 0) & (sx <= 3) & (sz >= 0) & (sz <= 3)
            sx, sz = np.where(inside, sx, vx), np.where(inside, sz, vz)
            faces = 1.0 + (sx == 0) + (sx == 3) + (sz == 0) + (sz == 3)
            p += wx * wz * faces / 6.0
    return p


def test_one_step_walk_closed_form(oracle, cube_camera):
    data = walk_scene(cube_world(), cube_camera, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 1)})
    ref = build(oracle, data)
    n_frames = 48
    c = frames(ref, cube_camera, 2, range(500, 500 + n_frames), TRACE_SUBSURFACE)[:, :, 0]
    assert np.all((c == 0.0) | (c == 1.0))
    p = one_step_survival(cube_camera)
    expect = p.mean()
    sigma = np.sqrt((p * (1.0 - p)).sum() * n_frames) / (len(p) * n_frames)
    assert abs(c.mean() - expect) < 5 * sigma, (c.mean(), expect, sigma)
    # per pixel too, where the start voxel decides: the corner columns survive more than the centre
    per_pixel = c.mean(axis=0)
    corner, centre = p > 0.4, p < 0.2
    assert per_pixel[corner].mean() > per_pixel[centre].mean() + 0.1
    # the entry jitter: on the pixels whose start voxel it moves toward the cube's sides (hits near the
    # inner voxel planes), the survival follows the jittered start, not the hit voxel
    p0 = one_step_survival(cube_camera, jitter_on=False)
    for moved in (p - p0 > 0.03, p - p0 < -0.03):
        assert moved.sum() > 100
        got = c[:, moved].mean()
        sigma = np.sqrt((p[moved] * (1.0 - p[moved])).sum() * n_frames) / (moved.sum() * n_frames)
        assert abs(got - p[moved].mean()) < 5 * sigma, (got, p[moved].mean(), p0[moved].mean())
        assert abs(got - p0[moved].mean()) > 5 * sigma


def test_branch_weights_and_guide_albedo(oracle, cube_camera):
    """p 0.5, sss_albedo 0.25 on albedo 0.5, one step: the guide albedo lerp(0.5, 0.25, 0.5) = 0.375 is
    the albedo output; the surface branch carries 0.5 / 0.375 into the sky (colour 0.5), a walk that leaves
    carries its weight 0.25 / 0.375 (colour 0.25), a dead one 0. The mean is 0.5 * 0.5 + 0.5 * 0.25 * P(walk
    survives) (the one-step closed form)."""
    data = walk_scene(cube_world(), cube_camera, {SSS_CODE: ((0.25, 0.25, 0.25), 0.0, 0.5, 1)})
    ref = build(oracle, data)
    n_frames = 48
    colors, radiances = [], []
    for seed in range(900, 900 + n_frames):
        color, radiance, _, _ = ref.trace_frame(cube_camera, 2, seed, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY)
        colors.append(color.reshape(-1, 4)[:, 0])
        radiances.append(radiance.reshape(-1, 4)[:, 0])
    c, r = np.stack(colors), np.stack(radiances)
    f32 = np.float32
    bsdf, walked = f32(0.5) / f32(0.375), f32(0.25) / f32(0.375)
    assert np.all(np.isin(r, [0.0, float(bsdf), float(walked)])), np.unique(r)
    assert np.allclose(c[r == float(bsdf)], 0.5) and np.allclose(c[r == float(walked)], 0.25) and np.all(c[r == 0] == 0)
    s = one_step_survival(cube_camera)
    pix = 0.5 * 0.5 + 0.5 * 0.25 * s
    var = 0.5 * 0.25 + 0.5 * s * 0.0625 - pix ** 2
    sigma = np.sqrt(var.sum() * n_frames) / (len(s) * n_frames)
    assert abs(c.mean() - pix.mean()) < 5 * sigma, (c.mean(), pix.mean(), sigma)


# ------------------------------------------------------------------------------------------ lifecycle


def floor_block_world() -> dict:
    world: dict = {}
    rf.fill_box(world, (0, -8, 0), (63, -1, 63), WHITE_FLOOR)
    rf.fill_box(world, (24, 0, 24), (39, 15, 39), PLAIN_CODE)
    return world


def test_unused_sss_entry_changes_nothing(oracle):
    cam = look_down_from(48.0, 6.0, 32.0, 32)
    plain = build(oracle, walk_scene(floor_block_world(), cam, None))
    unused = build(oracle, walk_scene(floor_block_world(), cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.0, 1.0, 8)}))
    for seed in range(40, 43):
        want = plain.trace_frame(cam, 3, seed, 1, sky=WHITE_SKY)
        for flags in (0, TRACE_SUBSURFACE):
            got = unused.trace_frame(cam, 3, seed, 1, flags=flags, sky=WHITE_SKY)
            for a, b in zip(got, want):
                assert np.array_equal(a, b)


def test_walks_at_bounce_vertices(oracle):
    cam = look_down_from(48.0, 6.0, 32.0, 32)
    data = walk_scene(floor_block_world(), cam, {PLAIN_CODE: ((0.9, 0.9, 0.9), 0.3, 0.8, 32)})
    ref = build(oracle, data)
    off = frames(ref, cam, 3, range(500, 516), 0)
    on = frames(ref, cam, 3, range(500, 516), TRACE_SUBSURFACE)
    assert not np.array_equal(on, off)
    assert on.mean() > 0.1
    # the camera sees only the floor: the primary vertices are the same
    attrib_on = ref.trace_frame(cam, 3, 500, 1, flags=TRACE_SUBSURFACE, sky=WHITE_SKY)[2]
    attrib_off = ref.trace_frame(cam, 3, 500, 1, sky=WHITE_SKY)[2]
    assert np.array_equal(attrib_on, attrib_off) and np.all((attrib_on >> 16) == WHITE_FLOOR)


def test_dead_walks_in_a_thick_slab(oracle):
    world: dict = {}
    rf.fill_box(world, (0, 0, 0), (31, 15, 31), SSS_CODE)
    cam = look_down_from(16.0, 24.0, 16.0, 32)
    data = walk_scene(world, cam, {SSS_CODE: ((1.0, 1.0, 1.0), 0.9, 1.0, 1)})
    slab = frames(build(oracle, data), cam, 4, range(700, 716), TRACE_SUBSURFACE)
    assert slab.mean() < 0.05, slab.mean()


def test_subsurface_table_packs_halves():
    t = subsurface_table({SSS_CODE: ((0.9, 0.5, 0.1), 0.3, 0.8, 300)})
    e = t[SSS_CODE]
    assert e["model"] == SSS_MODEL_HG and e["steps"] == 300 & 0xFF
    assert np.array_equal(e["one_minus_albedo"], np.float16(1.0 - np.float32([0.9, 0.5, 0.1])).astype(np.float32))
    assert e["probability"] == np.float32(np.float16(0.8)) and e["anisotropy"] == np.float32(np.float16(0.3))
    assert np.all(t["model"][np.arange(65536) != SSS_CODE] == 0)
