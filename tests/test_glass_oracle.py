"""Pins the oracle's full material model (ShadeVoxelHit's transparent and specular branches, the
delta chain, media) without the GPU, on a glass slab (thickness 4, transparent id 0x42) lying on an
emissive floor (emission E, no NEE) under a constant sky C, camera above it, no bounces:

  a camera ray meets the slab's top face at incidence cos_i; it reflects with the Schlick probability
  F = r0 + (1 - r0)(1 - cos_i)^5, r0 = ((1 - ior)/(1 + ior))^2, and then leaves into the sky: color C,
  depth = the distance to the top face; or it refracts (Snell: sin_t = sin_i / ior), crosses the glass
  over L = (4 - push) / cos_t and ends on the floor: color = exp(-(1 - tint) extinction L) E per
  channel, depth = distance to the top + L. The share of reflections matches the mean F (5 sigma).
  ior 1: no bending, F = (1 - cos_i)^5, almost never a reflection.
  Rough glass (smoothness < 1, shaded stochastically) at one bounce gives exactly the colours of the
  smooth delta chain at none: the same random number decides the same interface.

The full model follows shaders/render/caelix_shading.hlsl:116-307 and caelix_path_trace.hlsl:595-978."""
from __future__ import annotations

import numpy as np
import pytest

import trace_checks as tc
from caelix_amd import scene

LOAD = 0.7
C = np.array([0.0, 0.0, 7.0])
E = 2.0
PUSH = 0.002


def build(oracle, data):
    return oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / LOAD)))


def constant_sky():
    return {"sun_dir": (0.0, 1.0, 0.0), "sun_radiance": (0.0, 0.0, 0.0), "sun_solid_angle": 0.0, "sun_cos_radius": 1.0,
            "sky_zenith": C, "sky_horizon": C, "sky_ground": C}


def slab_expectation(data, ior, extinction, tint):
    """Per pixel (float64): reflection probability, refracted colour, depths (reflected, refracted)."""
    d = tc.primary_dirs(data.camera).astype(np.float64)
    cos_i = -d[:, 1]
    r0 = ((1.0 - ior) / (1.0 + ior)) ** 2
    F = r0 + (1.0 - r0) * (1.0 - cos_i) ** 5
    sin_t = np.sqrt(1.0 - cos_i ** 2) / ior
    cos_t = np.sqrt(1.0 - sin_t ** 2)
    L = (4.0 - PUSH) / cos_t
    T = np.exp(-(1.0 - np.asarray(tint))[None, :] * extinction * L[:, None])
    t_top = (float(data.camera["position"][1]) - 12.0) / cos_i
    return F, T * E, t_top, t_top + L


@pytest.mark.parametrize("ior", [1.5, 1.0])
def test_glass_slab_closed_form(oracle, ior):
    tint, extinction = (0.9, 0.7, 0.5), 0.3
    data = scene.glass_slab(size_bricks=8, ior=ior, extinction=extinction, tint=tint)
    ref = build(oracle, data)
    F, refracted, depth_reflected, depth_refracted = slab_expectation(data, ior, extinction, tint)
    reflections = 0
    frames = 8
    for seed in range(frames):
        color, radiance, attrib, counters = ref.trace_frame(data.camera, 0, 100 + seed, 1, sky=constant_sky(),
                                                            bounces_transparent=2)
        assert np.all(attrib != 0) and np.all(radiance == 0.0)
        refl = color[:, 0] == 0.0
        reflections += int(refl.sum())
        assert np.all(color[refl, :3] == C[None, :])
        assert np.allclose(color[refl, 3], depth_reflected[refl], rtol=1e-5)
        assert np.allclose(color[~refl, :3], refracted[~refl], rtol=1e-4)
        assert np.allclose(color[~refl, 3], depth_refracted[~refl], rtol=1e-5)
        blocks = (attrib >> 16) & 0xFFFF
        assert np.all(blocks[refl] & 0x1FF == scene.GLASS) and np.all(blocks[~refl] == 0x8123)
        # one camera ray and one chain ray per pixel
        assert counters[0] == 2 * len(color)
    mean, var = F.sum() * frames, (F * (1 - F)).sum() * frames
    assert abs(reflections - mean) < 5 * np.sqrt(var) + 1, (reflections, mean, np.sqrt(var))
    if ior == 1.5:
        assert reflections > 500


def test_rough_glass_one_bounce_equals_delta_chain(oracle):
    smooth = scene.glass_slab(size_bricks=8, smoothness=1.0)
    rough = scene.glass_slab(size_bricks=8, smoothness=0.5)
    a = build(oracle, smooth).trace_frame(smooth.camera, 0, 17, 1, sky=constant_sky(), bounces_transparent=2)
    b = build(oracle, rough).trace_frame(rough.camera, 1, 17, 1, sky=constant_sky(), bounces_transparent=2)
    assert np.array_equal(a[0][:, :3], b[0][:, :3])
    # the rough slab is the primary surface: its depth is the distance to the top face
    _, _, depth_top, _ = slab_expectation(smooth, 1.5, 0.3, (0.9, 0.7, 0.5))
    assert np.allclose(b[0][:, 3], depth_top, rtol=1e-5)


def test_delta_budget_zero_shades_the_glass(oracle):
    """Budget 0: the slab's top face is the primary surface, shaded stochastically like rough glass."""
    data = scene.glass_slab(size_bricks=8)
    ref = build(oracle, data)
    color, _, attrib, counters = ref.trace_frame(data.camera, 1, 17, 1, sky=constant_sky(), bounces_transparent=0)
    assert np.all((attrib >> 16) & 0x1FF == scene.GLASS)
    chain = ref.trace_frame(data.camera, 0, 17, 1, sky=constant_sky(), bounces_transparent=2)
    assert np.array_equal(color[:, :3], chain[0][:, :3])


def test_glass_terrain_renders_finite(oracle):
    """The small terrain with media, rough glass and specular lobes: finite, and different from the
    opaque-only render of the same scene."""
    data = scene.glass_terrain()
    ref = build(oracle, data)
    color, radiance, attrib, counters = ref.trace_frame(data.camera, 3, 9, 1, sky=constant_sky(), bounces_transparent=4)
    assert np.isfinite(color).all() and np.isfinite(radiance).all()
    media = ((attrib >> 16) & 0x8000) == 0
    assert (media & (attrib != 0)).sum() > 100  # the chain ends on (rough) glass for some pixels


SLAB_TINT, SLAB_EXTINCTION = (0.8, 0.6, 0.4), 0.2
# the shadow ray enters the slab's underside and is pushed 0.002 through it along the normal
GLASS_PATH = 4.0 - PUSH


def sun_through_slab(L, r_deg, tint, extinction):
    """E[radiance] of the floor's sun NEE through the slab: 2 L int_0^r cos t sin t exp(-s d / cos t) dt
    per channel, s = (1 - tint) extinction, d the vertical glass path (the shadow ray goes straight
    through, no fog)."""
    r = np.radians(r_deg)
    t = (np.arange(400000) + 0.5) / 400000 * r
    sigma = (1.0 - np.asarray(tint))[:, None] * extinction
    f = np.cos(t) * np.sin(t) * np.exp(-sigma * GLASS_PATH / np.cos(t))
    return 2.0 * L * f.sum(axis=1) * (r / len(t))


def test_transmissive_sun_shadow_closed_form(oracle):
    """A floor under a floating glass slab (closed at the sides), the sun overhead (radius 10 degrees):
    with transmissive shadows the floor's radiance is the sun NEE through the slab (5 sigma + 1e-4);
    the claim keeps the refracted bounce rays that reach the sun disk from counting it again, so three
    bounces give exactly the radiance of one (without the claim about 3 % of them would add the disk)."""
    data = scene.sun_floor(glass_above=(SLAB_TINT, SLAB_EXTINCTION))
    ref = build(oracle, data)
    L, r_deg, frames = 100.0, 10.0, 8
    sky = {"sun_dir": (0.0, 1.0, 0.0), "sun_radiance": (L, L, L),
           "sun_solid_angle": 2 * np.pi * (1 - np.cos(np.radians(r_deg))), "sun_cos_radius": np.cos(np.radians(r_deg)),
           "sky_zenith": (0, 0, 0), "sky_horizon": (0, 0, 0), "sky_ground": (0, 0, 0)}
    cam = dict(data.camera)
    cam["width"] = cam["height"] = 96
    radiance, floor = [], []
    for seed in range(frames):
        _, rad, attrib, counters = ref.trace_frame(cam, 3, 5 + seed, 1, sky=sky, bounces_transparent=2,
                                                   flags=oracle.SHADOW_TRANSMISSION)
        _, one, _, _ = ref.trace_frame(cam, 1, 5 + seed, 1, sky=sky, bounces_transparent=2,
                                       flags=oracle.SHADOW_TRANSMISSION)
        blocks = (attrib >> 16) & 0xFFFF
        # the chain crossed the slab to the floor, or reflected (into the sky, or inside the slab until
        # the budget ended it on an interface)
        assert np.all((blocks == 0x8001) | ((blocks & 0x8000) == 0))
        on_floor = blocks == 0x8001
        assert np.array_equal(rad[on_floor], one[on_floor])
        radiance.append(rad[on_floor, :3])
        floor.append(on_floor)
    assert np.concatenate(floor).mean() > 0.9
    px = np.concatenate(radiance)
    want = sun_through_slab(L, r_deg, SLAB_TINT, SLAB_EXTINCTION)
    se = px.std(axis=0) / np.sqrt(len(px))
    assert np.all(np.abs(px.mean(axis=0) - want) < 5 * se + 1e-4 * want), (px.mean(axis=0), want, se)
    # without the transmission the slab blocks the floor's shadow rays
    _, blocked, attrib, _ = ref.trace_frame(cam, 1, 5, 1, sky=sky, bounces_transparent=2)
    assert np.all(blocked[((attrib >> 16) & 0xFFFF) == 0x8001, :3] == 0.0)


def checkerboard_expectation(data):
    """Per pixel (float64) with the delta checkerboard on the slab: px + py even refracts with weight
    2 (1 - F) (colour 2 (1 - F) T E, depth top + L), odd reflects with weight 2 F (colour 2 F C, depth
    top)."""
    F, refracted, depth_reflected, depth_refracted = slab_expectation(data, 1.5, 0.3, (0.9, 0.7, 0.5))
    w, h = data.camera["width"], data.camera["height"]
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    even = ((px + py).ravel() & 1) == 0
    color = np.where(even[:, None], 2 * (1 - F)[:, None] * refracted, 2 * F[:, None] * C[None, :])
    depth = np.where(even, depth_refracted, depth_reflected)
    return color, depth


def test_delta_checkerboard_closed_form(oracle):
    data = scene.glass_slab(size_bricks=8)
    ref = build(oracle, data)
    want_color, want_depth = checkerboard_expectation(data)
    for seed in (1, 2):
        color, _, attrib, _ = ref.trace_frame(data.camera, 0, seed, 1, sky=constant_sky(), bounces_transparent=2,
                                              flags=oracle.DELTA_CHECKERBOARD)
        assert np.allclose(color[:, :3], want_color, rtol=1e-4, atol=1e-7)
        assert np.allclose(color[:, 3], want_depth, rtol=1e-5)
    # the checkerboard's mean over pixel pairs is the random mode's expectation per pixel
    F, refracted, _, _ = slab_expectation(data, 1.5, 0.3, (0.9, 0.7, 0.5))
    expectation = (1 - F)[:, None] * refracted + F[:, None] * C[None, :]
    assert abs(want_color.mean() / expectation.mean() - 1) < 1e-2
