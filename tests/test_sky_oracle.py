"""Pins the oracle's procedural sky and sun (CX_TRACE_SKY) without the GPU, on a flat floor under an
open sky where the expected radiance has a closed form:

  sun straight overhead, sun NEE     E[radiance] = L_sun sin^2 r (the cone's cosine-weighted solid
                                     angle / pi: Omega (1 + cos r) / (2 pi) = 1 - cos^2 r), the same
                                     for 1 and 4 bounces (the bounce ray leaves into a black sky)
  sun NEE off                        the bounce ray finds the disk with probability sin^2 r: the same
                                     expectation (statistical)
  constant sky C, no sun             radiance = C exactly (every bounce ray misses with throughput 1)
  graded sky, no sun                 E[radiance] = H + 0.8 (Z - H): a cosine-weighted direction has
                                     cos theta = sqrt(u), so E[sqrt(cos theta)] = E[u^(1/4)] = 4/5
  primary misses                     color = ProceduralSky(dir) + the sun disk, recomputed in numpy
  sun + lamp (one-sample RIS)        E[both] = E[lamp alone] + E[sun alone] (unbiasedness of the
                                     choice weight w_sum / w_picked; statistical)

and the fog of the air (density s, radiance F):

  camera segment, 0 bounces          color = F (1 - exp(-s t)) at the pixel's depth t
  camera segment + constant sky C    color = F (1 - T) + T albedo C, radiance = C (bounce rays miss:
                                     no fog on them)
  isothermal fogged furnace          walls of emission E and albedo a in fog of radiance F = E / (1 - a):
                                     every segment adds theta F (1 - a T) and passes theta a T on, so
                                     the pixel is F (1 - a T_0 prod_j a T_j), F within float rounding at
                                     24 bounces without roulette, whatever the segment lengths

The procedural sky and sun follow shaders/render/caelix_shading.hlsl:364-485 and
caelix_path_trace.hlsl:276-290, 356-413, 737-846; the fog caelix_shading.hlsl:150-175 and
caelix_path_trace.hlsl:1273-1283."""
from __future__ import annotations

import numpy as np
import pytest

import trace_checks as tc
from caelix_amd import scene

LOAD = 0.7


def sky(sun_dir=(0.0, 1.0, 0.0), sun_radiance=(0.0, 0.0, 0.0), radius_deg=0.5, zenith=(0, 0, 0), horizon=(0, 0, 0),
        ground=(0, 0, 0)) -> dict:
    d = np.asarray(sun_dir, np.float64)
    d = d / np.linalg.norm(d)
    r = np.radians(radius_deg)
    return {"sun_dir": d, "sun_radiance": sun_radiance, "sun_solid_angle": 2 * np.pi * (1 - np.cos(r)),
            "sun_cos_radius": np.cos(r), "sky_zenith": zenith, "sky_horizon": horizon, "sky_ground": ground}


def build(oracle, data):
    return oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / LOAD)))


@pytest.fixture(scope="module")
def floor(oracle):
    data = scene.sun_floor(lamp="below")
    return data, build(oracle, data)


def test_sky_params_helper_matches(cx):
    """The package's sky_params gives the fields this file's helper gives."""
    got = cx.sky_params(sun_dir=(1.0, 2.0, 0.5), sun_radiance=(3.0, 2.0, 1.0), sun_radius_deg=1.5,
                        zenith=(0.1, 0.2, 0.3), horizon=(0.4, 0.5, 0.6), ground=(0.7, 0.8, 0.9))
    want = sky((1.0, 2.0, 0.5), (3.0, 2.0, 1.0), 1.5, (0.1, 0.2, 0.3), (0.4, 0.5, 0.6), (0.7, 0.8, 0.9))
    for k, v in want.items():
        assert np.allclose(np.asarray(got[k], np.float64), np.asarray(v, np.float64), rtol=1e-12, atol=0), k


@pytest.mark.parametrize("bounces", [1, 4])
def test_sun_floor_closed_form(floor, bounces):
    data, ref = floor
    L, r_deg, spp = 1000.0, 2.0, 4
    k = sky(sun_radiance=(L, L, L), radius_deg=r_deg)
    color, radiance, attrib, counters = ref.trace_frame(data.camera, bounces, 7, spp, sky=k)
    assert np.all(attrib != 0)
    albedo = np.float32(123 / 255)
    rad = radiance[:, 0] / spp
    want = L * np.sin(np.radians(r_deg)) ** 2
    # cos_l lies in [cos r, 1]: per-sample relative spread below 2e-4, the mean is far tighter
    assert abs(rad.mean() - want) < 2e-5 * want, (rad.mean(), want)
    assert abs(color[:, 0].mean() / spp - float(albedo) * want) < 2e-5 * want
    assert counters[1] == data.camera["width"] * data.camera["height"] * spp  # one sun shadow ray per pixel


def test_sun_floor_bounces_agree_and_zero_without_nee(floor):
    data, ref = floor
    k = sky(sun_radiance=(1000.0, 800.0, 600.0), radius_deg=2.0)
    one = ref.trace_frame(data.camera, 1, 11, 2, sky=k)
    four = ref.trace_frame(data.camera, 4, 11, 2, sky=k)
    assert np.array_equal(one[1], four[1]) and np.array_equal(one[0], four[0])
    zero = ref.trace_frame(data.camera, 0, 11, 2, sky=k)
    assert np.all(zero[1] == 0.0) and np.all(zero[0][:, :3] == 0.0)


def test_sun_disk_seen_by_bounce_rays_without_sun_nee(oracle, floor):
    data, ref = floor
    L, r_deg, spp = 100.0, 10.0, 16
    k = sky(sun_radiance=(L, L, L), radius_deg=r_deg)
    cam = dict(data.camera)
    cam["width"] = cam["height"] = 96
    _, radiance, _, counters = ref.trace_frame(cam, 1, 3, spp, sky=k, flags=oracle.NO_SUN_NEE)
    assert counters[1] == 0  # no shadow ray: the lamp is out of reach and the sun is not sampled
    samples = radiance[:, 0] / spp
    want = L * np.sin(np.radians(r_deg)) ** 2
    se = samples.std() / np.sqrt(len(samples))
    assert se > 0 and abs(samples.mean() - want) < 5 * se, (samples.mean(), want, se)
    # every sample is 0 or L (a bounce ray in the disk): pixel sums are multiples of L
    sums = radiance[:, 0]
    assert np.allclose(sums / L, np.round(sums / L), atol=1e-9)


@pytest.mark.parametrize("bounces", [1, 3])
def test_constant_sky_floor_exact(floor, bounces):
    data, ref = floor
    c = (0.3, 0.5, 0.7)
    color, radiance, attrib, counters = ref.trace_frame(data.camera, bounces, 5, 1, sky=sky(zenith=c, horizon=c, ground=c))
    assert np.array_equal(radiance[:, :3], np.tile(np.float32(c).astype(np.float64), (len(radiance), 1)))
    albedo = np.float64(np.float32(123 / 255))
    assert np.allclose(color[:, :3], albedo * np.float32(c).astype(np.float64), rtol=1e-15, atol=0)
    assert counters[1] == 0  # no lamp in reach, no sun: no shadow ray


def test_graded_sky_floor_expectation(floor):
    data, ref = floor
    zenith, horizon = np.array([0.2, 0.4, 1.0]), np.array([1.0, 0.8, 0.6])
    spp = 8
    cam = dict(data.camera)
    cam["width"] = cam["height"] = 96
    _, radiance, _, _ = ref.trace_frame(cam, 2, 21, spp, sky=sky(zenith=zenith, horizon=horizon, ground=(5, 5, 5)))
    samples = radiance[:, :3] / spp
    want = horizon + 0.8 * (zenith - horizon)
    se = samples.std(axis=0) / np.sqrt(len(samples))
    assert np.all(np.abs(samples.mean(axis=0) - want) < 5 * se + 1e-6), (samples.mean(axis=0), want, se)
    # without the square root the expectation would be H + (2/3)(Z - H): far outside the band
    wrong = horizon + (2 / 3) * (zenith - horizon)
    assert np.all(np.abs(wrong - want) > 50 * se)


def test_primary_miss_shows_sky_and_sun_disk(floor):
    data, ref = floor
    sun_dir = np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2])
    position = np.array([64.0, 40.0, 64.0])
    cam = scene.look_at(position, position + sun_dir * 100, 64, 64, fov_y_deg=30.0)
    zenith, horizon, ground = (0.1, 0.3, 0.9), (0.8, 0.7, 0.6), (0.2, 0.1, 0.0)
    k = sky(sun_dir=sun_dir, sun_radiance=(50.0, 40.0, 30.0), radius_deg=2.0, zenith=zenith, horizon=horizon,
            ground=ground)
    color, radiance, attrib, counters = ref.trace_frame(cam, 2, 1, 1, sky=k)
    assert np.all(attrib == 0) and np.all(radiance == 0.0) and np.all(color[:, 3] == 0.0)
    assert counters[1] == 0
    d = tc.primary_dirs(cam).astype(np.float64)
    y = d[:, 1:2]
    want = np.where(y >= 0, np.asarray(horizon) + (np.asarray(zenith) - horizon) * np.sqrt(np.clip(y, 0, 1)),
                    np.asarray(horizon) + (np.asarray(ground) - horizon) * np.sqrt(np.clip(-y, 0, 1)))
    cos_sun = d @ np.float32(sun_dir).astype(np.float64)
    margin = np.abs(cos_sun - np.cos(np.radians(2.0))) > 1e-6
    in_disk = cos_sun > np.cos(np.radians(2.0))
    want = want + np.where(in_disk[:, None], np.array([50.0, 40.0, 30.0]), 0.0)
    assert in_disk.sum() > 20  # the disk covers a few dozen pixels
    assert np.allclose(color[margin, :3], want[margin], rtol=1e-5, atol=1e-6)


def test_sun_and_lamp_choice_is_unbiased(oracle):
    """One shadow ray per vertex, sun or lamp kept in proportion to its unoccluded luminance and
    weighted by w_sum / w_kept: the image mean with both lights equals the sum of the means with each
    alone (the lamp switched off by an emission scale of 0), to five standard errors."""
    data = scene.sun_floor(lamp="above", lamp_emission=30.0)
    ref = build(oracle, data)
    cam = dict(data.camera)
    cam["width"] = cam["height"] = 96
    spp = 8
    k = sky(sun_dir=(0.4, 1.0, 0.3), sun_radiance=(200.0, 200.0, 200.0), radius_deg=3.0)
    lamp_brick = int(data.emissive_bricks[0])

    def mean_se(**kw):
        _, rad, _, _ = ref.trace_frame(cam, 1, 41, spp, **kw)
        px = rad[:, 0] / spp
        return px.mean(), px.std() / np.sqrt(len(px))

    both, se_b = mean_se(sky=k)
    lamp, se_l = mean_se()
    scale = np.ones(len(data.keys), np.float32)
    scale[lamp_brick] = 0.0
    ref.set_parameters(brick_scale=scale)
    try:
        sun, se_s = mean_se(sky=k)
    finally:
        ref.set_parameters(brick_scale=None)
    tol = 5 * np.sqrt(se_b ** 2 + se_l ** 2 + se_s ** 2)
    assert lamp > 10 * tol and sun > 10 * tol  # both lights matter
    assert abs(both - (lamp + sun)) < tol, (both, lamp, sun, tol)


FOG = (np.array([0.5, 0.6, 0.7]), 0.01)


@pytest.mark.parametrize("density", [1e-4, 3e-3, 0.01, 0.04])
def test_fog_camera_segment_closed_form(floor, density):
    """Also pins the exponential (density x depth from 0.011 to 5) to 1e-6 relative."""
    data, ref = floor
    color, radiance, attrib, _ = ref.trace_frame(data.camera, 0, 3, 1, fog=(FOG[0], density))
    assert np.all(attrib != 0) and np.all(radiance == 0.0)
    t = color[:, 3]
    assert t.min() > 100.0  # the camera is 112 voxels above the floor
    want = FOG[0][None, :] * (1.0 - np.exp(-density * t))[:, None]
    # 1 - T is taken in float32: a few ulps of T (6e-8 each) absolute, then 1e-6 relative
    assert np.all(np.abs(color[:, :3] - want) <= 1e-6 * want + 3e-7 * FOG[0][None, :])


def test_fog_camera_segment_with_constant_sky(floor):
    data, ref = floor
    c = np.array([0.3, 0.5, 0.7])
    color, radiance, _, _ = ref.trace_frame(data.camera, 1, 3, 1, sky=sky(zenith=c, horizon=c, ground=c), fog=FOG)
    assert np.array_equal(radiance[:, :3], np.tile(np.float32(c).astype(np.float64), (len(radiance), 1)))
    T = np.exp(-FOG[1] * color[:, 3])[:, None]
    albedo = float(np.float32(123 / 255))
    want = FOG[0][None, :] * (1.0 - T) + T * albedo * c[None, :]
    assert np.allclose(color[:, :3], want, rtol=2e-6, atol=0)


def isothermal_furnace():
    """The furnace box with walls of emission 1 and albedo 0.5 that are not lights (no NEE)."""
    data = scene.furnace_box(width=64, height=64)
    data.light_codes = np.zeros_like(data.light_codes)
    data.emissive_bricks = np.zeros(0, np.int64)
    return data


def test_fog_isothermal_furnace(oracle):
    data = isothermal_furnace()
    ref = build(oracle, data)
    F = 1.0 / (1.0 - 0.5)
    fog = ((F, F, F), 0.02)
    color, radiance, attrib, counters = ref.trace_frame(data.camera, 24, 8, 1, flags=oracle.NO_RUSSIAN_ROULETTE, fog=fog)
    assert np.all(attrib != 0) and counters[1] == 0
    assert np.all(np.abs(color[:, :3] - F) < 1e-5 * F), np.abs(color[:, :3] - F).max()


def test_fog_isothermal_furnace_one_bounce(oracle):
    """One bounce: radiance R = F (1 - a T1) and color = F (1 - a^2 T0 T1), T0 = exp(-s depth). Both
    outputs give the same bounce-segment transmission T1, inside [exp(-s * the box diagonal), 1];
    without the fog R would be E = 1 and color 1.5."""
    data = isothermal_furnace()
    ref = build(oracle, data)
    a, s_ = 0.5, 0.02
    F = 1.0 / (1.0 - a)
    color, radiance, _, _ = ref.trace_frame(data.camera, 1, 8, 1, flags=oracle.NO_RUSSIAN_ROULETTE,
                                            fog=((F, F, F), s_))
    T0 = np.exp(-s_ * color[:, 3])
    t1_from_radiance = (F - radiance[:, 0]) / (F * a)
    t1_from_color = (F - color[:, 0]) / (F * a * a * T0)
    assert np.allclose(t1_from_radiance, t1_from_color, rtol=2e-5, atol=2e-6)
    diagonal = 48 * np.sqrt(3.0)
    assert np.all(t1_from_radiance > np.exp(-s_ * diagonal) - 1e-6) and np.all(t1_from_radiance < 1.0)
    assert t1_from_radiance.mean() < 0.8  # the bounce segments lose a visible share to the fog


@pytest.mark.parametrize("q", [0.5, 0.25])
def test_bounce_nee_budget_under_the_sun(oracle, floor, q):
    """The bounce NEE budget draws its number from the sun's stream only at bounce vertices: on the sun
    floor (whose bounce rays leave into the black sky) the primary vertex's sun NEE is untouched, so
    the closed form L sin^2 r holds at q < 1 too."""
    data, ref = floor
    L, r_deg, spp = 1000.0, 2.0, 4
    k = sky(sun_radiance=(L, L, L), radius_deg=r_deg)
    _, radiance, _, counters = ref.trace_frame(data.camera, 4, 7, spp, sky=k, bounce_nee_probability=q)
    want = L * np.sin(np.radians(r_deg)) ** 2
    assert abs(radiance[:, 0].mean() / spp - want) < 2e-5 * want


def test_bounce_nee_budget_unbiased_with_sky_sun_and_glass(oracle):
    """On the glass terrain under the sky and the sun, two bounces: the image at q = 0.5 (bounce vertices
    run their sun / lamp NEE half the time, scaled by 2) has the mean of the image at q = 1, within 5 sigma
    of the per-pixel differences of two independent renders."""
    data = scene.glass_terrain()
    ref = build(oracle, data)
    k = sky(sun_dir=(0.3, 0.9, 0.2), sun_radiance=(40.0, 36.0, 30.0), radius_deg=3.0, zenith=(0.2, 0.3, 0.6),
            horizon=(0.6, 0.6, 0.7), ground=(0.1, 0.1, 0.1))
    spp = 16
    half = ref.trace_frame(data.camera, 2, 100, spp, sky=k, bounces_transparent=4, bounce_nee_probability=0.5)
    full = ref.trace_frame(data.camera, 2, 900, spp, sky=k, bounces_transparent=4)
    assert half[3][1] < full[3][1]  # fewer shadow rays
    d = (half[1][:, :3] - full[1][:, :3]).sum(axis=1) / spp
    assert abs(d.mean()) < 5 * d.std() / np.sqrt(len(d)), (d.mean(), d.std())
