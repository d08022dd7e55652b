"""The oracle on the reference tests' own scenes (tests/ref_fixtures.py rebuilds them from their seeds and
literals), against what those tests check and what the geometry fixes:

random opaque world   tests/trace_test.cpp:961-1031 (mt19937(1234), 600 bricks): primary rays hit, and never
                      an air voxel; every hit is the first opaque voxel of a float64 voxel walk
red brick             trace_test.cpp:449-536: the brick covers exactly the pixels whose ray crosses its -Z
                      face square, seen as red through that face at depth 20 / cos
emissive ceiling      trace_test.cpp:1199-1288: one bounce from a grey floor under an emissive ceiling; a
                      1 spp sample is exactly 0 or the albedo; the mean is the albedo times the point-to-
                      rectangle form factor (the textbook closed form, checked here by Monte Carlo), within
                      5 sigma and within the reference's 3 %
lamp box              trace_test.cpp:3100-3219: the light NEE (MIS, no MIS, bounce NEE 0.25) against the
                      plain path tracer (lamps found by BSDF hits only), lamps alone and with the sun
                      through the glazed opening: mean ratio within 0.97..1.03 and 16 x 16 block PSNR
                      above 30 dB over the pixels that do not see a lamp directly
"""
from __future__ import annotations

import numpy as np
import pytest

import ref_fixtures as rf
import trace_checks as tc
from caelix_amd import sky_params

LOAD = 0.7


def build(oracle, data, mis: bool = True):
    return oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / LOAD)), mis=mis)


def primary_rays(cam) -> np.ndarray:
    d = tc.primary_dirs(cam)
    o = np.broadcast_to(np.asarray(cam["position"], np.float32), d.shape)
    return np.concatenate([o, d, np.full((len(d), 1), 1e30, np.float32)], axis=1).astype(np.float32)


# ------------------------------------------------------------------------------------------ worlds


def random_world():
    world = rf.random_opaque_world(1234, 600)
    for brick in world.values():
        brick[brick == 0xFC1F] = 0xFC1E  # the reference's air marker (pure magenta) is kept out
    return world


def test_random_world_generator_matches_the_reference_tests_asserts():
    """RandomBlobs + RandomBrick over mt19937(20261014), as the finalize test draws them: 300 keys in at
    least 15 groups of shift 3, some negative (finalize_test.cpp:325-342); blobs grow by face steps."""
    rng = rf.Mt19937Stream(20261014)
    keys = rf.random_blobs(rng, 300)
    assert len(keys) == len(set(keys)) == 300
    assert len({(x >> 3, y >> 3, z >> 3) for x, y, z in keys}) >= 15
    assert any(min(k) < 0 for k in keys)
    brick = rf.random_brick(rng)
    air = (brick == 0).mean()
    assert 0.2 < air < 0.4 and np.all(brick[brick != 0] & 0x8000)
    world = random_world()
    assert len(world) == 600


def test_random_world_hits_no_air_and_first_opaque_voxel(oracle):
    cam = rf.look_plus_z(0.0, 0.0, -900.0)
    data = rf.scene_from_world(random_world(), cam, lights=False)
    o = build(oracle, data)
    _, _, attrib, _ = o.trace_frame(cam, 0, 1, 1)
    hits = attrib != 0
    assert hits.sum() > 500
    block = attrib >> 16
    assert np.all(block[hits] & 0xC000) and not np.any(block[hits] == 0xFC1F)
    rays = primary_rays(cam)
    h = o.trace_rays(rays)
    assert np.array_equal(h[:, 1], attrib)  # the frame's primary hit is TraceScene's
    cell = (h[:, 1] >> 6) & 0x1FF
    payload = o.a["payload"].reshape(-1, 512)
    assert np.array_equal(payload[h[hits, 2], cell[hits]], block[hits])
    voxel, t, axis, hit, tie = tc.walk_sparse(data.keys, data.payload, 0xC000, rays[:, 0:3], rays[:, 3:6])
    got = np.stack([h[:, 3].view(np.int32) * 8 + (cell & 7), h[:, 4].view(np.int32) * 8 + ((cell >> 3) & 7),
                    h[:, 5].view(np.int32) * 8 + (cell >> 6)], axis=1)
    d = rays[:, 3:6]
    flags = h[:, 1] & 0x3F
    expect = np.where(hit, 1 << (2 * axis + (d[np.arange(len(d)), np.maximum(axis, 0)] > 0)), 0)
    t_got = h[:, 0].view(np.float32).astype(np.float64)
    same = (hit == hits) & np.where(hits, np.all(got == voxel, axis=1) & (flags == expect)
                                    & (np.abs(t_got - t) < 1e-3 * np.maximum(1.0, t)), True)
    # Every ray whose walk never crosses two planes within 5e-4 of each other agrees exactly; the others
    # may differ where the float32 DDA's 1e-4 push cuts a corner.
    assert np.all(same[tie >= 5e-4])
    assert (~same).mean() < 0.01, (~same).sum()
    assert (hit & hits).sum() > 500


def red_brick_expectation(cam):
    """Per pixel (float64 from the float32 directions): whether the ray crosses the brick's -Z face
    square [0, 8]^2 at z = 0, the hit distance, and the distance to the square's edge."""
    d = tc.primary_dirs(cam).astype(np.float64)
    o = np.asarray(cam["position"], np.float64)
    t = (0.0 - o[2]) / d[:, 2]
    x = o[0] + d[:, 0] * t
    y = o[1] + d[:, 1] * t
    inside = (x >= 0) & (x <= 8) & (y >= 0) & (y <= 8)
    edge = np.minimum.reduce([np.abs(x), np.abs(x - 8), np.abs(y), np.abs(y - 8)])
    return inside, t, edge


def check_red_brick(color, attrib, data, cam):
    inside, t, edge = red_brick_expectation(cam)
    clear = edge > 1e-3
    hits = attrib != 0
    assert np.array_equal(hits[clear], inside[clear])
    assert hits.sum() > 5000
    centre = (cam["height"] // 2) * cam["width"] + cam["width"] // 2
    assert hits[centre] and (attrib[centre] >> 16) == rf.RED
    assert hits[0] == 0 and hits[-1] == 0  # corners see the sky
    assert np.all(attrib[hits] >> 16 == rf.RED)
    assert np.all(attrib[hits] & 0x3F == 1 << 5)  # entered through the -Z face: normal (0, 0, -1)
    depth = color.reshape(-1, 4)[:, 3]
    assert np.all(np.abs(depth[hits] - t[hits]) <= 1e-5 * t[hits])
    assert list(data.materials[rf.RED, 0:3]) == [1.0, 0.0, 0.0]  # the albedo view's red


def test_red_brick(oracle):
    cam = rf.look_plus_z(4.0, 4.0, -20.0)
    data = rf.scene_from_world(rf.red_brick(), cam, lights=False)
    o = build(oracle, data)
    color, _, attrib, _ = o.trace_frame(cam, 0, 1, 1)
    check_red_brick(color, attrib, data, cam)


# ------------------------------------------------------------------------------- emissive ceiling


def rectangle_form_factor(px, pz, h, x0, x1, z0, z1):
    """The form factor from an upward differential area at (px, 0, pz) to the downward rectangle
    [x0, x1] x [z0, z1] at height h (plane element to a parallel rectangle with its normal through a
    corner, summed over the four corner rectangles)."""
    def corner(a, b):
        ah, bh = a / h, b / h
        sa, sb = np.sqrt(1 + ah * ah), np.sqrt(1 + bh * bh)
        return (ah / sa * np.arctan(bh / sa) + bh / sb * np.arctan(ah / sb)) / (2 * np.pi)
    return corner(x1 - px, z1 - pz) - corner(x0 - px, z1 - pz) - corner(x1 - px, z0 - pz) + corner(x0 - px, z0 - pz)


def test_form_factor_against_monte_carlo():
    rng = np.random.default_rng(42)
    n = 2_000_000
    for px, pz in ((32.0, 32.0), (5.0, 60.0), (-3.0, 10.0)):
        r1, phi = rng.uniform(size=n), 2 * np.pi * rng.uniform(size=n)
        s = np.sqrt(r1)
        dx, dz, dy = s * np.cos(phi), s * np.sin(phi), np.sqrt(1 - r1)
        t = 24.0 / dy
        x, z = px + dx * t, pz + dz * t
        p = ((x >= 0) & (x < 64) & (z >= 0) & (z < 64)).mean()
        f = rectangle_form_factor(px, pz, 24.0, 0.0, 64.0, 0.0, 64.0)
        assert abs(p - f) < 5 * np.sqrt(f * (1 - f) / n), (px, pz, p, f)


def ceiling_expectation(cam, push=0.002):
    """Per pixel: the probability that the cosine-sampled bounce from the pixel's floor point reaches
    the ceiling (its bottom face at y = 24, the bounce leaving from 0.002 above the floor)."""
    d = tc.primary_dirs(cam).astype(np.float64)
    o = np.asarray(cam["position"], np.float64)
    t = (0.0 - o[1]) / d[:, 1]
    return rectangle_form_factor(o[0] + d[:, 0] * t, o[2] + d[:, 2] * t, rf.CEILING_Y - push, 0.0,
                                 rf.ROOM_SIZE, 0.0, rf.ROOM_SIZE)


def centre(values, size, n=256):
    img = np.asarray(values).reshape(n, n)
    a = (n - size) // 2
    return img[a:a + size, a:a + size]


def check_ceiling(radiance_1spp, color_1spp, radiance, spp, cam, albedo):
    r1 = radiance_1spp.reshape(-1, 4)[:, 0]
    assert np.all((r1 == 0.0) | (r1 == 1.0))  # one bounce, emission 1: a sample is 0 or 1
    assert np.array_equal(color_1spp.reshape(-1, 4)[:, 0], (np.float32(albedo) * r1.astype(np.float32)))
    p = ceiling_expectation(cam)
    mean = radiance.reshape(-1, 4)[:, 0] / spp
    for size in (128, 32):
        got = centre(mean, size).mean()
        want = centre(p, size).mean()
        sigma = np.sqrt((centre(p, size) * (1 - centre(p, size))).sum() / spp) / size**2
        assert abs(got - want) < 5 * sigma + 1e-9, (size, got, want, sigma)
        assert abs(got * albedo / (want * albedo) - 1) < 0.03


def test_emissive_ceiling_one_bounce(oracle):
    cam = rf.look_down()
    data = rf.scene_from_world(rf.ceiling_room(), cam)
    assert len(data.emissive_bricks) == 0  # emission 1 is below the default light threshold of 20
    o = build(oracle, data)
    albedo = float(data.materials[rf.GREY_FLOOR, 0])
    assert albedo == float(np.float16(15 / 31))
    c1, r1, _, _ = o.trace_frame(cam, 1, 77, 1)
    _, r, _, _ = o.trace_frame(cam, 1, 1000, 64)
    check_ceiling(r1, c1, r, 64, cam, albedo)


# ------------------------------------------------------------------------------------- lamp box


SUN_DIR = np.array([0.0, 1.0, -0.4]) / np.linalg.norm([0.0, 1.0, -0.4])


def lamp_box_sky(sun: float):
    if sun == 0.0:
        return None
    return sky_params(sun_dir=tuple(SUN_DIR), sun_radiance=(sun, sun, sun), sun_radius_deg=2.0)


def compare_light(ref: np.ndarray, test: np.ndarray, size: int = rf.LAMP_BOX_IMAGE, block: int = 16,
                  direct: float = 32.0):
    """CompareLight (trace_test.cpp:3019-3087): (reference mean, ratio of means, 16 x 16 block PSNR) over
    the pixels whose reference colour is at most `direct` in every channel."""
    r = ref.reshape(size, size, 4)[:, :, :3].astype(np.float64)
    t = test.reshape(size, size, 4)[:, :, :3].astype(np.float64)
    keep = ~(r > direct).any(axis=2)
    squared, blocks = 0.0, 0
    for by in range(0, size, block):
        for bx in range(0, size, block):
            k = keep[by:by + block, bx:bx + block]
            used = int(k.sum())
            if used == 0:
                continue
            rb = (r[by:by + block, bx:bx + block] * k[:, :, None]).sum(axis=(0, 1)) / used
            tb = (t[by:by + block, bx:bx + block] * k[:, :, None]).sum(axis=(0, 1)) / used
            squared += ((np.clip(tb, 0, 1) - np.clip(rb, 0, 1)) ** 2).sum()
            blocks += 1
    values = 3.0 * keep.sum()
    ref_mean = (r * keep[:, :, None]).sum() / values
    test_mean = (t * keep[:, :, None]).sum() / values
    mse = squared / (3.0 * blocks)
    return ref_mean, test_mean / ref_mean, 10 * np.log10(1 / mse) if mse > 0 else 99.0


LAMP_VARIANTS = [("mis", True, 1.0), ("no mis", False, 1.0), ("bounce nee 0.25", True, 0.25)]


@pytest.mark.parametrize("sun", [0.0, 1000.0])
def test_lamp_box_light_nee_matches_path_tracer(oracle, sun):
    cam = rf.lamp_box_camera()
    plain = rf.scene_from_world(rf.lamp_box(), cam, glass=True, lights=False)
    lit = rf.scene_from_world(rf.lamp_box(), cam, glass=True)
    assert len(lit.emissive_bricks) == 5 and len(plain.emissive_bricks) == 0
    sky = lamp_box_sky(sun)
    spp_pt, spp_nee = 1024, 256
    pt = build(oracle, plain).trace_frame(cam, 4, 100000, spp_pt, bounces_transparent=4, sky=sky)[0] / spp_pt
    for name, mis, q in LAMP_VARIANTS:
        o = build(oracle, lit, mis=mis)
        assert int(o.lt.control[0]["light_bricks"]) == 5
        nee = o.trace_frame(cam, 4, 200000, spp_nee, bounces_transparent=4, sky=sky,
                            bounce_nee_probability=q)[0] / spp_nee
        ref_mean, ratio, psnr = compare_light(pt, nee)
        assert ref_mean > 0.01, name
        assert 0.97 < ratio < 1.03, (name, ratio)
        assert psnr > 30.0, (name, psnr)
