"""The host's sky cubemap in place of the procedural sky (caelix_shading.hlsl:346-447,
shaders/render/caelix_sky_sh.hlsl), as the reference's trace test checks it ("sky cubemap, fog and
emission scale", tests/trace_test.cpp:2721-2800), on the red brick (tests/ref_fixtures.py), on the oracle
(CPU) and on the HIP path through the C ABI (gpu):

uniform cube     a cube of one radiance L (8 x 8 faces): a primary miss is L (1e-4) with depth 0, the
                 brick is black at 0 bounces; without the cube the procedural sky is back
SH irradiance    at 1 bounce every diffuse bounce ray from the brick's front face misses and receives
                 E(n) / pi = L: the red brick is albedo * L within 2e-3, with no variance (two seeds
                 give the same value within 1e-6)
linear sky       L(w) = a + b w.z is exactly bands 0 and 1: E(n) / pi = a + 2 b n.z / 3, so the front
                 face (n = -Z) sees a - 2 b / 3 (1e-2: the 8 x 8 cube interpolates the gradient)
sampling         a cube whose texels hold (face, x, y) returns each texel at its centre direction (within
                 1e-4: the direction's rounding leaves bilinear weights of a few 1e-8; a wrong face or
                 orientation is off by 1 or more; the reference test's CubeTexelDirection layout), a lod
                 between two constant mips their linear blend, lods clamp to the mips
SH closed forms  a constant cube projects to E(n) = pi L at every normal (1e-6; bands 1 and 2 vanish up
                 to the quadrature, 2e-3 L); the linear sky to a + 2 b n.z / 3 at six normals (1e-2)
GPU = oracle     the probes bit for bit; frames with a three-mip cube (diffuse misses through the SH,
                 specular misses at the roughness mip on a glass terrain) as the other frame comparisons
"""
from __future__ import annotations

import numpy as np
import pytest

import ref_fixtures as rf
from caelix_amd import cube_texel_direction, make_sky_cube, sky_params
from test_ref_scenes_oracle import build

L = np.array([0.5, 1.0, 2.0])
A, B = 1.0, 0.6


def red_brick():
    cam = rf.look_plus_z(4.0, 4.0, -20.0)
    return cam, rf.scene_from_world(rf.red_brick(), cam, lights=False)


def sky():
    return sky_params(sun_dir=(0.0, 1.0, 0.0), sun_radiance=(0.0, 0.0, 0.0), zenith=(0.2, 0.4, 0.9),
                      horizon=(0.7, 0.7, 0.8), ground=(0.3, 0.25, 0.2))


def constant_cube(size=8, mips=1):
    return make_sky_cube(size, lambda d: np.broadcast_to(L, d.shape), mips), size, mips


def linear_cube(size=8):
    return make_sky_cube(size, lambda d: np.repeat(A + B * d[..., 2:3], 3, axis=-1), 1), size, 1


def check_uniform(frame, cam):
    """frame(cube or None, bounces, seed) -> color [n, 4] (float64)."""
    w = cam["width"]
    centre = (cam["height"] // 2) * w + w // 2
    color = frame(constant_cube(), 0, 1)
    assert np.all(np.abs(color[0, :3] - L) < 1e-4) and color[0, 3] == 0.0
    assert np.all(color[centre, :3] == 0.0) and color[centre, 3] > 0.0
    plain = frame(None, 0, 1)
    assert np.all(np.abs(plain[0, :3] - L) > 1e-2)  # the procedural sky, not the cube


def check_sh(frame, cam):
    w = cam["width"]
    cx, cy = w // 2, cam["height"] // 2
    probes = [(cx, cy), (cx - 7, cy), (cx + 7, cy), (cx, cy - 7), (cx, cy + 7)]
    color = frame(constant_cube(), 1, 2)
    for x, y in probes:
        c = color[y * w + x]
        assert abs(c[0] - L[0]) < 2e-3 and c[1] == 0.0 and c[2] == 0.0, c
    again = frame(constant_cube(), 1, 101)
    assert abs(again[cy * w + cx, 0] - color[cy * w + cx, 0]) < 1e-6
    linear = frame(linear_cube(), 1, 5)
    assert abs(linear[cy * w + cx, 0] - (A - 2.0 * B / 3.0)) < 1e-2


def index_cube(size):
    """Texel (face, x, y) of mip 0 holds (face, x, y) (and mip 1, when present, the constant 100)."""
    faces = []
    for f in range(6):
        y, x = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
        faces.append(np.stack([np.full_like(x, f), x, y, np.ones_like(x)], -1).reshape(-1, 4))
    mip0 = np.concatenate(faces).astype(np.float32)
    e = size // 2
    mip1 = np.tile(np.array([100.0, 100.0, 100.0, 1.0], np.float32), (6 * e * e, 1))
    return np.concatenate([mip0, mip1])


def centre_queries(size, lod=0.0):
    rows, want = [], []
    c = (np.arange(size) + 0.5) / size * 2.0 - 1.0
    for f in range(6):
        for y in range(size):
            for x in range(size):
                d = cube_texel_direction(f, np.float64(c[x]), np.float64(c[y]))
                d = d / np.linalg.norm(d)
                rows.append([d[0], d[1], d[2], lod])
                want.append([f, x, y])
    return np.array(rows, np.float32), np.array(want, np.float32)


def check_sampling(sample):
    """sample(texels, size, mips, queries) -> [n, 3]."""
    size = 8
    texels = index_cube(size)
    q, want = centre_queries(size)
    got = sample(texels, size, 2, q)
    assert np.allclose(got, want, rtol=0, atol=1e-4)
    q[:, 3] = 0.25  # a quarter of the way to the constant mip 1
    blended = sample(texels, size, 2, q)
    assert np.allclose(blended, want + 0.25 * (100.0 - want), rtol=0, atol=1e-4)
    q[:, 3] = 7.0  # clamped to the last mip
    assert np.all(sample(texels, size, 2, q) == 100.0)
    q[:, 3] = -3.0
    assert np.array_equal(sample(texels, size, 2, q), got)


def sh_irradiance(sh, n):
    n = np.asarray(n, np.float64)
    basis = np.array([0.282095, 0.488603 * n[1], 0.488603 * n[2], 0.488603 * n[0], 1.092548 * n[0] * n[1],
                      1.092548 * n[1] * n[2], 0.315392 * (3 * n[2] ** 2 - 1), 1.092548 * n[0] * n[2],
                      0.546274 * (n[0] ** 2 - n[1] ** 2)])
    return basis @ sh.astype(np.float64)


def check_sh_closed_forms(project):
    """project(texels, size, mips) -> [9, 3]."""
    normals = [(0, 0, -1), (0, 0, 1), (1, 0, 0), (0, -1, 0), (0.6, 0.0, 0.8), (-0.48, 0.6, 0.64)]
    texels, size, mips = constant_cube(32, 3)
    sh = project(texels, size, mips)
    for n in normals:
        assert np.all(np.abs(sh_irradiance(sh, n) / np.pi - L) <= 1e-6 * L + 2e-3 * L)
    assert np.all(np.abs(sh[1:]) < 2e-3 * np.pi * L.max())
    e0 = sh[0, 0] * 0.282095 / np.pi
    assert abs(e0 - L[0]) < 1e-6 * L[0] + 1e-7
    texels, size, mips = linear_cube(64)
    sh = project(texels, size, mips)
    for n in normals:
        assert abs(sh_irradiance(sh, n)[0] / np.pi - (A + 2.0 * B * n[2] / 3.0)) < 1e-2


# ------------------------------------------------------------------------------------------- oracle


def test_cube_sampling_oracle(oracle):
    check_sampling(oracle.sky_cube_sample)


def test_sh_closed_forms_oracle(oracle):
    check_sh_closed_forms(oracle.sky_sh)


@pytest.fixture(scope="module")
def brick_oracle(oracle):
    cam, data = red_brick()
    return cam, build(oracle, data)


def oracle_frame(o, cam):
    def frame(cube, bounces, seed):
        return o.trace_frame(cam, bounces, seed, 1, sky=sky(), sky_cube=cube)[0]
    return frame


def test_uniform_cube_oracle(brick_oracle):
    cam, o = brick_oracle
    check_uniform(oracle_frame(o, cam), cam)


def test_sh_irradiance_oracle(brick_oracle):
    cam, o = brick_oracle
    check_sh(oracle_frame(o, cam), cam)


# ------------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_cube_probes_gpu(cx, oracle):
    check_sampling(cx.sky_cube_sample)
    check_sh_closed_forms(cx.sky_sh)
    rng = np.random.default_rng(4)
    texels = make_sky_cube(16, lambda d: np.stack([np.sin(3 * d[..., 0]) + 1.5, d[..., 1] ** 2 + 0.2,
                                                   np.exp(d[..., 2])], -1), 5)
    q = np.concatenate([rng.normal(size=(20000, 3)), rng.uniform(-1, 5, size=(20000, 1))], axis=1).astype(np.float32)
    assert np.array_equal(cx.sky_cube_sample(texels, 16, 5, q).view(np.uint32),
                          oracle.sky_cube_sample(texels, 16, 5, q).view(np.uint32))
    assert np.array_equal(cx.sky_sh(texels, 16, 5).view(np.uint32), oracle.sky_sh(texels, 16, 5).view(np.uint32))
    big = make_sky_cube(128, lambda d: np.abs(d) + 0.1, 8)
    assert np.array_equal(cx.sky_sh(big, 128, 8).view(np.uint32), oracle.sky_sh(big, 128, 8).view(np.uint32))


@pytest.mark.gpu
def test_cube_frames_gpu(cx, oracle):
    from test_glass_gpu import compare

    cam, data = red_brick()
    gpu = cx.TraceScene(data, load_factor=0.7)

    def frame(cube, bounces, seed):
        return gpu.trace_frame(cam, bounces, seed, 1, sky=sky(), sky_cube=cube)[0].cpu().numpy().reshape(-1, 4).astype(
            np.float64)

    check_uniform(frame, cam)
    check_sh(frame, cam)
    o = build(oracle, data)
    cube = (make_sky_cube(16, lambda d: np.stack([d[..., 0] + 1.2, d[..., 1] + 1.1, np.abs(d[..., 2])], -1), 3), 16, 3)
    compare(gpu.trace_frame(cam, 2, 9, 2, sky=sky(), sky_cube=cube),
            o.trace_frame(cam, 2, 9, 2, sky=sky(), sky_cube=cube))


@pytest.mark.gpu
def test_cube_glass_terrain_gpu(cx, oracle):
    """Specular bounces that miss read the cube at their roughness mip (a glass terrain: transparent
    interfaces and opaque specular lobes), with the sun: GPU = oracle."""
    from caelix_amd import scene
    from test_glass_gpu import compare

    data = scene.glass_terrain()
    gpu = cx.TraceScene(data, load_factor=0.7)
    o = build(oracle, data)
    s = sky_params(sun_dir=(0.3, 1.0, 0.2), sun_radiance=(20.0, 20.0, 20.0), sun_radius_deg=2.0)
    cube = (make_sky_cube(32, lambda d: np.stack([d[..., 1] + 1.5, np.abs(d[..., 0]) + 0.3, d[..., 2] ** 2], -1), 6),
            32, 6)
    for seed in (3, 4):
        compare(gpu.trace_frame(data.camera, 3, seed, 1, sky=s, sky_cube=cube, bounces_transparent=4),
                o.trace_frame(data.camera, 3, seed, 1, sky=s, sky_cube=cube, bounces_transparent=4))


@pytest.mark.gpu
def test_cube_rejects_bad_arguments(cx):
    cam, data = red_brick()
    gpu = cx.TraceScene(data, load_factor=0.7)
    texels, size, _ = constant_cube()
    for mips, flags in ((0, cx.TRACE_SKY), (5, cx.TRACE_SKY), (1, 0)):
        with pytest.raises(cx.CxError):
            gpu.trace_frame(cam, 1, 1, 1, sky=sky() if flags else None, sky_cube=(texels, size, mips))
