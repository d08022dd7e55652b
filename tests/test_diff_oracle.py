"""Pins the oracle's float64 path-replay adjoint (oracle_trace_frame_adjoint) against central finite
differences (h = 1e-3) of the oracle's own forward render, with the Russian roulette off so the
estimator is a smooth function of the parameters at fixed random numbers."""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd import scene

NO_RR = 1
SPP = 2
BOUNCES = 3


@pytest.fixture(scope="module")
def setup(oracle):
    data = scene.small_scene()
    cam = dict(data.camera)
    cam["width"], cam["height"] = 40, 32
    o = oracle.OracleScene.build(data, int(np.ceil(len(data.keys) / 0.7)))
    bricks = len(data.keys)
    scale = np.ones(bricks, np.float32)
    scale[data.emissive_bricks] = np.random.default_rng(1).uniform(0.5, 1.5, len(data.emissive_bricks)).astype(np.float32)
    albedo = data.materials[0x8000:0x8100, 0:3].copy()
    o.set_parameters(albedo, scale)
    color, _, _, _ = o.trace_frame(cam, BOUNCES, 9, SPP, flags=NO_RR)
    target = 0.8 * color[:, :3] / SPP
    return data, cam, o, albedo, scale, target


def loss(o, cam, target):
    color, _, _, _ = o.trace_frame(cam, BOUNCES, 9, SPP, flags=NO_RR)
    image = color[:, :3] / SPP
    return float(((image - target) ** 2).sum()), image


def test_oracle_adjoint_matches_finite_differences(setup, oracle):
    data, cam, o, albedo, scale, target = setup
    o.set_parameters(albedo, scale)
    l0, image = loss(o, cam, target)
    d_color = np.zeros((image.shape[0], 4), np.float32)
    d_color[:, :3] = 2.0 * (image - target)
    ga, gs = o.trace_frame_adjoint(cam, BOUNCES, 9, SPP, d_color, flags=NO_RR)
    h = 1e-3
    # the largest albedo and emitter gradients
    ia = np.argsort(-np.abs(ga).ravel())[:5]
    for flat in ia:
        mid, ch = divmod(int(flat), 3)
        up = albedo.copy()
        dn = albedo.copy()
        up[mid, ch] += h
        dn[mid, ch] -= h
        o.set_parameters(up, scale)
        lu, _ = loss(o, cam, target)
        o.set_parameters(dn, scale)
        ld, _ = loss(o, cam, target)
        fd = (lu - ld) / (2 * h)
        assert abs(fd - ga[mid, ch]) <= 1e-3 * abs(fd) + 1e-6, (mid, ch, fd, ga[mid, ch])
    ie = np.argsort(-np.abs(gs))[:5]
    for b in ie:
        up = scale.copy()
        dn = scale.copy()
        up[b] += h
        dn[b] -= h
        o.set_parameters(albedo, up)
        lu, _ = loss(o, cam, target)
        o.set_parameters(albedo, dn)
        ld, _ = loss(o, cam, target)
        fd = (lu - ld) / (2 * h)
        assert abs(fd - gs[b]) <= 1e-3 * abs(fd) + 1e-6, (b, fd, gs[b])
    o.set_parameters(albedo, scale)
    # only emitters and materials that occur get gradients
    assert np.all(gs[np.setdiff1d(np.arange(len(gs)), data.emissive_bricks)] == 0.0)
    assert np.all(ga[0] == 0.0)


def test_oracle_scale_ones_is_the_plain_render(setup, oracle):
    data, cam, o, albedo, scale, target = setup
    o.set_parameters(albedo, None)
    plain = o.trace_frame(cam, BOUNCES, 9, SPP)[0]
    o.set_parameters(albedo, np.ones(len(data.keys), np.float32))
    ones = o.trace_frame(cam, BOUNCES, 9, SPP)[0]
    o.set_parameters(albedo, scale)
    assert np.array_equal(plain, ones)
