"""GPU checks of the differentiable form: the path-replay adjoint (cx_trace_frame_adjoint) against the
oracle's float64 gradient, against central finite differences (h = 1e-3) of the GPU forward render
with the roulette off, and through the torch.autograd.Function harness."""
from __future__ import annotations

import numpy as np
import pytest

from caelix_amd import scene

pytestmark = pytest.mark.gpu

NO_RR = 1
SPP = 2
BOUNCES = 3
SEED = 21


@pytest.fixture(scope="module")
def setup(cx, oracle):
    import torch

    data = scene.small_scene()
    cam = dict(data.camera)
    cam["width"], cam["height"] = 64, 48
    ts = cx.TraceScene(data, load_factor=0.7)
    o = oracle.OracleScene.from_gpu_arrays(ts.host_arrays(), data)
    rng = np.random.default_rng(4)
    scale = np.ones(len(data.keys), np.float32)
    scale[data.emissive_bricks] = rng.uniform(0.5, 1.5, len(data.emissive_bricks)).astype(np.float32)
    albedo = data.materials[0x8000:0x8100, 0:3].copy()
    scale_dev = torch.from_numpy(scale).cuda()
    ts.set_brick_scale(scale_dev)
    o.set_parameters(albedo, scale)
    return data, cam, ts, o, albedo, scale, scale_dev


def gpu_image(ts, cam, flags, seed=SEED):
    color, _, _, _ = ts.trace_frame(cam, BOUNCES, seed, SPP, flags=flags)
    return color.view(-1, 4)[:, :3].double().cpu().numpy() / SPP


def gpu_grads(ts, cam, d_color, flags):
    import torch

    ga = torch.zeros(256, 4, dtype=torch.float32, device="cuda")
    gs = torch.zeros(ts.bricks, dtype=torch.float32, device="cuda")
    d = torch.from_numpy(np.ascontiguousarray(d_color, np.float32).reshape(-1)).cuda()
    ts.trace_frame_adjoint(cam, BOUNCES, SEED, SPP, d, ga, gs, flags=flags)
    return ga[:, :3].double().cpu().numpy(), gs.double().cpu().numpy()


def test_scale_ones_is_the_plain_render(setup):
    import torch

    data, cam, ts, o, albedo, scale, scale_dev = setup
    ts.set_brick_scale(None)
    plain = ts.trace_frame(cam, BOUNCES, SEED, SPP)[0].cpu().numpy()
    ones = torch.ones(len(data.keys), dtype=torch.float32, device="cuda")
    ts.set_brick_scale(ones)
    with_ones = ts.trace_frame(cam, BOUNCES, SEED, SPP)[0].cpu().numpy()
    ts.set_brick_scale(scale_dev)
    assert np.array_equal(plain, with_ones)


@pytest.mark.parametrize("flags", [0, NO_RR])
def test_scaled_forward_matches_oracle(setup, flags):
    data, cam, ts, o, albedo, scale, scale_dev = setup
    color, radiance, attrib, counters = ts.trace_frame(cam, BOUNCES, SEED, SPP, flags=flags)
    w_color, w_rad, w_attrib, w_counters = o.trace_frame(cam, BOUNCES, SEED, SPP, flags=flags)
    assert np.array_equal(counters.cpu().numpy().view(np.uint64)[:5], w_counters[:5])
    got = color.view(-1, 4)[:, :3].double().cpu().numpy()
    assert np.all(np.abs(got - w_color[:, :3]) <= 2e-5 * (1.0 + np.abs(w_color[:, :3])))


@pytest.mark.parametrize("flags", [0, NO_RR])
def test_adjoint_matches_oracle_float64_gradient(setup, flags):
    data, cam, ts, o, albedo, scale, scale_dev = setup
    image = gpu_image(ts, cam, flags)
    target = 0.8 * image
    d_color = np.zeros((image.shape[0], 4), np.float32)
    d_color[:, :3] = 2.0 * (image - target)
    ga, gs = gpu_grads(ts, cam, d_color, flags)
    wa, ws = o.trace_frame_adjoint(cam, BOUNCES, SEED, SPP, d_color, flags=flags)
    assert np.abs(wa).max() > 0 and np.abs(ws).max() > 0
    assert np.abs(ga - wa).max() <= 1e-4 * np.abs(wa).max(), np.abs(ga - wa).max() / np.abs(wa).max()
    assert np.abs(gs - ws).max() <= 1e-4 * np.abs(ws).max(), np.abs(gs - ws).max() / np.abs(ws).max()
    big = np.abs(wa) > 1e-2 * np.abs(wa).max()
    np.testing.assert_allclose(ga[big], wa[big], rtol=1e-3)


def test_adjoint_matches_gpu_finite_differences(setup):
    import torch

    data, cam, ts, o, albedo, scale, scale_dev = setup
    mats = ts.materials.view(65536, 8)
    image = gpu_image(ts, cam, NO_RR)
    target = 0.8 * image

    def loss():
        return float(((gpu_image(ts, cam, NO_RR) - target) ** 2).sum())

    d_color = np.zeros((image.shape[0], 4), np.float32)
    d_color[:, :3] = 2.0 * (image - target)
    ga, gs = gpu_grads(ts, cam, d_color, NO_RR)
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
        assert abs(fd - ga[mid, ch]) <= 1e-3 * abs(fd) + 1e-4, (mid, ch, fd, ga[mid, ch])
    for b in np.argsort(-np.abs(gs))[:4]:
        base = float(scale_dev[b])
        scale_dev[b] = base + h
        lu = loss()
        scale_dev[b] = base - h
        ld = loss()
        scale_dev[b] = base
        fd = (lu - ld) / (2 * h)
        assert abs(fd - gs[b]) <= 1e-3 * abs(fd) + 1e-4, (b, fd, gs[b])
    torch.cuda.synchronize()


def test_autograd_function(setup, cx):
    import torch

    from caelix_amd.diff import DifferentiableRenderer

    data, cam, ts, o, albedo, scale, scale_dev = setup
    r = DifferentiableRenderer(ts, cam, BOUNCES, SPP, SEED, data.emissive_bricks, flags=NO_RR)
    a = torch.from_numpy(albedo).cuda().requires_grad_(True)
    s = torch.from_numpy(scale[data.emissive_bricks]).cuda().requires_grad_(True)
    with torch.no_grad():
        target = 0.8 * r.render(a, s)
    img = r.render(a, s)
    loss = ((img.double() - target.double()) ** 2).sum()
    loss.backward()
    d_color = np.zeros((cam["width"] * cam["height"], 4), np.float32)
    d_color[:, :3] = (2.0 * (img.detach().double() - target.double())).cpu().numpy().reshape(-1, 3)
    ga, gs = gpu_grads(ts, cam, d_color, NO_RR)
    np.testing.assert_allclose(a.grad.double().cpu().numpy(), ga, rtol=1e-5, atol=1e-6 * np.abs(ga).max())
    np.testing.assert_allclose(s.grad.double().cpu().numpy(), gs[data.emissive_bricks], rtol=1e-5,
                               atol=1e-6 * np.abs(gs).max())
    ts.set_brick_scale(scale_dev)
