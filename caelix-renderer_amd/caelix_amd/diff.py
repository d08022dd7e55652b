"""The differentiable form: radiance gradients with respect to the material albedos and the emitters'
emission scales, behind a torch.autograd.Function.

forward  cx_trace_frame renders the image I = color / spp (rgb) with the current parameters;
backward cx_trace_frame_adjoint replays the same paths (the same random streams, rays re-traced and
         lights re-drawn) and adds dL/d(albedo) and dL/d(scale) for dL/dI = grad_output.

Parameters: `albedo` [256, 3], the albedo of the opaque materials 0x8000 | id (id 0 unused), and
`scales` [E], the emission scale of emitter e = brick record emitters[e] (every other brick keeps
scale 1). The light tree samples with the unscaled emission (a fixed proposal); sampling pdfs, MIS
weights and the Russian roulette normalisation are held constant in the gradient. With
flags=TRACE_NO_RUSSIAN_ROULETTE the image is a smooth function of the parameters at fixed random
numbers, which central finite differences check exactly.
"""
from __future__ import annotations

import numpy as np

from . import TraceScene, TRACE_NO_RUSSIAN_ROULETTE, _torch

__all__ = ["DifferentiableRenderer", "TRACE_NO_RUSSIAN_ROULETTE"]


class DifferentiableRenderer:
    def __init__(self, scene: TraceScene, camera: dict, bounces: int, spp: int, seed: int, emitters, flags: int = 0):
        torch = _torch()
        self.ts = scene
        self.camera = camera
        self.bounces = int(bounces)
        self.spp = int(spp)
        self.seed = int(seed)
        self.flags = int(flags)
        self.width, self.height = int(camera["width"]), int(camera["height"])
        self.emitters = torch.as_tensor(np.asarray(emitters, np.int64), device="cuda")
        self.brick_scale = torch.ones(scene.bricks, dtype=torch.float32, device="cuda")
        scene.set_brick_scale(self.brick_scale)
        n = self.width * self.height
        self.color = torch.empty(n * 4, dtype=torch.float32, device="cuda")
        self.d_color = torch.zeros(n * 4, dtype=torch.float32, device="cuda")
        self.counters = torch.zeros(8, dtype=torch.int64, device="cuda")

    def initial_albedo(self):
        """The scene's current albedos as a [256, 3] tensor."""
        m = self.ts.materials.view(65536, 8)
        return m[0x8000:0x8100, 0:3].clone()

    def set_parameters(self, albedo, scales) -> None:
        m = self.ts.materials.view(65536, 8)
        m[0x8000:0x8100, 0:3] = albedo.detach().to(torch_dtype()).contiguous()
        self.brick_scale.fill_(1.0)
        self.brick_scale[self.emitters] = scales.detach().to(torch_dtype())

    def forward_image(self):
        self.ts.trace_frame(self.camera, self.bounces, self.seed, self.spp, color=self.color, radiance=None,
                            attrib=None, counters=self.counters, flags=self.flags)
        return self.color.view(self.height, self.width, 4)[..., :3] / float(self.spp)

    def adjoint(self, grad_image):
        torch = _torch()
        self.d_color.view(self.height, self.width, 4)[..., :3] = grad_image.to(torch.float32)
        grad_albedo = torch.zeros(256, 4, dtype=torch.float32, device="cuda")
        grad_scale = torch.zeros(self.ts.bricks, dtype=torch.float32, device="cuda")
        self.ts.trace_frame_adjoint(self.camera, self.bounces, self.seed, self.spp, self.d_color, grad_albedo, grad_scale,
                                    flags=self.flags, counters=self.counters)
        return grad_albedo[:, :3], grad_scale[self.emitters]

    def render(self, albedo, scales):
        return _Render.apply(albedo, scales, self)


def torch_dtype():
    return _torch().float32


def _make_function():
    torch = _torch()

    class Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, albedo, scales, renderer):
            renderer.set_parameters(albedo, scales)
            image = renderer.forward_image().clone()
            ctx.renderer = renderer
            ctx.save_for_backward(albedo, scales)
            return image

        @staticmethod
        def backward(ctx, grad_image):
            renderer = ctx.renderer
            albedo, scales = ctx.saved_tensors
            renderer.set_parameters(albedo, scales)
            grad_albedo, grad_scales = renderer.adjoint(grad_image.contiguous())
            return grad_albedo.to(albedo.dtype), grad_scales.to(scales.dtype), None

    return Render


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _make_function()
        return cls.fn.apply(*args)


_Render = _Lazy
