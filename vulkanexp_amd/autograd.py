"""torch.autograd.Function over the differentiable probe update of libddgi.

    probes.diff_prepare(R, gbuffer, n_materials)             # once per frame: trace, factors, binning
    loss = shaded_gbuffer_loss(probes, albedo, sun, target)    # albedo (M, 3), sun (3,), target (N, 3)
    loss.backward()                                            # albedo.grad, sun.grad

Forward and backward are HIP kernels (ddgi_diff_forward / ddgi_diff_backward); torch only
provides the device buffers and the stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, load


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class DiffContext:
    """The prepared differentiable state of one IrradianceProbes handle."""

    def __init__(self, probes, orientation, position, normal, to_camera, material, n_materials):
        self.probes = probes
        self.lib = load()
        self.N = int(position.shape[0])
        self.M = int(n_materials)
        self.generation = 0
        keep = [t.contiguous() for t in (position, normal, to_camera)]
        mat = material.to(torch.int32).contiguous()
        R = np.ascontiguousarray(orientation, np.float32).reshape(9)
        check(self.lib.ddgi_diff_prepare(probes.handle, R.ctypes.data_as(C.c_void_p), _ptr(keep[0]), _ptr(keep[1]),
                                         _ptr(keep[2]), _ptr(mat), self.N, self.M, _stream()), probes.handle)

    def forward(self, albedo, sun, target):
        shade = torch.empty_like(target)
        loss = torch.empty(1, dtype=torch.float32, device=target.device)
        check(self.lib.ddgi_diff_forward(self.probes.handle, _ptr(albedo), _ptr(sun), _ptr(target), _ptr(shade),
                                         _ptr(loss), _stream()), self.probes.handle)
        self.generation += 1
        return loss, shade

    def backward(self, albedo, sun, target, grad_loss):
        g_albedo = torch.empty_like(albedo)
        g_sun = torch.empty_like(sun)
        check(self.lib.ddgi_diff_backward(self.probes.handle, _ptr(albedo), _ptr(sun), _ptr(target), _ptr(grad_loss),
                                          _ptr(g_albedo), _ptr(g_sun), _stream()), self.probes.handle)
        return g_albedo, g_sun


class _ShadedGBufferLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, sun, target, dctx):
        albedo, sun, target = (t.contiguous().float() for t in (albedo, sun, target))
        loss, shade = dctx.forward(albedo, sun, target)
        ctx.save_for_backward(albedo, sun, target)
        ctx.dctx = dctx
        ctx.generation = dctx.generation
        ctx.mark_non_differentiable(shade)
        return loss.reshape(()), shade

    @staticmethod
    def backward(ctx, grad_loss, grad_shade):
        albedo, sun, target = ctx.saved_tensors
        dctx = ctx.dctx
        if dctx.generation != ctx.generation:   # another forward ran since: restore this one's state
            dctx.forward(albedo, sun, target)
        g = grad_loss.reshape(1).contiguous().float()
        g_albedo, g_sun = dctx.backward(albedo, sun, target, g)
        return g_albedo, g_sun, None, None


def shaded_gbuffer_loss(dctx: DiffContext, albedo, sun, target):
    """MSE between irradiance * albedo * 0.996 at the G-buffer texels and `target`; returns
    (loss, shade).  Differentiable with respect to albedo (M, 3) and sun (3,)."""
    return _ShadedGBufferLoss.apply(albedo, sun, target, dctx)
