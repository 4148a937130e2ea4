"""torch.autograd wrapper of the bake's differentiable part, for the tests.

The bake maps light colours, per-patch albedo and per-patch emission to the
per-patch light after the bounces:

    E = ambient + sum_l colour_l * w_l + emission      (direct light, checkpointed w)
    T = sum_{k=0..B} K^k E,  K = F diag(area) diag(albedo_c)

Forward and backward run on the GPU through the lgx_* C-ABI:
lgx_RadSolveGrad (reverse mode of the bounces) and lgx_DirectLightGrad
(reverse mode of the light accumulation).  The weights w come from the
device's last direct-light run (lgx_ReadWeights); patch i is device sample i.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from lighter_amd import _lib

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)


def _d(a):
    return a.ctypes.data_as(dp)


class BakeContext:
    """Device, direct-light work and patch data shared by forward and backward."""

    def __init__(self, lib, dev, items, light_off, weights, area, form=0, bounces=4, ambient=None):
        self.lib, self.dev = lib, dev
        self.items = np.asarray(items, np.uint32)
        self.light_off = np.asarray(light_off, np.uint32)
        self.w = np.asarray(weights, np.float32)
        self.area = np.ascontiguousarray(area, np.float32)
        self.N = len(self.area)
        self.form, self.bounces = form, bounces
        self.ambient = np.zeros((self.N, 3)) if ambient is None else np.asarray(ambient, np.float64)
        skip = self.w.view(np.uint32) == np.uint32(0xFFFFFFFF)
        self.wv = np.where(skip, 0.0, self.w.astype(np.float64))
        self.light_of = self.items >> np.uint32(26)
        self.sample_of = (self.items & np.uint32((1 << 26) - 1)).astype(np.int64)

    def direct(self, colors):
        """E without emission: ambient + sum_l colour_l * w_l (FP64)."""
        E = self.ambient.copy()
        contrib = self.wv[:, None] * colors[self.light_of]
        np.add.at(E, self.sample_of, contrib)
        return E

    def radiosity(self, E, albedo, g=None):
        N = self.N
        E = np.ascontiguousarray(E, np.float64)
        alb = np.ascontiguousarray(albedo, np.float32)
        total = np.zeros((N, 3))
        t = (C.c_double * 4)()
        if g is None:
            _lib.check(self.lib.lgx_RadSolve(self.dev, _d(E), 3, alb.ctypes.data_as(fp),
                                             self.area.ctypes.data_as(fp), self.form, self.bounces, 0,
                                             _d(total), t), self.lib)
            return total
        g = np.ascontiguousarray(g, np.float64)
        lam = np.zeros((N, 3))
        dalb = np.zeros((N, 3))
        _lib.check(self.lib.lgx_RadSolveGrad(self.dev, _d(E), 3, alb.ctypes.data_as(fp),
                                             self.area.ctypes.data_as(fp), self.form, self.bounces, _d(g),
                                             _d(total), _d(lam), _d(dalb), t), self.lib)
        return total, lam, dalb

    def color_grad(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        L = len(self.light_off) - 1
        out = np.zeros((L, 3))
        _lib.check(self.lib.lgx_DirectLightGrad(self.dev, _d(lam), _d(out)), self.lib)
        return out


class BakeFn(torch.autograd.Function):
    """(colours L x 3, albedo N x 3, emission N x 3) -> T (N x 3), FP64."""

    @staticmethod
    def forward(ctx, colors, albedo, emission, bake: BakeContext):
        E = bake.direct(colors.detach().cpu().numpy()) + emission.detach().cpu().numpy()
        alb = albedo.detach().cpu().numpy()
        T = bake.radiosity(E, alb)
        ctx.bake = bake
        ctx.save_for_backward(colors, albedo, emission)
        ctx.E = E
        return torch.from_numpy(T)

    @staticmethod
    def backward(ctx, gT):
        bake = ctx.bake
        _, albedo, _ = ctx.saved_tensors
        _, lam, dalb = bake.radiosity(ctx.E, albedo.detach().cpu().numpy(), gT.detach().cpu().numpy())
        dcol = bake.color_grad(lam)
        return torch.from_numpy(dcol), torch.from_numpy(dalb), torch.from_numpy(lam), None
