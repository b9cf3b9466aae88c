"""Unrolled shearlet / wavelet image separation with learnable per-band soft
thresholds (gradient harness).

The reference separates an image into a curve part (shearlets) and a second
part (undecimated wavelets) by alternating transform -> threshold -> inverse
steps (shearcuda/separate.m:162-257, hard thresholds on a fixed schedule).
This unrolled form runs K alternating soft-threshold steps
    c <- S_sh(soft(A_sh(x - p), th_sh)),   p <- S_w(soft(A_w(x - c), th_w))
from c = p = 0.  Gradients reach x and the thresholds through the
hand-written adjoint kernels of the C ABI.  torch.autograd.Function is used
here only to chain those kernels; every forward and backward operator is a
shx call.  Each step's coefficients are recomputed in the backward pass
rather than stored.
"""
import ctypes

import numpy as np
import torch

from . import api
from ._lib import check, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def soft_bands(x, th, nimg, img_elems, rowlen, seglen, band_off=0, out=None):
    out = torch.empty_like(x) if out is None else out
    check(lib.shx_soft_bands(_vp(x), _vp(out), nimg, img_elems, rowlen, seglen, band_off, _vp(th), _stream()))
    return out


def soft_bands_backward(x, gy, th, gth, nimg, img_elems, rowlen, seglen, band_off=0):
    gx = torch.empty_like(x)
    check(lib.shx_soft_bands_backward(_vp(x), _vp(gy), _vp(gx), _vp(gth), nimg, img_elems, rowlen, seglen, band_off,
                                      _vp(th), _stream()))
    return gx


def hard_bands(x, th, nimg, img_elems, rowlen, seglen, band_off=0, out=None):
    out = torch.empty_like(x) if out is None else out
    check(lib.shx_hard_bands(_vp(x), _vp(out), nimg, img_elems, rowlen, seglen, band_off, _vp(th), _stream()))
    return out


def hard_bands_backward(x, gy, th, nimg, img_elems, rowlen, seglen, band_off=0):
    gx = torch.empty_like(x)
    check(lib.shx_hard_bands_backward(_vp(x), _vp(gy), _vp(gx), nimg, img_elems, rowlen, seglen, band_off, _vp(th),
                                      _stream()))
    return gx


class ShearShrink(torch.autograd.Function):
    """a (B, L, L) -> S_sh(soft(A_sh(a), th)) with per-band thresholds th."""

    @staticmethod
    def forward(ctx, a, th, plan):
        a = a.contiguous()
        B, L = a.shape[0], a.shape[-1]
        C = plan.alloc_coeffs()
        plan.shear_trans(a, C)
        n = C[0].numel()
        Cs = soft_bands(C, th, B, n, n, L * L)
        out = torch.empty_like(a)
        plan.inverse_shear(Cs, out)
        ctx.save_for_backward(a, th)
        ctx.plan = plan
        return out

    @staticmethod
    def backward(ctx, g):
        a, th = ctx.saved_tensors
        plan = ctx.plan
        g = g.contiguous()
        B, L = a.shape[0], a.shape[-1]
        C = plan.alloc_coeffs()
        plan.shear_trans(a, C)                      # recomputed bands
        Gs = plan.alloc_coeffs()
        plan.inverse_shear_adjoint(g, Gs)           # S^T g
        n = C[0].numel()
        gth = torch.zeros_like(th)
        gC = soft_bands_backward(C, Gs, th, gth, B, n, n, L * L)
        ga = torch.empty_like(a)
        plan.shear_trans_adjoint(gC, ga)            # A^T
        return ga, gth, None


class WaveShrink(torch.autograd.Function):
    """b (B, m, n) -> S_w(soft(A_w(b), th)); th: 1 + 3*levels thresholds
    (yl first, then [lh hl hh] per level)."""

    @staticmethod
    def forward(ctx, b, th, h, levels):
        b = b.contiguous()
        B, m, n = b.shape
        yl, yh = api.mrdwt_cuda(b, h, levels)
        yls = soft_bands(yl, th, B, m * n, m * n, m * n, 0)
        yhs = soft_bands(yh, th, B, m * 3 * n * levels, 3 * n * levels, n, 1)
        out = api.mirdwt_cuda(yls, yhs, h, levels)
        ctx.save_for_backward(b, th)
        ctx.h, ctx.levels = h, levels
        return out

    @staticmethod
    def backward(ctx, g):
        b, th = ctx.saved_tensors
        h, levels = ctx.h, ctx.levels
        B, m, n = b.shape
        yl, yh = api.mrdwt_cuda(b, h, levels)       # recomputed bands
        gyl, gyh = api.mirdwt_adjoint_cuda(g.contiguous(), h, levels)
        gth = torch.zeros_like(th)
        gyl = soft_bands_backward(yl, gyl, th, gth, B, m * n, m * n, m * n, 0)
        gyh = soft_bands_backward(yh, gyh, th, gth, B, m * 3 * n * levels, 3 * n * levels, n, 1)
        gb = api.mrdwt_adjoint_cuda(gyl, gyh, h, levels)
        return gb, gth, None, None


def unrolled_separation(x, th_sh, th_w, plan, h, levels=3, iters=10):
    """K alternating soft-threshold steps; returns (curve part c, other part p)."""
    c = torch.zeros_like(x)
    p = torch.zeros_like(x)
    for _ in range(iters):
        c = ShearShrink.apply(x - p, th_sh, plan)
        p = WaveShrink.apply(x - c, th_w, h, levels)
    return c, p


def wavelet_scalars(noise, h, levels):
    """MAD/0.6745 of each wavelet band of a unit white-noise image (exact
    medians on the device, shx_band_mad); order: yl, then [lh hl hh] per level."""
    yl, yh = api.mrdwt_cuda(noise, h, levels)
    m, n = noise.shape[-2:]
    bands = yh.reshape(m, 3 * levels, n).permute(1, 0, 2).contiguous()
    return np.concatenate([api.band_mad_cuda(yl.reshape(1, m, n).contiguous()).cpu().numpy(),
                           api.band_mad_cuda(bands).cpu().numpy()])
