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


def _check_layout(x, th, nimg, img_elems, rowlen, seglen, band_off, *same):
    """Buffers of the per-band threshold kernels: x holds nimg images of
    img_elems elements; element e of an image belongs to band
    band_off + (e mod rowlen) // seglen, whose threshold th holds (one table
    for all images); each tensor in `same` has x's shape."""
    api._require_gpu(x, "x")
    api._require_gpu(th, "th")
    if nimg < 1 or img_elems < 1 or rowlen < 1 or seglen < 1 or img_elems % rowlen or rowlen % seglen:
        raise ValueError("invalid band layout")
    if x.numel() != nimg * img_elems:
        raise ValueError("x holds %d elements, the layout %d" % (x.numel(), nimg * img_elems))
    if band_off < 0 or th.numel() < band_off + rowlen // seglen:
        raise ValueError("th must hold at least %d thresholds" % (band_off + rowlen // seglen))
    for t in same:
        api._require_shape(t, x.shape, "buffer")


def soft_bands(x, th, nimg, img_elems, rowlen, seglen, band_off=0, out=None):
    out = torch.empty_like(x) if out is None else out
    _check_layout(x, th, nimg, img_elems, rowlen, seglen, band_off, out)
    check(lib.shx_soft_bands(_vp(x), _vp(out), nimg, img_elems, rowlen, seglen, band_off, _vp(th), _stream()))
    return out


def soft_bands_backward(x, gy, th, gth, nimg, img_elems, rowlen, seglen, band_off=0):
    gx = torch.empty_like(x)
    _check_layout(x, th, nimg, img_elems, rowlen, seglen, band_off, gy)
    api._require_shape(gth, th.shape, "gth")
    check(lib.shx_soft_bands_backward(_vp(x), _vp(gy), _vp(gx), _vp(gth), nimg, img_elems, rowlen, seglen, band_off,
                                      _vp(th), _stream()))
    return gx


def hard_bands(x, th, nimg, img_elems, rowlen, seglen, band_off=0, out=None):
    out = torch.empty_like(x) if out is None else out
    _check_layout(x, th, nimg, img_elems, rowlen, seglen, band_off, out)
    check(lib.shx_hard_bands(_vp(x), _vp(out), nimg, img_elems, rowlen, seglen, band_off, _vp(th), _stream()))
    return out


def hard_bands_backward(x, gy, th, nimg, img_elems, rowlen, seglen, band_off=0):
    gx = torch.empty_like(x)
    _check_layout(x, th, nimg, img_elems, rowlen, seglen, band_off, gy)
    check(lib.shx_hard_bands_backward(_vp(x), _vp(gy), _vp(gx), nimg, img_elems, rowlen, seglen, band_off, _vp(th),
                                      _stream()))
    return gx


class ShearShrink(torch.autograd.Function):
    """a (B, L, L) -> S_sh(soft(A_sh(a), th)) with per-band thresholds th,
    (nbands,) shared by the images or (B, nbands).

    Forward: one shx_batch_shrink call, the soft threshold fused into the
    reconstruction's reads; when a gradient is wanted it also keeps the
    coefficients A_sh(a).  Backward: one shx_batch_shrink_backward call, the
    threshold's mask and gradient fused between S^T and A^T."""

    @staticmethod
    def forward(ctx, a, th, plan, keep=True):
        a = a.contiguous()
        th = th.contiguous()
        out = torch.empty_like(a)
        C = plan.alloc_coeffs() if keep else None
        plan.shrink(a, th, out, C)
        ctx.plan = plan
        if keep:
            ctx.save_for_backward(C, th)
        return out

    @staticmethod
    def backward(ctx, g):
        if not ctx.saved_tensors:
            raise RuntimeError("ShearShrink ran without keeping its coefficients (keep=False)")
        C, th = ctx.saved_tensors
        g = g.contiguous()
        ga = torch.empty_like(g)
        gth = torch.zeros_like(th)
        ctx.plan.shrink_backward(C, th, g, ga, gth)
        return ga, gth, None, None


def shear_shrink(a, th, plan):
    """ShearShrink.apply, keeping the coefficients only when autograd will
    call the backward."""
    keep = torch.is_grad_enabled() and (a.requires_grad or th.requires_grad)
    return ShearShrink.apply(a, th, plan, keep)


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
        c = shear_shrink(x - p, th_sh, plan)
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


def tv_correction_cuda(x, gamma, haar=None):
    """TVCorrection_cuda (shearcuda/separate.m:272-278): one-level undecimated
    Haar transform, soft threshold of its detail bands, inverse."""
    from .wavelets import MakeONFilter
    haar = MakeONFilter("Haar") if haar is None else haar
    ll, wc = api.mrdwt_cuda(x, haar, 1)
    api.softth_cuda(wc, gamma)
    return api.mirdwt_cuda(ll, wc, haar, 1)


def separate(img, level2, itermax, stop, gamma, coeff, shear, trace=None, info=None):
    """[C, P] = separate(img, level2, itermax, stop, gamma, coeff, 0, shear)
    (shearcuda/separate.m:1-257, GPU branch): curve part C by hard-thresholded
    shearlets with a Haar TV correction, point part P by the hard-thresholded
    undecimated Symmlet-4 wavelet, on an exponentially decreasing threshold.

    Every transform, threshold and reduction is a shx call; torch supplies
    the buffers and the elementwise sums of the iteration.  The noise level
    sigma is the MAD of the finest diagonal band of a one-level decimated D4
    transform (separate.m:86-104, computed on the host in the reference); that
    band is the undecimated band at even rows and columns, so it is taken from
    shx_mrdwt and its median from shx_band_mad.

    trace, if given, is called after every iteration with a dict holding the
    iteration's inputs (part1_in, part2_in, delta) and results (part1, part2),
    the thresholded shearlet cells (coeff) and wavelet details (yh).  info, if
    a dict, receives the schedule: sigma, deltamax, lambda and the two
    maxima (buff) deltamax is the smaller of.  Returns (C, P)."""
    C, P, _, _ = _separate(img, level2, itermax, stop, gamma, coeff, shear, ("Symmlet", 4), 0.0, None, trace, info)
    return C, P


def separate_hairline_crack(img, level2, param, shear, E=None, trace=None, info=None):
    """[C, P, residual, Ct] = separateHairlineCrack(img, level2, param, 0,
    shear, 0, E) (crack_detection/separateHairlineCrack.m): separate() with
    param.numIter / stop / gamma / coeff, an optional wavelet
    (param.wavelet_name / wavelet_val, default Symmlet 4), an optional fixed
    starting threshold param.deltaMax > 0 (then sigma is taken as 2.5 and no
    transform estimates it), and band norms E passed in.  residual is the last
    iteration's residual and Ct its thresholded shearlet cells."""
    wav = (param.get("wavelet_name", "Symmlet"), param.get("wavelet_val", 4))
    return _separate(img, level2, int(param["numIter"]), float(param["stop"]), float(param["gamma"]),
                     param["coeff"], shear, wav, float(param.get("deltaMax", 0.0)), E, trace, info)


def _separate(img, level2, itermax, stop, gamma, coeff, shear, wavelet, delta_max, E, trace, info):
    from .wavelets import MakeONFilter
    L = shear.L
    if tuple(img.shape) != (L, L):
        raise ValueError("image must be %dx%d" % (L, L))
    if itermax < 2:
        raise ValueError("itermax must be at least 2 (the schedule divides by 1 - itermax)")
    qmf = MakeONFilter(*wavelet)
    haar = MakeONFilter("Haar")
    ones = np.ones(shear.nscales + 1)
    if E is None:
        E = api.com_norm("maxflat", (L, L), shear)                               # separate.m:69
    x = api.re_weight_band_cuda(img.contiguous(), coeff, shear)                  # :74
    Ct = shear.alloc_coeffs()
    if delta_max > 0:                                                            # separateHairlineCrack.m:63-66
        sigma, b1, b2 = 2.5, None, None
        stop_t = stop * sigma
        delta = delta_max
    else:
        _, yh1 = api.mrdwt_cuda(x, MakeONFilter("Daubechies", 4), 1)             # separate.m:81-101
        hh = yh1[0::2, 2 * L::2].contiguous().reshape(1, -1)
        sigma = float(api.band_mad_cuda(hh).cpu()[0])                            # :103-104
        stop_t = stop * sigma                                                    # :105
        api.shear_trans_cuda(x, "maxflat", shear, Ct)                            # :116-117
        b1 = api.thresh_cuda(Ct, 0.0, 2, E, ones, 1)
        Ctw = list(api.mrdwt_cuda(x, qmf, level2))                               # :127-128
        b2 = api.thresh_cuda(Ctw, 0.0, 1, E, np.zeros(1), 1)                     # :131, max |Ctw{2}|
        delta = min(b1, b2)                                                      # :133
    dmax = delta
    lam = (delta / stop_t) ** (1.0 / (1 - itermax))                              # :135
    part1 = torch.zeros_like(x)
    part2 = torch.zeros_like(x)
    residual = x
    for it in range(itermax):
        residual = x - part1 - part2                                             # :164-166
        api.shear_trans_cuda(part1 + residual, "maxflat", shear, Ct)             # :175-177
        api.thresh_cuda(Ct, 1.4 * delta, 2, E, ones, 0)
        p1 = tv_correction_cuda(api.inverse_shear_cuda(Ct, "maxflat", shear), gamma, haar)  # :187-191
        yl, yh = api.mrdwt_cuda(part2 + residual, qmf, level2)                   # :206-212
        api.thresh_cuda([yl, yh], 1.4 * delta, 1, E, np.zeros(1), 0)
        p2 = api.mirdwt_cuda(yl, yh, qmf, level2)
        if trace is not None:
            trace({"iter": it, "delta": delta, "part1_in": part1, "part2_in": part2, "part1": p1, "part2": p2,
                   "coeff": [c.clone() for c in Ct], "yh": yh})
        part1, part2 = p1, p2
        delta *= lam                                                             # :252
    if info is not None:
        info.update({"sigma": sigma, "deltamax": dmax, "lambda": lam, "buff": (b1, b2)})
    return part1, part2, residual, Ct


def separate_native(img, level2, itermax, stop, gamma, coeff, shear, wavelet=("Symmlet", 4), delta_max=0.0, E=None,
                    info=None):
    """The same loop as separate() / separate_hairline_crack() run inside the
    library (shx_separate): one call, one workspace, no per-step host work.
    Returns (C, P); info, if a dict, receives sigma, deltamax and lambda."""
    from .wavelets import MakeONFilter
    api._require_gpu(img, "img")
    L = shear.L
    if tuple(img.shape) != (L, L) or img.dtype != torch.float32:
        raise ValueError("image must be a %dx%d float32 tensor" % (L, L))
    img = img.contiguous()
    qmf = np.ascontiguousarray(MakeONFilter(*wavelet), dtype=np.float64)
    coeff = np.ascontiguousarray(np.asarray(coeff, dtype=np.float64).ravel())
    dp = ctypes.POINTER(ctypes.c_double)
    if E is not None:
        Ef = np.asfortranarray(np.asarray(E, dtype=np.float64))
        Eflat = np.ascontiguousarray(Ef.ravel(order="F"))
        if Ef.ndim != 2 or Ef.shape[0] != shear.nscales + 1:
            raise ValueError("E must have %d rows (lowpass and one per scale)" % (shear.nscales + 1))
        eptr, erows, ecols = Eflat.ctypes.data_as(dp), Ef.shape[0], Ef.shape[1]
    else:
        Eflat, eptr, erows, ecols = None, None, 0, 0
    C = torch.empty_like(img)
    P = torch.empty_like(img)
    inf = np.zeros(3)
    check(lib.shx_separate(shear.handle, _vp(img), int(level2), int(itermax), float(stop), float(gamma),
                           coeff.ctypes.data_as(dp), coeff.size, qmf.ctypes.data_as(dp), qmf.size, float(delta_max),
                           eptr, erows, ecols, _vp(C), _vp(P), inf.ctypes.data_as(dp), _stream()))
    if info is not None:
        info.update({"sigma": float(inf[0]), "deltamax": float(inf[1]), "lambda": float(inf[2])})
    return C, P
