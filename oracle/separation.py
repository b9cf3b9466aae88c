"""FP64 CPU oracle of the unrolled shearlet / wavelet separation.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The alternating transform -> threshold -> inverse structure follows
shearcuda/separate.m:162-257 (curve part by shearlets, second part by the
undecimated wavelet); the thresholds are per-band soft thresholds
(shearcuda/SoftThresh.m:11-13) and the number of steps is fixed:
    c <- rec(soft(dec(x - p), th_sh)),  p <- mirdwt(soft(mrdwt(x - c), th_w))
Used for FP64 central finite differences of the loss.
"""
import numpy as np

from . import nsst as o
from . import rdwt as w


def soft(v, t):
    return np.sign(v) * np.maximum(np.abs(v) - t, 0.0)


def shear_shrink(a, H, th):
    d = o.nsst_dec2(a, H, strict=False)
    d = [soft(d[0], th[0])] + [soft(b, np.asarray(t)[:, None, None]) for b, t in
                               zip(d[1:], np.split(np.asarray(th[1:]), np.cumsum([h.shape[0] for h in H])[:-1]))]
    return o.nsst_rec2(d, H, strict=False)


def wave_shrink(b, h, levels, th):
    yl, yh = w.mrdwt(b, h, levels)
    m, n = b.shape
    yl = soft(yl, th[0])
    yh = yh.copy()
    for k in range(3 * levels):
        yh[:, k * n:(k + 1) * n] = soft(yh[:, k * n:(k + 1) * n], th[1 + k])
    return w.mirdwt(yl, yh, h, levels)


def unrolled(x, H, th_sh, h, levels, th_w, iters):
    c = np.zeros_like(x)
    p = np.zeros_like(x)
    for _ in range(iters):
        c = shear_shrink(x - p, H, th_sh)
        p = wave_shrink(x - c, h, levels, th_w)
    return c, p
