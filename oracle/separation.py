"""FP64 CPU oracle of the unrolled shearlet / wavelet separation.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The alternating transform -> threshold -> inverse structure follows
shearcuda/separate.m:162-257 (curve part by shearlets, second part by the
undecimated wavelet); the thresholds are per-band soft thresholds
(shearcuda/SoftThresh.m:11-13) and the number of steps is fixed:
    c <- rec(soft(dec(x - p), th_sh)),  p <- mirdwt(soft(mrdwt(x - c), th_w))
Used for FP64 central finite differences of the loss.

``separate_setup`` / ``separate_iter`` / ``separate`` restate the shipped
loop itself, shearcuda/separate.m:41-257 (CPU branch): re_weight_band
pre-emphasis, the noise level from the decimated D4 transform, the starting
threshold deltamax and its exponential schedule, the hard-thresholded
shearlet part with its Haar TV correction (separate.m:262-270), and the
hard-thresholded Symmlet-4 undecimated wavelet part.  ``separate_iter``
accepts keep masks for its two hard thresholds, so that a test can follow
another implementation's threshold decisions and count the flips.
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


def mad(x):
    """MAD(x) (shearcuda/MAD.m): median(|x - median(x)|) / .6745."""
    x = np.asarray(x, dtype=np.float64).ravel()
    return np.median(np.abs(x - np.median(x))) / 0.6745


def tv_correction(x, gamma):
    """TVCorrection (separate.m:262-270): soft threshold of the one-level
    undecimated Haar detail bands."""
    haar = w.make_on_filter("Haar")
    ll, wc = w.mrdwt(x, haar, 1)
    return w.mirdwt(ll, soft(wc, gamma), haar, 1)


def separate_setup(img, level2, itermax, stop, coeff, H, strict=True, wavelet=("Symmlet", 4), delta_max=0.0):
    """separate.m:56-133: the pre-emphasised image, the band norms E, the
    noise level sigma and the threshold schedule (deltamax, lambda).
    delta_max > 0 is separateHairlineCrack.m:63-66's fixed start (sigma 2.5)."""
    img = np.asarray(img, dtype=np.float64)
    N = img.shape[0]
    qmf = w.make_on_filter(*wavelet)
    E = o.com_norm(H, N, strict=strict)                                # :69
    x = o.re_weight_band(img, coeff, strict=strict)                    # :76
    if delta_max > 0:
        lam = (delta_max / (stop * 2.5)) ** (1.0 / (1 - itermax))
        return {"x": x, "E": E, "sigma": 2.5, "deltamax": delta_max, "lambda": lam, "buff": (None, None)}
    wc = w.mdwt1(x, w.make_on_filter("Daubechies", 4))                 # :81-101
    sigma = mad(wc[N // 2:N // 2 + N // 2, N // 2:N // 2 + N // 2])    # :103-104
    stop_t = stop * sigma                                              # :105
    Ct = o.nsst_dec2(x, H, strict=strict)                              # :116-117
    _, b1 = o.thresh(Ct, 0.0, 2, E, np.ones(len(H) + 1), 1)
    _, yh = w.mrdwt(x, qmf, level2)                                    # :127-128
    b2 = np.max(np.abs(yh))                                            # :131
    dmax = min(b1, b2)                                                 # :133
    lam = (dmax / stop_t) ** (1.0 / (1 - itermax))                     # :135
    return {"x": x, "E": E, "sigma": sigma, "deltamax": dmax, "lambda": lam, "buff": (b1, b2)}


def separate_iter(x, part1, part2, delta, E, H, level2, gamma, keep_sh=None, keep_w=None, strict=True,
                  wavelet=("Symmlet", 4)):
    """One outer iteration of separate.m:162-252: the residual is taken once,
    then the curve part (k = 1) and the wavelet part (k = 2) are re-estimated
    from it.  keep_sh (list of boolean arrays shaped like the coefficient
    cells) and keep_w (boolean, shaped like yh) replace the hard-threshold
    decisions when given.  Returns (part1, part2, flips) with flips the
    number of decisions that differ from this function's own."""
    qmf = w.make_on_filter(*wavelet)
    residual = x - (part1 + part2)                                     # :166
    Ra = part1 + residual                                              # :172
    Ct = o.nsst_dec2(Ra, H, strict=strict)                             # :183
    flips = 0
    out = []
    for s, c in enumerate(Ct):                                         # :185, thresh.m option 2
        c3 = c if c.ndim == 3 else c[None]
        own = np.stack([np.abs(c3[k]) > 1.4 * delta * E[s, k] for k in range(c3.shape[0])])
        keep = own if keep_sh is None else np.asarray(keep_sh[s]).reshape(own.shape)
        flips += int(np.count_nonzero(own != keep))
        out.append((c3 * keep).reshape(c.shape))
    p1 = o.nsst_rec2(out, H, strict=strict)                            # :198
    p1 = tv_correction(p1, gamma)                                      # :200
    Ra = part2 + residual
    yl, yh = w.mrdwt(Ra, qmf, level2)                                  # :221
    own = np.abs(yh) > 1.4 * delta                                     # :224, thresh.m option 1
    keep = own if keep_w is None else np.asarray(keep_w)
    flips += int(np.count_nonzero(own != keep))
    p2 = w.mirdwt(yl, yh * keep, qmf, level2)                          # :225
    return p1, p2, flips


def separate(img, level2, itermax, stop, gamma, coeff, H, strict=True):
    """[C, P] = separate(img, level2, itermax, stop, gamma, coeff, 0, H)."""
    st = separate_setup(img, level2, itermax, stop, coeff, H, strict)
    x = st["x"]
    part1 = np.zeros_like(x)
    part2 = np.zeros_like(x)
    delta = st["deltamax"]
    for _ in range(itermax):
        part1, part2, _ = separate_iter(x, part1, part2, delta, st["E"], H, level2, gamma, strict=strict)
        delta *= st["lambda"]                                          # :252
    return part1, part2
