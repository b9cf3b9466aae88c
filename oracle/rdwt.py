"""FP64 CPU oracle for the undecimated periodic 2D wavelet transform.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates the Rice Wavelet Toolbox code the reference ships:
* ``mrdwt``  shearcuda/src/mrdwt_r.h:92-207 (fpconv + MRDWT)
* ``mirdwt`` shearcuda/src/mirdwt_r.h:92-218 (bpconv + MIRDWT)
* ``make_on_filter`` shearcuda/MakeONFilter.m:55-57,111-114,174-177 (Haar,
  Daubechies 4, Symmlet 4), normalised as MakeONFilter.m:270.
* ``mdwt1``  the decimated transform at one level, shearcuda/src/mdwt_r.h:80-176
  (fpsconv + MDWT with L = 1), which separate.m:89-101 uses for its noise
  estimate.

Arrays are indexed like the MATLAB matrices (row, column); yh is m x 3nL with
level l in columns [3n(l-1), 3nl) as [lh hl hh].  The loops follow the C
code: each level splits a row (column) into sample_f interleaved
subsequences and filters each periodically with the undilated filter.

Pinned by: perfect reconstruction mirdwt(mrdwt(x)) = x, the Haar one-level
closed form, and orthonormality of the filters (tests/test_oracle_rdwt.py);
mdwt1 by the Haar closed form and energy preservation of the orthonormal
decimated transform.
"""
import numpy as np


def make_on_filter(name, par=0):
    """Orthonormal QMF (shearcuda/MakeONFilter.m)."""
    if name == "Haar":
        f = np.array([1.0, 1.0]) / np.sqrt(2.0)
    elif name == "Daubechies" and par == 4:
        f = np.array([.482962913145, .836516303738, .224143868042, -.129409522551])
    elif name == "Symmlet" and par == 4:
        f = np.array([-.107148901418, -.041910965125, .703739068656, 1.136658243408,
                      .421234534204, -.140317624179, -.017824701442, .045570345896])
    else:
        raise ValueError("filter not restated: %s %s" % (name, par))
    return f / np.linalg.norm(f)


def _fpconv(x, h0, h1):
    """fpconv (mrdwt_r.h:92-112): periodic, out[i] = sum_j x[i+j] f[lh-1-j];
    x is (..., lx), filtered along the last axis."""
    lh = len(h0)
    lx = x.shape[-1]
    lo = np.zeros_like(x)
    hi = np.zeros_like(x)
    for j in range(lh):
        xs = np.take(x, (np.arange(lx) + j) % lx, axis=-1)
        lo += xs * h0[lh - 1 - j]
        hi += xs * h1[lh - 1 - j]
    return lo, hi


def _bpconv(xl, xh, g0, g1):
    """bpconv (mirdwt_r.h:92-110): out[i] = sum_t xl[i-t] g0[t] + xh[i-t] g1[t]."""
    lh = len(g0)
    lx = xl.shape[-1]
    out = np.zeros_like(xl)
    for t in range(lh):
        idx = (np.arange(lx) - t) % lx
        out += np.take(xl, idx, axis=-1) * g0[t] + np.take(xh, idx, axis=-1) * g1[t]
    return out


def mrdwt(x, h, L):
    """[yl, yh] = mrdwt(x, h, L) for a 2D x (MRDWT, mrdwt_r.h:114-207)."""
    x = np.asarray(x, dtype=np.float64)
    m, n = x.shape
    lh = len(h)
    h0 = np.array([h[lh - 1 - i] for i in range(lh)])
    h1 = np.array(h, dtype=np.float64).copy()
    h1[0::2] = -h1[0::2]
    yl = x.copy()
    yh = np.zeros((m, 3 * n * L))
    sample_f = 1
    for lev in range(1, L + 1):
        coa = 3 * n * (lev - 1)
        # rows: each row split into sample_f interleaved subsequences
        A = np.zeros_like(yl)
        B = np.zeros_like(yl)
        for nc in range(sample_f):
            lo, hi = _fpconv(yl[:, nc::sample_f], h0, h1)
            A[:, nc::sample_f] = lo
            B[:, nc::sample_f] = hi
        # columns
        for nr in range(sample_f):
            ll, lhb = _fpconv(A[nr::sample_f, :].T, h0, h1)
            hl, hh = _fpconv(B[nr::sample_f, :].T, h0, h1)
            yl[nr::sample_f, :] = ll.T
            yh[nr::sample_f, coa:coa + n] = lhb.T
            yh[nr::sample_f, coa + n:coa + 2 * n] = hl.T
            yh[nr::sample_f, coa + 2 * n:coa + 3 * n] = hh.T
        sample_f *= 2
    return yl, yh


def mirdwt(yl, yh, h, L):
    """x = mirdwt(yl, yh, h, L) for 2D signals (MIRDWT, mirdwt_r.h:112-218)."""
    yl = np.asarray(yl, dtype=np.float64)
    yh = np.asarray(yh, dtype=np.float64)
    m, n = yl.shape
    lh = len(h)
    g0 = np.array(h, dtype=np.float64) / 2
    g1 = np.array([h[lh - 1 - i] for i in range(lh)]) / 2
    g1[1::2] = -g1[1::2]
    x = yl.copy()
    sample_f = 2 ** (L - 1)
    for lev in range(L, 0, -1):
        coa = 3 * n * (lev - 1)
        X = np.zeros_like(x)
        XH = np.zeros_like(x)
        for nr in range(sample_f):
            sl = slice(nr, None, sample_f)
            X[sl, :] = _bpconv(x[sl, :].T, yh[sl, coa:coa + n].T, g0, g1).T
            XH[sl, :] = _bpconv(yh[sl, coa + n:coa + 2 * n].T, yh[sl, coa + 2 * n:coa + 3 * n].T, g0, g1).T
        for nc in range(sample_f):
            sl = slice(nc, None, sample_f)
            x[:, sl] = _bpconv(X[:, sl], XH[:, sl], g0, g1)
        sample_f //= 2
    return x


def _fpsconv(x, h0, h1):
    """fpsconv (mdwt_r.h:80-99): periodic filtering kept at even positions,
    out[i/2] = sum_j x[i+j] f[lh-1-j] for i = 0, 2, ...; last axis."""
    lh = len(h0)
    lx = x.shape[-1]
    i = np.arange(0, lx, 2)
    lo = np.zeros(x.shape[:-1] + (lx // 2,))
    hi = np.zeros_like(lo)
    for j in range(lh):
        xs = np.take(x, (i + j) % lx, axis=-1)
        lo += xs * h0[lh - 1 - j]
        hi += xs * h1[lh - 1 - j]
    return lo, hi


def mdwt1(x, h):
    """wc = mdwt(x, h, 1) for a 2D x (MDWT, mdwt_r.h:102-176 with L = 1): rows
    split into [low | high] halves, then columns into [low ; high]; the
    finest diagonal band is wc[m/2:, n/2:]."""
    x = np.asarray(x, dtype=np.float64)
    m, n = x.shape
    lh = len(h)
    h0 = np.array([h[lh - 1 - i] for i in range(lh)])
    h1 = np.array(h, dtype=np.float64).copy()
    h1[0::2] = -h1[0::2]
    lo, hi = _fpsconv(x, h0, h1)
    y = np.concatenate([lo, hi], axis=1)
    lo, hi = _fpsconv(y.T, h0, h1)
    return np.concatenate([lo, hi], axis=1).T
