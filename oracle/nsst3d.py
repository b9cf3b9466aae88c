"""FP64 CPU oracle for the 3D shearlet volume transform.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP library is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``shearcuda_amd``) never does.

It restates the reference's MATLAB 3D path in numpy, in FP64, one function per
MATLAB file, each citing the file:line it follows (paths under
``shearcuda3d/``):

* ``rcos``               3DBP/rcos.m:10-13
* ``pyr_params``         3DBP/PyrNDDec_mm.m:52-70 (Pyr_mode 1, 1.5, 2)
* ``passband``           3DBP/PrySDdec_onestep.m:63-88 (1D factor of Lf)
* ``pyr_dec_onestep``    3DBP/PrySDdec_onestep.m:56-125
* ``pyr_rec_onestep``    3DBP/PrySDrec_onestep.m:56-118
* ``pyr_nd_dec``         3DBP/PyrNDDec_mm.m:48-109 (with ccsym.m:20-52 as rfftn)
* ``pyr_nd_rec``         3DBP/PyrNDRec_mm.m:46-95
* ``generate_xyz``       3DShearTrans/GenerateXYZ.m:12-54
* ``pyramid_sections``   3DShearTrans/GeneratePyramidSection.m:1-14
* ``polar_to_rec``       3DShearTrans/PolarToRec.m:10-19 (last write wins)
* ``meyer_filters_3d``   3DShearTrans/GetMeyerBasedFilter.m:12-151
* ``convnfft_same``      CONVNFFT_Folder/convnfft.m ('same': offset ceil((n-1)/2))
* ``sh_dec``             3DShearTrans/ShDec.m:18-34
* ``threshold_sh_coeff`` DenoiseDemo/ThresholdShCoeff.m:1-10
* ``sh_rec``             3DShearTrans/ShRec.m:11-23
* ``dec_thresh_rec``     3DShearTrans/src_cuda/ShDecThRec_cuda.cpp:120-171,178-335
* ``nsst_scalars_3d``    DenoiseDemo/NsstScalars.m:1-19 (noise passed in)
* ``denoise_3d``         DenoiseDemo/main.m:84-150

Volumes are numpy arrays indexed like the MATLAB arrays (``x[i, j, k]`` is
``X(i+1, j+1, k+1)``).  The transform is the same in any memory order as long
as volumes and filters share it.

How each function is pinned (no MATLAB exists here):

* band-pass: rcos closed form; Lf^2 + Hf^2 = 1, so pyr_nd_rec(pyr_nd_dec(x)) = x
  to 1e-12 for every mode; Lf vanishes on the edge of its box.
* generate_xyz: the closed form X = k, Y = round(i + (k-1)(n+1-2i)/(n-1)),
  Z = round(j + (k-1)(n+1-2j)/(n-1)) (derived from GenerateXYZ.m's geometry).
* meyer_filters_3d: the windows over all cones and directions sum to one at
  every frequency, so the spatial filters sum to a unit impulse at the centre
  voxel; each cone's windows lie in that cone's pyramid; the reference's
  own noise-scalar fixtures (DenoiseDemo/nsstScalarsData*.mat, Pyr_mode 1.5)
  agree with the band norms of this restatement.
* convnfft_same: scipy.signal.convolve (direct and FFT methods) sliced at the
  'same' offset.
* dec_thresh_rec / denoise_3d: compositions of the above; thresholding checked
  against a closed form on small arrays.
"""
import math

import numpy as np
from scipy.fft import next_fast_len


# --- 3DBP: Fourier-domain multiscale band-pass pyramid --------------------------
def rcos(x):
    """Raised cosine 0.5 (1 - cos(pi x)) clipped to [0, 1] (3DBP/rcos.m:10-13)."""
    x = np.asarray(x, dtype=np.float64)
    t = 0.5 * (1.0 - np.cos(np.pi * x))
    t = np.where(x <= 0, 0.0, t)
    return np.where(x >= 1, 1.0, t)


def pyr_params(mode, nlevels):
    """(w, tbw, D) per level for Pyr_mode 1, 1.5 or 2 (3DBP/PyrNDDec_mm.m:52-70)."""
    L = nlevels
    if mode == 1:
        return [0.5] + [0.25] * (L - 1), [1 / 6] + [1 / 12] * (L - 1), [1] + [2] * (L - 1)
    if mode == 1.5:
        return [0.5] + [3 / 8] * (L - 1), [1 / 7] + [1 / 9] * (L - 1), [1.5] + [2] * (L - 1)
    if mode == 2:
        return [1 / 3] * L, [1 / 7] * L, [2] * L
    raise ValueError("Unsupported Pyr mode.")


def passband(nr, w, tbw):
    """Box half-width P and the 1D passband values v[0..P] of one axis
    (3DBP/PrySDdec_onestep.m:73-79): v = sqrt(rcos((P - k) / (P - floor(nr/2 (w - tbw)))))."""
    P = int(math.ceil(nr / 2 * (w + tbw)))
    k = np.arange(P + 1, dtype=np.float64)
    den = math.ceil(nr / 2 * (w + tbw)) - math.floor(nr / 2 * (w - tbw))
    return P, np.sqrt(rcos((P - k) / den))


def _box(nr, P, half):
    """Frequency indices of the passband box on one axis and their |f|."""
    if half:
        idx = np.arange(P + 1)
        return idx, idx
    idx = np.concatenate([np.arange(P + 1), np.arange(nr - P, nr)])
    return idx, np.concatenate([np.arange(P + 1), np.arange(P, 0, -1)])


def _lowpass_box(szF, w, tbw):
    """Index arrays of the box per axis and Lf on it (PrySDdec_onestep.m:63-88)."""
    nd = len(szF)
    idx, vals = [], []
    for d in range(nd):
        P, v = passband(szF[d], w, tbw)
        ix, af = _box(szF[d], P, d == nd - 1)
        idx.append(ix)
        vals.append(v[af])
    Lf = vals[0][:, None, None] * vals[1][None, :, None] * vals[2][None, None, :]
    return idx, Lf


def pyr_dec_onestep(X, w, tbw, D):
    """One analysis step on a half spectrum X (last axis halved)
    (3DBP/PrySDdec_onestep.m:56-125).  Returns (Lp, Hp)."""
    szX = X.shape
    szF = list(szX)
    szF[-1] = (szX[-1] - 1) * 2
    idx, Lf = _lowpass_box(szF, w, tbw)
    szLp = [s / D for s in szF]
    if any(int(s) != s for s in szLp):
        raise ValueError("The downsampling factor must be able to divide the size of the FFT matrix!")
    szLp = [int(s) for s in szLp]
    szLp[-1] = szLp[-1] // 2 + 1
    sml = []
    for d in range(3):
        ix = idx[d].copy()
        if d < 2:
            half = (len(ix) + 3) // 2 - 1  # 0-based start of the negative part
            ix[half:] += szLp[d] - szX[d]
        sml.append(ix)
    Lp = np.zeros(szLp, dtype=np.complex128)
    Lp[np.ix_(*sml)] = (Lf / D ** 1.5) * X[np.ix_(*idx)]
    Hp = X.copy()
    Hp[np.ix_(*idx)] = np.sqrt(1.0 - Lf ** 2) * X[np.ix_(*idx)]
    return Lp, Hp


def pyr_rec_onestep(Lp, Hp, w, tbw, D):
    """One synthesis step (3DBP/PrySDrec_onestep.m:56-118)."""
    szX = Hp.shape
    szF = list(szX)
    szF[-1] = (szX[-1] - 1) * 2
    idx, Lf = _lowpass_box(szF, w, tbw)
    rec = Hp.copy()
    rec[np.ix_(*idx)] = np.sqrt(1.0 - Lf ** 2) * rec[np.ix_(*idx)]
    szLp = [int(s / D) for s in szF]
    szLp[-1] = szLp[-1] // 2 + 1
    sml = []
    for d in range(3):
        ix = idx[d].copy()
        if d < 2:
            half = (len(ix) + 3) // 2 - 1
            ix[half:] += szLp[d] - szX[d]
        sml.append(ix)
    rec[np.ix_(*idx)] += Lp[np.ix_(*sml)] * (Lf * D ** 1.5)
    return rec


def pyr_nd_dec(x, nlevels, mode=1):
    """Band-pass subbands, finest first, lowpass last (3DBP/PyrNDDec_mm.m:48-109,
    ccsym.m:20-52: the compact half spectrum and real(ifftn(expand)) are
    numpy's rfftn / irfftn)."""
    w, tbw, D = pyr_params(mode, nlevels)
    X = np.fft.rfftn(np.asarray(x, dtype=np.float64))
    subs = []
    for n in range(nlevels):
        Lp, Hp = pyr_dec_onestep(X, w[n], tbw[n], D[n])
        X = Lp
        full = (Hp.shape[0], Hp.shape[1], (Hp.shape[2] - 1) * 2)
        subs.append(np.fft.irfftn(Hp, s=full, axes=(0, 1, 2)))
    full = (X.shape[0], X.shape[1], (X.shape[2] - 1) * 2)
    subs.append(np.fft.irfftn(X, s=full, axes=(0, 1, 2)))
    return subs


def pyr_nd_rec(subs, mode=1):
    """Reconstruction from band-pass subbands (3DBP/PyrNDRec_mm.m:46-95)."""
    L = len(subs) - 1
    w, tbw, D = pyr_params(mode, L)
    Lp = np.fft.rfftn(subs[-1])
    for n in range(L - 1, -1, -1):
        Hp = np.fft.rfftn(subs[n])
        Lp = pyr_rec_onestep(Lp, Hp, w[n], tbw[n], D[n])
    full = (Lp.shape[0], Lp.shape[1], (Lp.shape[2] - 1) * 2)
    return np.fft.irfftn(Lp, s=full, axes=(0, 1, 2))


def level_sizes(n, nlevels, mode=1):
    """Side of each band-pass subband (finest first) and of the lowpass."""
    _, _, D = pyr_params(mode, nlevels)
    sizes, s = [], float(n)
    for d in D:
        sizes.append(int(s))
        s = s / d
    sizes.append(int(s))
    return sizes


# --- 3DShearTrans: pseudo-polar grid and Meyer-based cone windows ----------------
def _mround(v):
    """MATLAB's conversion of a nonnegative double to uint16 (round half away)."""
    return np.floor(np.asarray(v, dtype=np.float64) + 0.5).astype(np.int64)


def generate_xyz(n):
    """1-based Cartesian coordinates X, Y, Z (each n x n x n, indexed (i, k, j))
    of the pseudo-polar grid points (3DShearTrans/GenerateXYZ.m:12-54)."""
    X = np.zeros((n, n, n), dtype=np.int64)
    Y = np.zeros((n, n, n), dtype=np.int64)
    Z = np.zeros((n, n, n), dtype=np.int64)
    mid = int(math.ceil(n / 2))
    for j in range(1, mid + 1):
        for i in range(1, mid + 1):
            d2 = (n - 1) ** 2 + (n - i + 1 - i) ** 2
            d3 = (d2 + (n - j + 1 - j) ** 2) ** 0.5
            d2 = d2 ** 0.5
            sin_a = (n - j + 1 - j) / d3
            cos_a = d2 / d3
            sin_b = (n - i + 1 - i) / d2
            cos_b = (n - 1) / d2
            dl3 = d3 / (n - 1)
            I, J, Im, Jm = i - 1, j - 1, n - i, n - j
            for k in range(1, n + 1):
                K = k - 1
                l = dl3 * (k - 1)
                temp = l * sin_a
                Z[I, K, J] = _mround(j + temp)
                Z[Im, K, J] = Z[I, K, J]
                Z[I, K, Jm] = _mround(n - j + 1 - temp)
                Z[Im, K, Jm] = _mround(n - j + 1 - temp)
                temp = l * cos_a
                t1 = temp * sin_b
                Y[I, K, J] = _mround(i + t1)
                Y[Im, K, J] = _mround(n - i + 1 - t1)
                X[I, K, J] = _mround(1 + temp * cos_b)
                X[Im, K, J] = X[I, K, J]
                X[I, K, Jm] = X[I, K, J]
                X[Im, K, Jm] = X[I, K, J]
                Y[I, K, Jm] = Y[I, K, J]
                Y[Im, K, Jm] = Y[Im, K, J]
    return X, Y, Z


def pyramid_sections(n):
    """The three cones' coordinate triples (3DShearTrans/GeneratePyramidSection.m:1-14)."""
    X, Y, Z = generate_xyz(n)
    return [(X, Y, Z), (Y, X, Z), (Z, Y, X)]


def polar_to_rec(ranges, weights, n, section):
    """Window on the n^3 grid: the polar points (i, k, j) in the index ranges get
    the weights; where several land on one voxel the last in MATLAB's loop order
    (j outer, i middle, k inner) wins (3DShearTrans/PolarToRec.m:10-19).
    ranges: ((i0, i1), (j0, j1)) 1-based inclusive; weights[i-i0, k, j-j0]."""
    X, Y, Z = section
    (i0, i1), (j0, j1) = ranges
    jj, ii, kk = np.meshgrid(np.arange(j0, j1 + 1), np.arange(i0, i1 + 1), np.arange(1, n + 1), indexing="ij")
    jj, ii, kk = jj.ravel(), ii.ravel(), kk.ravel()  # loop order: j outer, i, k inner
    vox = ((X[ii - 1, kk - 1, jj - 1] - 1) * n + (Y[ii - 1, kk - 1, jj - 1] - 1)) * n + (Z[ii - 1, kk - 1, jj - 1] - 1)
    val = weights[ii - i0, kk - 1, jj - j0]
    rev = vox[::-1]
    _, first = np.unique(rev, return_index=True)
    last = len(vox) - 1 - first
    block = np.zeros(n ** 3)
    block[vox[last]] = val[last]
    return block.reshape(n, n, n)


def _radial(a, b, n, corner=None):
    """mRadial: a x n x b ones, 0.5 on the first/last row and column, and 1/3 at
    the (row, column) `corner` when given (GetMeyerBasedFilter.m:27-30 etc.)."""
    m = np.ones((a, n, b))
    m[[0, a - 1], :, :] = 0.5
    m[:, :, [0, b - 1]] = 0.5
    if corner is not None:
        m[corner[0], :, corner[1]] = 1.0 / 3.0
    return m


def meyer_window_specs(num_dir, n=24):
    """The (l2, l1) -> (index ranges, weights) table of one cone's windows
    (3DShearTrans/GetMeyerBasedFilter.m:26-137), 1-based l2, l1."""
    s = n // num_dir
    D = num_dir
    specs = {}

    def rng(l, last):
        return ((l - 1) * s + 1, l * s) if last else ((l - 1) * s + 1, l * s + 1)

    for l2 in range(1, D + 1):
        for l1 in range(1, D + 1):
            last1, last2 = l1 == D, l2 == D
            a = s if last1 else s + 1
            b = s if last2 else s + 1
            corner = None
            if l1 == 1 and l2 == 1:
                corner = (0, 0)
            elif l1 == D and l2 == 1:
                corner = (a - 1, 0)
            elif l1 == 1 and l2 == D:
                corner = (0, b - 1)
            elif l1 == D and l2 == D:
                corner = (a - 1, b - 1)
            specs[(l2, l1)] = ((rng(l1, last1), rng(l2, last2)), _radial(a, b, n, corner))
    return specs


def meyer_filters_3d(num_dirs, n=24):
    """Spatial n^3 filters F[c][l] (a D x D nested list indexed [l2-1][l1-1]) for
    cones c = 0, 1, 2 and levels with num_dirs[l] directions per axis
    (3DShearTrans/GetMeyerBasedFilter.m:12-151, the final swap of cones 1 and 2
    included).  A level's filters depend only on its direction count."""
    sections = pyramid_sections(n)
    cache = {}
    F = [[None] * len(num_dirs) for _ in range(3)]
    for l, D in enumerate(num_dirs):
        if D not in cache:
            specs = meyer_window_specs(D, n)
            W = [[[None] * D for _ in range(D)] for _ in range(3)]
            A = np.zeros((n, n, n))
            for c in range(3):
                for (l2, l1), (ranges, wts) in specs.items():
                    W[c][l2 - 1][l1 - 1] = polar_to_rec(ranges, wts, n, sections[c])
                    A += W[c][l2 - 1][l1 - 1]
            if np.any(A == 0):
                raise ValueError("direction count %d leaves frequencies uncovered (F./A divides by zero)" % D)
            out = [[[None] * D for _ in range(D)] for _ in range(3)]
            for c in range(3):
                for a in range(D):
                    for b in range(D):
                        f = W[c][a][b] / A
                        out[c][a][b] = np.real(np.fft.fftshift(np.fft.ifftn(np.fft.fftshift(f))))
            out[0], out[1] = out[1], out[0]
            cache[D] = out
        for c in range(3):
            F[c][l] = cache[D][c]
    return F


def band_list(F, level):
    """Filters of one level in the reference's band order: cone-major, then the
    MATLAB linear (column-major) index over the (l2, l1) cell
    (ShDecThRec_cuda.cpp:128-145,260-291)."""
    out = []
    for c in range(3):
        cell = F[c][level]
        D = len(cell)
        for l1 in range(D):
            for l2 in range(D):
                out.append(cell[l2][l1])
    return out


# --- CONVNFFT: N-D linear convolution through the FFT ---------------------------
def convnfft_same(A, B):
    """'same' linear convolution of A by B: the full convolution sliced at
    ceil((nb-1)/2) on every axis (CONVNFFT_Folder/convnfft.m, ifun for 'same').
    Any FFT length >= na + nb - 1 gives the linear convolution; convnfft pads to
    a power of two, this uses the next fast length."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    full = [next_fast_len(a + b - 1) for a, b in zip(A.shape, B.shape)]
    C = np.fft.irfftn(np.fft.rfftn(A, s=full, axes=(0, 1, 2)) * np.fft.rfftn(B, s=full, axes=(0, 1, 2)), s=full, axes=(0, 1, 2))
    sl = tuple(slice(int(math.ceil((b - 1) / 2)), int(math.ceil((b - 1) / 2)) + a) for a, b in zip(A.shape, B.shape))
    return C[sl]


class _SameConv:
    """convnfft_same of one volume by many filters, sharing the volume's FFT."""

    def __init__(self, A, fshape):
        self.shape = A.shape
        self.fshape = fshape
        self.full = [next_fast_len(a + b - 1) for a, b in zip(A.shape, fshape)]
        self.Ah = np.fft.rfftn(np.asarray(A, dtype=np.float64), s=self.full, axes=(0, 1, 2))
        self.sl = tuple(slice(int(math.ceil((b - 1) / 2)), int(math.ceil((b - 1) / 2)) + a)
                        for a, b in zip(A.shape, fshape))

    def __call__(self, B):
        C = np.fft.irfftn(self.Ah * np.fft.rfftn(B, s=self.full, axes=(0, 1, 2)), s=self.full, axes=(0, 1, 2))
        return C[self.sl]


# --- ShDec / ThresholdShCoeff / ShRec --------------------------------------------
def sh_dec(cone, F, subs, nlevels):
    """Bands of one cone: coeff[l][l2][l1] = convnfft(subs[l], F[cone][l][l2][l1], 'same')
    (3DShearTrans/ShDec.m:18-34)."""
    out = []
    for l in range(nlevels):
        cell = F[cone][l]
        conv = _SameConv(subs[l], cell[0][0].shape)
        out.append([[conv(f) for f in row] for row in cell])
    return out


def threshold_sh_coeff(coeff, band_est, cone, sigma, T):
    """Zero |c| <= T[l] sigma bandEst{cone,l}(l2,l1) (DenoiseDemo/ThresholdShCoeff.m:1-10)."""
    out = []
    for l, cell in enumerate(coeff):
        rows = []
        for l2, row in enumerate(cell):
            r = []
            for l1, c in enumerate(row):
                t = T[l] * sigma * band_est[cone][l][l2][l1]
                r.append(np.where(np.abs(c) <= t, 0.0, c))
            rows.append(r)
        out.append(rows)
    return out


def sh_rec(coeff):
    """Plain sum of a level's bands (3DShearTrans/ShRec.m:11-23)."""
    return [sum(c for row in cell for c in row) for cell in coeff]


def dec_thresh_rec(F, subs, nlevels, band_est, sigma, T):
    """Per level: sum over cones and directions of the hard-thresholded 'same'
    convolutions (ShDecThRec_cuda.cpp:120-171,178-335, = main.m:113-131)."""
    rec = []
    for l in range(nlevels):
        conv = _SameConv(subs[l], F[0][l][0][0].shape)
        acc = np.zeros(subs[l].shape)
        for c in range(3):
            cell = F[c][l]
            for l2, row in enumerate(cell):
                for l1, f in enumerate(row):
                    b = conv(f)
                    t = T[l] * sigma * band_est[c][l][l2][l1]
                    acc += np.where(np.abs(b) <= t, 0.0, b)
        rec.append(acc)
    return rec


def mad_sigma(v):
    """median(|v - median(v)|) / 0.6745 (DenoiseDemo/NsstScalars.m:7,15)."""
    v = np.asarray(v, dtype=np.float64).ravel()
    return float(np.median(np.abs(v - np.median(v))) / 0.6745)


def nsst_scalars_3d(noise, F, nlevels, mode=1):
    """bandEst[c][l][l2][l1] and lowPassEst from a unit-variance noise volume
    (DenoiseDemo/NsstScalars.m:1-19; the noise is passed in)."""
    subs = pyr_nd_dec(noise, nlevels, mode)
    low = mad_sigma(subs[nlevels])
    est = [[None] * nlevels for _ in range(3)]
    for l in range(nlevels):
        conv = _SameConv(subs[l], F[0][l][0][0].shape)
        for c in range(3):
            cell = F[c][l]
            est[c][l] = [[mad_sigma(conv(f)) for f in row] for row in cell]
    return est, low


def denoise_3d(xn, F, nlevels, band_est, sigma, T, mode=1):
    """Band-pass, per-level thresholded shearlet sum, lowpass passed through,
    reconstruction (DenoiseDemo/main.m:84-150)."""
    subs = pyr_nd_dec(xn, nlevels, mode)
    rec = dec_thresh_rec(F, subs, nlevels, band_est, sigma, T)
    rec.append(subs[nlevels])
    return pyr_nd_rec(rec, mode)


# --- linear analysis operator and its transpose (for adjoint checks) -------------
def analysis_3d(x, F, nlevels, mode=1):
    """x -> (bands per level in band_list order, lowpass)."""
    subs = pyr_nd_dec(x, nlevels, mode)
    bands = []
    for l in range(nlevels):
        conv = _SameConv(subs[l], F[0][l][0][0].shape)
        bands.append([conv(f) for f in band_list(F, l)])
    return bands, subs[nlevels]


def analysis_3d_adjoint(bands, low, F, nlevels, mode=1):
    """Transpose of analysis_3d: the 'same' convolution's transpose is the
    correlation with the filter at the same offset, and the band-pass
    synthesis is the transpose of its analysis (real masks with
    Lf^2 + Hf^2 = 1)."""
    subs = []
    for l in range(nlevels):
        fl = band_list(F, l)
        n = bands[l][0].shape
        acc = np.zeros(n)
        for b, f in zip(bands[l], fl):
            full = [next_fast_len(a + s - 1) for a, s in zip(n, f.shape)]
            o = [int(math.ceil((s - 1) / 2)) for s in f.shape]
            G = np.zeros(full)
            G[o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]] = b
            # correlation: sum_k G(k) f(k - m)
            C = np.fft.irfftn(np.fft.rfftn(G, s=full, axes=(0, 1, 2)) * np.conj(np.fft.rfftn(f, s=full, axes=(0, 1, 2))), s=full, axes=(0, 1, 2))
            acc += C[:n[0], :n[1], :n[2]]
        subs.append(acc)
    subs.append(np.asarray(low, dtype=np.float64))
    return pyr_nd_rec(subs, mode)
