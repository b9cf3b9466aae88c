"""Spatial Meyer shearing filters and dictionary parameters (host, FP64).

Host-side mirror of the MATLAB front end that builds the directional filters
before they are handed to the GPU (shearcuda/shearing_filters_Myer.m:29-40,
which calls gen_x_y_cordinates.m, windowing.m, meyer_wind.m, rec_from_pol.m).
The spectra themselves are formed by the C ABI
(shx_prepare_shearing_filters), as prepare_shearing_filters_cuda does.
"""
import numpy as np

# Default parameters of the 2D denoise path
# (shearlet_toolbox_1/denoise_shearlet_example.m:26-38).
DEFAULT_DCOMP = (3, 3, 4, 4)
DEFAULT_DSIZE = (32, 32, 16, 16)
DEFAULT_TSCALARS = (0.0, 3.0, 4.0)


def meyer_window(x):
    """Vectorised Meyer window (meyer_wind.m:11-19)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    inner = (x > -1.0 / 3 + 0.5) & (x < 1.0 / 3 + 0.5)
    ramp = ~inner & (((x >= 1.0 / 3 + 0.5) & (x <= 2.0 / 3 + 0.5)) | ((x >= -2.0 / 3 + 0.5) & (x <= 1.0 / 3 + 0.5)))
    w = 3.0 * np.abs(x - 0.5) - 1.0
    z = w ** 4 * (35.0 - 84.0 * w + 70.0 * w ** 2 - 20.0 * w ** 3)
    out[inner] = 1.0
    out[ramp] = np.cos(np.pi / 2.0 * z[ramp]) ** 2
    return out


def meyer_bands(n, nbands):
    """n x nbands matrix of Meyer band windows on a length-n unit signal
    (windowing.m:11-27 applied to ones(n,1))."""
    T = n // nbands
    g = meyer_window((np.arange(2 * T) - T / 2.0) / T)
    y = np.zeros((n, nbands))
    ks = np.arange(-T / 2.0, 1.5 * T - 1 + 0.5)  # k = -T/2 : 1.5T-1
    for j in range(nbands):
        idx = np.floor(np.mod(ks + j * T, n)).astype(int)
        # later writes win, as in the MATLAB loop
        for i, g_i in zip(idx, g[: len(idx)]):
            y[i, j] = g_i
    return y


def _round_half_away(v):
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def pseudo_polar_grid(n):
    """1-based grids (x1, y1, x2, y2) and the hit counts D of the n x n
    pseudo-polar sampling (gen_x_y_cordinates.m:22-111)."""
    m = n + 1
    i = np.arange(1, m + 1)[:, None].astype(np.float64)      # slice index
    j = np.arange(m)[None, :].astype(np.float64)              # 0..m-1
    x0 = i
    xn = m - i + 1
    # linspace(x0, xn, m) as MATLAB forms it: x0 + (0:m-2)*(xn-x0)/(m-1), last = xn
    xt = x0 + (j * (xn - x0)) / (m - 1)
    xt[:, -1] = xn[:, 0]
    vertical = (xn == x0)[:, 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        slope = (m - 1) / (xn - x0)
        ya = _round_half_away(slope * (xt - x0) + 1)
    xa = _round_half_away(xt)
    mid = (m - 1) / 2 + 1
    ya[vertical] = np.arange(1, m + 1)
    xa[vertical] = mid
    x1, y1 = xa, ya
    x2 = ya.copy()
    y2 = xa.copy()
    x2[vertical] = np.arange(1, m + 1)
    y2[vertical] = mid
    x1n = np.flipud(x1[:n, :n]).astype(int)
    y1n = y1[:n, :n].astype(int)
    x2n = x2[1:n + 1, :n].astype(int)
    y2n = y2[1:n + 1, :n].astype(int)
    y2n[n - 1, 0] = n
    D = np.zeros((n, n))
    np.add.at(D, (y1n - 1, x1n - 1), 1)
    np.add.at(D, (y2n - 1, x2n - 1), 1)
    return x1n, y1n, x2n, y2n, D


def shearing_filters_myer(n1, level):
    """(2**level, n1, n1) FP64 spatial shearing filters of one scale
    (shearcuda/shearing_filters_Myer.m:29-40)."""
    x1, y1, x2, y2, D = pseudo_polar_grid(n1)
    wf = meyer_bands(2 * n1, 2 ** level)
    out = np.empty((2 ** level, n1, n1))
    for k in range(2 ** level):
        C = np.zeros((n1, n1))
        # rec_from_pol.m:26-34: accumulate the two radial halves, average by D
        np.add.at(C, (y1 - 1, x1 - 1), wf[:n1, k][:, None] * np.ones((1, n1)))
        np.add.at(C, (y2 - 1, x2 - 1), wf[n1:, k][:, None] * np.ones((1, n1)))
        C /= D
        out[k] = np.real(np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(C)))) / np.sqrt(n1)
    return out


def dictionary_filters(dcomp=DEFAULT_DCOMP, dsize=DEFAULT_DSIZE):
    """Spatial filters of every scale, coarse to fine."""
    if len(dcomp) != len(dsize):
        raise ValueError("dcomp and dsize must have the same length")
    return [shearing_filters_myer(int(n1), int(lev)) for lev, n1 in zip(dcomp, dsize)]
