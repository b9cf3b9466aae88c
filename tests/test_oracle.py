"""Pins of the FP64 oracle (CPU).  Each test checks the restatement against
something that fixes it independently: a closed form, an identity of the
filter design, a library routine, or brute force on tiny inputs."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from oracle import nsst as o


def test_maxflat_perfect_reconstruction():
    # atrousfilters.m:16-18: conv2(h0,g0) + conv2(h1,g1) = delta
    h0, h1, g0, g1 = o.atrousfilters()
    p = convolve2d(h0, g0) + convolve2d(h1, g1)
    d = np.zeros_like(p)
    d[p.shape[0] // 2, p.shape[1] // 2] = 1.0
    assert np.abs(p - d).max() < 1e-14
    assert h0.shape == (13, 13) and g1.shape == (13, 13)
    assert h1.shape == (19, 19) and g0.shape == (19, 19)


def test_maxflat_symmetry_and_dc():
    for h in o.atrousfilters():
        assert np.array_equal(h, h.T)
        assert np.array_equal(h, h[::-1, :])
        assert np.array_equal(h, h[:, ::-1])
    h0, h1, g0, g1 = o.atrousfilters()
    assert abs(h0.sum() - 1) < 1e-14 and abs(g0.sum() - 1) < 1e-14
    assert abs(h1.sum()) < 1e-14 and abs(g1.sum()) < 1e-14


def test_maxflat_tampered_table_breaks_pr():
    # a single wrong digit in one table must break the identity the test uses
    h0, h1, g0, g1 = o.atrousfilters()
    h0 = h0.copy()
    h0[6, 6] += 1e-9
    p = convolve2d(h0, g0) + convolve2d(h1, g1)
    d = np.zeros_like(p)
    d[p.shape[0] // 2, p.shape[1] // 2] = 1.0
    assert np.abs(p - d).max() > 1e-10


def _symext_bruteforce(x, p, shift):
    # symext.m:29-31 by explicit index lists (1-based MATLAB indices)
    m, n = x.shape
    ss = p // 2 - shift + 1
    cols = list(range(ss, 0, -1)) + list(range(1, n + 1)) + list(range(n, n - p - shift, -1))
    rows = list(range(ss, 0, -1)) + list(range(1, m + 1)) + list(range(m, m - p - shift, -1))
    y = np.array([[x[r - 1, c - 1] for c in cols] for r in rows])
    return y[: m + p - 1, : n + p - 1]


@pytest.mark.parametrize("n,p,shift", [(16, 13, 1), (24, 19, 1), (40, 26, 1), (64, 38, 1), (80, 76, 0), (160, 152, -2)])
def test_symext_matches_bruteforce(n, p, shift):
    x = np.random.default_rng(n + p).standard_normal((n, n))
    ref = _symext_bruteforce(x, p, shift)
    assert np.array_equal(o.symext(x, p, p, (shift, shift), strict=True), ref)
    assert np.array_equal(o.symext(x, p, p, (shift, shift), strict=False), ref)


def test_symext_strict_raises_like_matlab():
    x = np.zeros((32, 32))
    with pytest.raises(IndexError):
        o.symext(x, 152, 152, (-2, -2), strict=True)
    # the folded extension is still defined
    assert o.symext(x, 152, 152, (-2, -2), strict=False).shape == (183, 183)


def test_fold_is_half_sample_reflection():
    n = 5
    got = o._fold(np.arange(-7, 13), n)
    fwd, rev = np.arange(n), np.arange(n)[::-1]
    ext = np.concatenate([fwd, rev, fwd, rev, fwd])
    # index s -> ext[s + 2n] for s in [-2n, 3n)
    assert np.array_equal(got, ext[np.arange(-7, 13) + 2 * n])


@pytest.mark.parametrize("power", [0, 1, 2, 3])
def test_atrousc_is_valid_convolution_with_upsampled_filter(power):
    rng = np.random.default_rng(power)
    h = rng.standard_normal((5, 7))
    M = 2 ** power
    hu = o.upsample2df(h, power)
    xe = rng.standard_normal((hu.shape[0] + 9, hu.shape[1] + 6))
    ref = convolve2d(xe, hu, mode="valid")
    assert np.abs(o.atrousc(xe, h, M) - ref).max() < 1e-12
    assert np.abs(o.conv2_valid(xe, hu) - ref).max() < 1e-12


def test_atrous_pyramid_round_trip():
    x = np.random.default_rng(1).standard_normal((256, 256)) * 40
    y = o.atrousdec(x, "maxflat", 4)
    assert len(y) == 5
    xr = o.atrousrec(y)
    assert np.abs(xr - x).max() / np.abs(x).max() < 1e-13


def test_atrous_pyramid_fold_equals_strict_when_defined():
    x = np.random.default_rng(2).standard_normal((256, 256))
    a = o.atrousdec(x, "maxflat", 4, strict=True)
    b = o.atrousdec(x, "maxflat", 4, strict=False)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_atrous_small_image_needs_fold():
    x = np.random.default_rng(3).standard_normal((64, 64))
    with pytest.raises(IndexError):
        o.atrousdec(x, "maxflat", 4, strict=True)
    y = o.atrousdec(x, "maxflat", 4, strict=False)
    assert np.abs(o.atrousrec(y, strict=False) - x).max() < 1e-12


def test_atrous_level_matches_centred_formula():
    # one level with dilation m equals sum h(a,b) x(fold(r+m(a-c0)), fold(c+m(b-c0)))
    rng = np.random.default_rng(4)
    x = rng.standard_normal((32, 32))
    h0, h1, _, _ = o.atrousfilters()
    for m, h in [(1, h0), (2, h1), (4, h0)]:
        p = h.shape[0] * m if m > 1 else h.shape[0]
        s = 1 if m == 1 else -(m // 2) + 2
        xe = o.symext(x, p, p, (s, s), strict=False)
        got = o.atrousc(xe, h, m) if m > 1 else o.conv2_valid(xe, h)
        c0 = h.shape[0] // 2
        ref = np.zeros_like(x)
        r, c = np.mgrid[0:32, 0:32]
        for a in range(h.shape[0]):
            for b in range(h.shape[1]):
                ref += h[a, b] * x[o._fold(r + m * (a - c0), 32), o._fold(c + m * (b - c0), 32)]
        assert np.abs(got - ref).max() < 1e-13


def test_meyer_window_closed_form():
    assert o.meyer_wind(0.5) == 1.0
    assert o.meyer_wind(5 / 6 + 1e-12) == pytest.approx(1.0, abs=1e-9)
    assert o.meyer_wind(7 / 6) == pytest.approx(0.0, abs=1e-15)
    assert o.meyer_wind(1.0) == pytest.approx(0.5, abs=1e-15)   # w=1/2 -> z=1/2
    assert o.meyer_wind(1.3) == 0.0 and o.meyer_wind(-0.3) == 0.0


@pytest.mark.parametrize("N,L", [(64, 8), (32, 16), (64, 16), (32, 4)])
def test_windowing_partition_of_unity(N, L):
    y = o.windowing(np.ones(N), L)
    assert np.abs(y.sum(axis=1) - 1.0).max() < 1e-15


@pytest.mark.parametrize("n1,level", [(32, 3), (16, 4), (16, 2), (8, 2)])
def test_shearing_filters_sum_to_centred_delta(n1, level):
    # the windows partition unity, so sum_k w_k = fftshift(ifft2(1))/sqrt(n1)
    w = o.shearing_filters_myer(n1, level)
    d = np.zeros((n1, n1))
    d[n1 // 2, n1 // 2] = 1.0 / np.sqrt(n1)
    assert np.abs(w.sum(axis=0) - d).max() < 1e-14


def test_pseudo_polar_grid_counts():
    x1, y1, x2, y2, D = o.gen_x_y_cordinates(32)
    assert D.sum() == 2 * 32 * 32
    assert D.min() >= 1
    for a in (x1, y1, x2, y2):
        assert a.min() >= 1 and a.max() <= 32


def test_meyer_filters_golden():
    g = np.load(__file__.replace("test_oracle.py", "golden/meyer_filters.npz"))
    for key in g.files:
        n1, lev = (int(v) for v in key.split("_")[1:])
        assert np.abs(o.shearing_filters_myer(n1, lev) - g[key]).max() < 1e-15


@pytest.fixture(scope="module")
def dict256():
    return o.make_dictionary(256)


def test_spectra_tight_frame(dict256):
    for H in dict256:
        assert np.abs(np.sum(np.abs(H) ** 2, axis=0) - 1.0).max() < 1e-13
        # Hermitian symmetry of real filters: H(-w) = conj(H(w))
        Hm = np.roll(H[:, ::-1, ::-1], 1, axis=(1, 2))
        assert np.abs(Hm - np.conj(H)).max() < 1e-13


def test_nsst_round_trip_and_adjoint(dict256):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((256, 256)) * 30
    d = o.nsst_dec2(x, dict256)
    assert [a.shape for a in d] == [(256, 256), (8, 256, 256), (8, 256, 256), (16, 256, 256), (16, 256, 256)]
    assert np.abs(o.nsst_rec2(d, dict256) - x).max() / np.abs(x).max() < 1e-13
    # directional stage: <F y, c> = <y, F^T c> with F^T the inverse stage
    y = rng.standard_normal((256, 256))
    c = rng.standard_normal((8, 256, 256))
    H = dict256[0]
    Fy = np.real(np.fft.ifft2(H * np.fft.fft2(y)[None]))
    Ftc = np.real(np.sum(np.fft.ifft2(np.conj(H) * np.fft.fft2(c)), axis=0))
    assert abs(np.sum(Fy * c) - np.sum(y * Ftc)) < 1e-9 * np.abs(np.sum(Fy * c))


def test_conv2p_is_periodic_convolution():
    rng = np.random.default_rng(6)
    y = rng.standard_normal((16, 16))
    w = rng.standard_normal((4, 4))
    pad = np.zeros((16, 16))
    pad[:4, :4] = w
    ref = np.zeros((16, 16))
    for a in range(4):
        for b in range(4):
            ref += w[a, b] * np.roll(y, (a, b), axis=(0, 1))
    assert np.abs(o.conv2p(np.fft.fft2(pad), y) - ref).max() < 1e-12


def test_center_shift_is_roll(dict256):
    x = np.random.default_rng(7).standard_normal((256, 256))
    d0 = o.nsst_dec2(x, dict256)
    d1 = o.nsst_dec2(x, o.make_dictionary(256, shift="center"))
    for j, n1 in enumerate([32, 32, 16, 16]):
        assert np.abs(d1[j + 1] - np.roll(d0[j + 1], (-(n1 // 2), -(n1 // 2)), axis=(1, 2))).max() < 1e-12


def test_nsst_HT_closed_form():
    dst = [np.array([[0.5, -2.0]]), np.array([[[1.0, -3.0]], [[2.5, 0.1]]]), np.array([[[4.0, -4.5]]])]
    sc = [1.0, np.array([1.0, 2.0]), np.array([1.0])]
    out = o.nsst_HT(dst, 1.0, [1.0, 1.5, 4.2], sc)
    assert np.array_equal(out[0], [[0.0, -2.0]])             # > 1
    assert np.array_equal(out[1][0], [[0.0, -3.0]])          # > 1.5
    assert np.array_equal(out[1][1], [[0.0, 0.0]])           # > 3.0
    assert np.array_equal(out[2][0], [[0.0, -4.5]])          # > 4.2 (finest)
    # strictness: equal to the threshold is zeroed
    assert o.nsst_HT([np.array([[1.0]])], 1.0, [1.0, 1, 1], [1.0])[0][0, 0] == 0.0


def test_thresh_closed_form():
    Ct = [np.array([[1.0, -5.0]]), np.array([[[3.0, 0.5]], [[-7.0, 2.0]]])]
    E = np.array([[1.0, 0.0], [2.0, 0.5]])
    out, maxi = o.thresh(Ct, 2.0, 2, E, [1.0, 1.0], 1)
    assert np.array_equal(out[0], [[0.0, -5.0]])             # th = 2
    assert np.array_equal(out[1][0], [[0.0, 0.0]])           # th = 4
    assert np.array_equal(out[1][1], [[-7.0, 2.0]])          # th = 1
    assert maxi == pytest.approx(14.0)                       # 7 / 0.5
    out, maxi = o.thresh(Ct, 2.5, 1, E, [1.0, 1.0], 1)
    assert np.array_equal(out[1], [[[3.0, 0.0]], [[-7.0, 0.0]]])
    assert maxi == 7.0


def test_soft_threshold_closed_form():
    assert np.array_equal(o.soft_threshold(np.array([-3.0, -1.0, 0.5, 2.0]), 1.0), [-2.0, 0.0, 0.0, 1.0])


def test_re_weight_band_linear_and_identity():
    x = np.random.default_rng(8).standard_normal((256, 256))
    assert np.abs(o.re_weight_band(x, [1, 1, 1, 1]) - x).max() < 1e-12
    a = o.re_weight_band(x, [1, 0, 0, 0])
    b = o.re_weight_band(x, [0, 1, 1, 1])
    assert np.abs(a + b - x).max() < 1e-12


def test_com_norm_equals_filter_energy(dict256):
    E = o.com_norm(dict256, 256)
    assert E.shape == (5, 16)
    # band (s,w) is the impulse response scaled by L; its RMS is the l2 norm of
    # the band's equivalent filter, i.e. sqrt(sum |G|^2)/L over frequencies
    h0, h1, _, _ = o.atrousfilters()
    y = o.atrousdec(np.pad(np.ones((1, 1)) * 256, ((128, 127), (128, 127))), "maxflat", 4)
    for j in range(1, 5):
        H = dict256[j - 1]
        for k in range(H.shape[0]):
            G = H[k] * np.fft.fft2(y[j])
            assert E[j, k] == pytest.approx(np.sqrt(np.sum(np.abs(G) ** 2)) / 256 / 256, rel=1e-10)
