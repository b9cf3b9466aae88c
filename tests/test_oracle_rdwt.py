"""Pins of the RDWT oracle (CPU)."""
import numpy as np
import pytest

from oracle import rdwt as w


@pytest.mark.parametrize("name,par", [("Haar", 0), ("Daubechies", 4), ("Symmlet", 4)])
def test_filters_orthonormal(name, par):
    h = w.make_on_filter(name, par)
    assert abs(np.sum(h * h) - 1) < 1e-14
    assert abs(np.sum(h) - np.sqrt(2)) < 1e-9          # lowpass: H(0) = sqrt(2)
    # double-shift orthogonality: sum_k h[k] h[k+2m] = 0, m != 0
    for m in range(1, len(h) // 2):
        assert abs(np.sum(h[:-2 * m] * h[2 * m:])) < 1e-9


@pytest.mark.parametrize("name,par,L,shape", [("Haar", 0, 1, (8, 8)), ("Daubechies", 4, 3, (32, 64)),
                                              ("Symmlet", 4, 2, (16, 16))])
def test_perfect_reconstruction(name, par, L, shape):
    h = w.make_on_filter(name, par)
    x = np.random.default_rng(0).standard_normal(shape)
    yl, yh = w.mrdwt(x, h, L)
    assert yh.shape == (shape[0], 3 * shape[1] * L)
    # the reference's taps carry 12 digits, so orthogonality (and PR) hold to ~1e-12
    assert np.abs(w.mirdwt(yl, yh, h, L) - x).max() < 1e-10


def test_haar_one_level_closed_form():
    h = w.make_on_filter("Haar")
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6))
    yl, yh = w.mrdwt(x, h, 1)
    # h0 = [1 1]/sqrt2, h1 = [-1 1]/sqrt2 applied as out[i] = x[i]*f[1] + x[i+1]*f[0]
    nxt = lambda a, ax: np.roll(a, -1, axis=ax)
    lo_r = (x + nxt(x, 1)) / np.sqrt(2)
    hi_r = (x - nxt(x, 1)) / np.sqrt(2)
    assert np.allclose(yl, (lo_r + nxt(lo_r, 0)) / np.sqrt(2), atol=1e-14)
    assert np.allclose(yh[:, 0:6], (lo_r - nxt(lo_r, 0)) / np.sqrt(2), atol=1e-14)
    assert np.allclose(yh[:, 6:12], (hi_r + nxt(hi_r, 0)) / np.sqrt(2), atol=1e-14)
    assert np.allclose(yh[:, 12:18], (hi_r - nxt(hi_r, 0)) / np.sqrt(2), atol=1e-14)


def test_product_filters_match_oracle():
    from shearcuda_amd.wavelets import MakeONFilter
    for name, par in [("Haar", 0), ("Daubechies", 4), ("Symmlet", 4)]:
        assert np.abs(MakeONFilter(name, par) - w.make_on_filter(name, par)).max() < 1e-15


@pytest.mark.parametrize("name,par", [("Haar", 0), ("Daubechies", 4), ("Symmlet", 4)])
def test_mdwt1_orthonormal(name, par):
    """The one-level decimated transform of an orthonormal filter preserves
    energy (a wrong phase, index or sign breaks it)."""
    x = np.random.default_rng(3).standard_normal((32, 16))
    wc = w.mdwt1(x, w.make_on_filter(name, par))
    assert wc.shape == x.shape
    assert abs(np.sum(wc * wc) / np.sum(x * x) - 1) < 1e-10


def test_mdwt1_haar_closed_form():
    x = np.random.default_rng(4).standard_normal((8, 8))
    wc = w.mdwt1(x, w.make_on_filter("Haar"))
    a, b, c, d = x[0::2, 0::2], x[0::2, 1::2], x[1::2, 0::2], x[1::2, 1::2]
    assert np.allclose(wc[:4, :4], (a + b + c + d) / 2, atol=1e-15)
    assert np.allclose(wc[4:, 4:], (a - b - c + d) / 2, atol=1e-15)


def test_mdwt1_is_mrdwt_at_even_samples():
    """The decimated finest diagonal band is the undecimated one at even rows
    and columns (what the GPU separation uses for its noise level)."""
    x = np.random.default_rng(5).standard_normal((16, 16))
    h = w.make_on_filter("Daubechies", 4)
    _, yh = w.mrdwt(x, h, 1)
    assert np.array_equal(w.mdwt1(x, h)[8:, 8:], yh[0::2, 32::2])
