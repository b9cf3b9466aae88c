"""Pins of the FP64 3D oracle (oracle/nsst3d.py) and of the product-side 3D
filter builder against it.  CPU only."""
import math
import os

import numpy as np
import pytest
import scipy.signal

from oracle import nsst3d as o3
from shearcuda_amd import filters3d

HERE = os.path.dirname(os.path.abspath(__file__))
SCALARS = os.path.join(HERE, "golden", "nsst3d_scalars_3884444.npz")


def test_rcos_closed_form():
    x = np.array([-0.5, 0.0, 0.25, 0.5, 0.75, 1.0, 1.5])
    want = [0.0, 0.0, (1 - math.sqrt(0.5)) / 2, 0.5, (1 + math.sqrt(0.5)) / 2, 1.0, 1.0]
    assert np.allclose(o3.rcos(x), want, atol=1e-15)


@pytest.mark.parametrize("mode,n", [(1, 32), (1.5, 48), (2, 64)])
def test_bandpass_perfect_reconstruction(mode, n):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    subs = o3.pyr_nd_dec(x, 3, mode)
    assert [s.shape[0] for s in subs] == o3.level_sizes(n, 3, mode)
    assert np.abs(o3.pyr_nd_rec(subs, mode) - x).max() < 1e-12


def test_bandpass_rejects_non_divisible_size():
    with pytest.raises(ValueError, match="downsampling factor"):
        o3.pyr_nd_dec(np.zeros((32, 32, 32)), 2, 1.5)


@pytest.mark.parametrize("nr,w,tbw", [(128, 0.5, 1 / 6), (128, 0.25, 1 / 12), (64, 1 / 3, 1 / 7), (96, 0.5, 1 / 7)])
def test_passband_edges(nr, w, tbw):
    P, v = o3.passband(nr, w, tbw)
    assert P == math.ceil(nr / 2 * (w + tbw))
    assert v[P] == 0.0  # Lf vanishes on the edge of its box
    assert v[0] == 1.0  # and is one inside the ideal passband
    assert np.all(np.diff(v) <= 0)


def test_bandpass_masks_sum_of_squares():
    # Hf^2 + Lf^2 = 1 as applied: Hp / X and Lp / X on the box
    n = 32
    X = np.fft.rfftn(np.random.default_rng(0).standard_normal((n, n, n)))
    w, tbw, D = o3.pyr_params(1, 1)
    Lp, Hp = o3.pyr_dec_onestep(X, w[0], tbw[0], D[0])
    ok = np.abs(X) > 1e-9
    r = (np.abs(Hp[ok]) ** 2 + np.abs(Lp[ok]) ** 2) / np.abs(X[ok]) ** 2
    assert np.allclose(r, 1.0, atol=1e-12)


def test_polar_grid_closed_form():
    n = 24
    X, Y, Z = o3.generate_xyz(n)
    i = np.arange(1, n + 1)[:, None, None]
    k = np.arange(1, n + 1)[None, :, None]
    j = np.arange(1, n + 1)[None, None, :]
    assert np.array_equal(X, np.broadcast_to(k, X.shape))
    assert np.array_equal(Y, np.broadcast_to(np.floor(i + (k - 1) * (n + 1 - 2 * i) / (n - 1) + 0.5), Y.shape))
    assert np.array_equal(Z, np.broadcast_to(np.floor(j + (k - 1) * (n + 1 - 2 * j) / (n - 1) + 0.5), Z.shape))
    # endpoints of polar line (i, j): (1, i, j) -> (n, n+1-i, n+1-j)
    assert np.array_equal(Y[:, 0, 0], np.arange(1, n + 1)) and np.array_equal(Y[:, -1, 0], np.arange(n, 0, -1))


@pytest.fixture(scope="module")
def filters():
    return o3.meyer_filters_3d([8, 8, 4])


@pytest.mark.parametrize("level", [0, 2])
def test_filters_partition_of_unity(filters, level):
    S = sum(f for c in range(3) for row in filters[c][level] for f in row)
    d = np.zeros((24, 24, 24))
    d[12, 12, 12] = 1.0
    assert np.abs(S - d).max() < 1e-14


def test_filters_cone_orientation(filters):
    # after the final swap, cone c has its radial axis along axis (1, 0, 2)[c]:
    # its windows' energy lies in the pyramid |f_ax| >= |f_other| (zero
    # frequency at the fftshift centre 12); windows on a cone's edge share
    # about a fifth with the neighbouring cone
    n = 24
    f = np.abs(np.arange(n) - 12)
    grid = np.meshgrid(f, f, f, indexing="ij")
    for c, ax in enumerate((1, 0, 2)):
        for wrong in (False, True):
            a = (ax + 1) % 3 if wrong else ax
            others = [grid[b] for b in range(3) if b != a]
            inside = (grid[a] >= others[0]) & (grid[a] >= others[1])
            frac = []
            for row in filters[c][0]:
                for flt in row:
                    W = np.abs(np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(flt)))) ** 2
                    frac.append(W[inside].sum() / W.sum())
            if wrong:
                assert np.mean(frac) < 0.2
            else:
                assert min(frac) > 0.75 and np.mean(frac) > 0.95


def test_filters_direction_count_not_dividing_fails():
    with pytest.raises(ValueError, match="uncovered"):
        o3.meyer_filters_3d([7])


@pytest.mark.parametrize("n", [4, 6])
def test_convnfft_same_matches_direct_convolution(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n + 1, n + 2))
    B = rng.standard_normal((24, 24, 24))
    full = scipy.signal.convolve(A, B, mode="full", method="direct")
    want = full[12:12 + n, 12:12 + n + 1, 12:12 + n + 2]
    assert np.abs(o3.convnfft_same(A, B) - want).max() < 1e-10


def test_convnfft_same_matches_fftconvolve_odd_filter():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((20, 20, 20))
    B = rng.standard_normal((5, 5, 5))
    want = scipy.signal.fftconvolve(A, B, mode="full")[2:22, 2:22, 2:22]
    assert np.abs(o3.convnfft_same(A, B) - want).max() < 1e-11


def test_threshold_closed_form():
    coeff = [[[np.array([[[-3.0, 2.0, 2.5, -0.1]]])]]]
    out = o3.threshold_sh_coeff(coeff, [[[[0.5]]]], 0, 2.0, [2.5])
    assert np.array_equal(out[0][0][0], np.array([[[-3.0, 0.0, 0.0, 0.0]]]))


def test_product_filters_match_oracle(filters):
    prod = filters3d.get_meyer_based_filter([8, 8, 4])
    assert prod[0] is prod[1]
    for level in (0, 2):
        want = np.stack(o3.band_list(filters, level))
        assert prod[level].shape == want.shape
        assert np.abs(prod[level] - want).max() < 1e-15


def test_analysis_adjoint_identity():
    n = 16
    F = o3.meyer_filters_3d([2, 2])
    rng = np.random.default_rng(9)
    x = rng.standard_normal((n, n, n))
    bands, low = o3.analysis_3d(x, F, 2)
    yb = [[rng.standard_normal(b.shape) for b in lvl] for lvl in bands]
    yl = rng.standard_normal(low.shape)
    lhs = sum(np.vdot(b, y) for lb, ly in zip(bands, yb) for b, y in zip(lb, ly)) + np.vdot(low, yl)
    rhs = np.vdot(x, o3.analysis_3d_adjoint(yb, yl, F, 2))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_noise_scalars_match_reference_fixture():
    """The reference's Monte-Carlo noise scalars (DenoiseDemo/nsstScalarsData3884444.mat,
    converted by tests/golden/make_3d_scalar_fixture.py: Pyr_mode 1.5, 8x8 / 4x4 /
    4x4 directions, data size unknown) against this restatement at 96^3: the
    per-band pattern correlates > 0.99 at the finest level and the values agree
    to a few percent (MC noise and the boundary share of the unknown size)."""
    fx = np.load(SCALARS)
    lp = float(fx["low_pass_est"])
    F = o3.meyer_filters_3d([8, 4, 4])
    noise = np.random.default_rng(1).standard_normal((96, 96, 96))
    est, low = o3.nsst_scalars_3d(noise, F, 3, 1.5)
    assert abs(low / lp - 1) < 0.02
    for c in range(3):
        for l in range(3):
            a = np.array(est[c][l])
            b = fx["band_est_c%d_l%d" % (c, l)].astype(np.float64)
            r = a / b
            assert abs(r.mean() - 1) < 0.07
            if l == 0:
                assert np.corrcoef(a.ravel(), b.ravel())[0, 1] > 0.99
