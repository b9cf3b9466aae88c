"""Buffer validation of the Python mirror: a tensor of the wrong shape, a
threshold or scalar table of the wrong length is rejected with ValueError
before anything reaches the C ABI (which trusts the sizes it is given).  The
library handle is replaced by one that fails on any use, so a check that came
after a launch would show up as an AssertionError instead."""
import ctypes

import numpy as np
import pytest
import torch

from shearcuda_amd import api
from shearcuda_amd import separation as sep
from shearcuda_amd._lib import lib as real_lib

pytestmark = pytest.mark.gpu

L = 64


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError("the library was called (%s) before the arguments were checked" % name)


@pytest.fixture(scope="module")
def shear(gpu):
    d = api.make_dictionary(L)
    yield d
    d.close()


@pytest.fixture(scope="module")
def batch(shear):
    b = api.Batch(shear, 2)
    yield b
    b.close()


@pytest.fixture
def nolib(monkeypatch):
    monkeypatch.setattr(api, "lib", _NoLib())
    monkeypatch.setattr(sep, "lib", _NoLib())


def _img(*shape):
    return torch.zeros(shape, device="cuda", dtype=torch.float32)


def _scalars(shear):
    return [1.0] + [np.ones(d) for d in shear.ndir]


def _cells(shear, bad=None):
    C = shear.alloc_coeffs()
    if bad is not None:
        j, shape = bad
        C[j] = _img(*shape)
    return C


def test_denoise_rejects_bad_buffers(shear, nolib):
    x, sc = _img(L, L), _scalars(shear)
    with pytest.raises(ValueError):
        api.denoise_cuda(_img(L, L + 1), shear, 20.0, sc)
    with pytest.raises(ValueError):
        api.denoise_cuda(_img(2 * L, 2 * L), shear, 20.0, sc)
    with pytest.raises(ValueError):
        api.denoise_cuda(x, shear, 20.0, sc[:-1])                      # one scale missing
    with pytest.raises(ValueError):
        api.denoise_cuda(x, shear, 20.0, [1.0, np.ones(3)] + sc[2:])   # wrong direction count
    with pytest.raises(ValueError):
        api.denoise_cuda(x, shear, 20.0, sc, Tscalars=(0.0, 3.0))
    with pytest.raises(ValueError):
        api.denoise_cuda(x, shear, 20.0, sc, out=_img(L, L - 1))
    with pytest.raises(ValueError):
        api.denoise_cuda(x, shear, 20.0, sc, Ct=_cells(shear, (2, (shear.ndir[1] - 1, L, L))))


def test_transforms_reject_bad_cells(shear, nolib):
    x = _img(L, L)
    bad_planes = _cells(shear, (3, (shear.ndir[2] + 1, L, L)))
    bad_side = _cells(shear, (1, (shear.ndir[0], L, L // 2)))
    bad_low = _cells(shear, (0, (L // 2, L // 2)))
    for C in (bad_planes, bad_side, bad_low):
        with pytest.raises(ValueError):
            api.shear_trans_cuda(x, "maxflat", shear, C)
        with pytest.raises(ValueError):
            api.inverse_shear_cuda(C, "maxflat", shear)
        with pytest.raises(ValueError):
            api.inverse_shear_thresh_cuda(C, np.zeros(shear.nbands), shear)
        with pytest.raises(ValueError):
            api.shear_trans_adjoint_cuda(C, shear)
        with pytest.raises(ValueError):
            api.inverse_shear_adjoint_cuda(x, shear, C)
        with pytest.raises(ValueError):
            api.applyfilters_cuda(C, shear)
    with pytest.raises(ValueError):
        api.inverse_shear_thresh_cuda(shear.alloc_coeffs(), np.zeros(shear.nbands - 1), shear)
    with pytest.raises(ValueError):
        api.inverse_shear_adjoint_cuda(_img(L, L + 1), shear)
    with pytest.raises(ValueError):
        api.inverse_shear_cuda(shear.alloc_coeffs(), "maxflat", shear, out=_img(L + 1, L))
    with pytest.raises(ValueError):
        api.convolution_cuda([_img(L, L)] * shear.nscales + [_img(L, L - 1)], shear)


def test_batch_rejects_bad_buffers(shear, batch, nolib):
    n, nb = batch.nimg, shear.nbands
    x, out = _img(n, L, L), _img(n, L, L)
    coeff = _img(n, nb, L, L)
    th = _img(n, nb)
    with pytest.raises(ValueError):
        batch.denoise(_img(n + 1, L, L), None, th, out)
    with pytest.raises(ValueError):
        batch.denoise(x, None, th, _img(n, L, L // 2))
    with pytest.raises(ValueError):
        batch.denoise(x, _img(n, nb - 1, L, L), th, out)
    with pytest.raises(ValueError):
        batch.denoise(x, None, _img(n, nb - 1), out)
    with pytest.raises(ValueError):
        batch.inverse_shear(coeff, out, th=_img(n - 1, nb))
    with pytest.raises(ValueError):
        batch.inverse_shear(_img(n, nb, L, L + 2), out)
    with pytest.raises(ValueError):
        batch.shear_trans(x, _img(n, nb, L // 2, L))
    with pytest.raises(ValueError):
        batch.shear_trans_adjoint(coeff, _img(n, L, L - 1))
    with pytest.raises(ValueError):
        batch.inverse_shear_adjoint(_img(n, L - 1, L), coeff)


def test_helpers_reject_bad_buffers(shear, nolib):
    with pytest.raises(ValueError):
        api.atrousrec_cuda([_img(L, L), _img(L, L), _img(L, L // 2)], "maxflat")
    with pytest.raises(ValueError):
        api.atrousdec_cuda(_img(L, L // 2), "maxflat", 2)
    with pytest.raises(ValueError):
        api.mirdwt_cuda(_img(L, L), _img(L, 3 * L * 2 - 1), np.ones(4), 2)
    with pytest.raises(ValueError):
        api.mrdwt_adjoint_cuda(_img(L, L), _img(L, 3 * L), np.ones(4), 2)
    x = _img(2, 4, L, L)
    with pytest.raises(ValueError):   # thresholds for 3 of 4 bands
        sep.soft_bands(x, _img(3), 2, 4 * L * L, 4 * L * L, L * L)
    with pytest.raises(ValueError):   # layout larger than x
        sep.hard_bands(x, _img(4), 3, 4 * L * L, 4 * L * L, L * L)
    with pytest.raises(ValueError):
        sep.soft_bands_backward(x, _img(2, 4, L, L - 1), _img(4), _img(4), 2, 4 * L * L, 4 * L * L, L * L)
    with pytest.raises(ValueError):   # E needs S+1 rows
        sep.separate_native(_img(L, L), 2, 3, 3.0, 3.0, [0.1, 0.1, 1.5], shear,
                            E=np.ones((shear.nscales, max(shear.ndir))))


def test_separate_rejects_wrong_E_rows_in_the_abi(shear):
    """The C entry point itself checks E's row count against the dictionary."""
    img = _img(L, L)
    C, P = _img(L, L), _img(L, L)
    dp = ctypes.POINTER(ctypes.c_double)
    coeff = np.array([0.1, 0.1, 1.5])
    qmf = np.array([0.48296291314469, 0.83651630373747, 0.22414386804186, -0.12940952255092])
    for rows in (shear.nscales, shear.nscales + 2):
        E = np.ones(rows * max(shear.ndir))
        rc = real_lib.shx_separate(shear.handle, ctypes.c_void_p(img.data_ptr()), 2, 3, 3.0, 3.0,
                                   coeff.ctypes.data_as(dp), coeff.size, qmf.ctypes.data_as(dp), qmf.size, 0.0,
                                   E.ctypes.data_as(dp), rows, max(shear.ndir), ctypes.c_void_p(C.data_ptr()),
                                   ctypes.c_void_p(P.data_ptr()), None, None)
        assert rc == -1
        assert b"rows" in real_lib.shx_last_error()
