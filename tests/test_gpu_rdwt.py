"""GPU parity of the undecimated wavelet transform and its adjoints."""
import numpy as np
import pytest
import torch

from oracle import rdwt as w

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def relmax(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name,par,L,shape", [("Daubechies", 4, 3, (512, 512)), ("Symmlet", 4, 2, (64, 128)),
                                              ("Haar", 0, 1, (32, 32))])
def test_mrdwt_mirdwt_match_oracle(api, name, par, L, shape):
    print("case rdwt", name, L, shape, flush=True)
    h = w.make_on_filter(name, par)
    x = np.random.default_rng(L).standard_normal(shape) * 50
    yl, yh = api.mrdwt_cuda(g(x), h, L)
    ryl, ryh = w.mrdwt(x, h, L)
    assert relmax(yl.cpu().numpy(), ryl) < 1e-5
    for lev in range(L):
        for q in range(3):
            c0 = (3 * lev + q) * shape[1]
            assert relmax(yh.cpu().numpy()[:, c0:c0 + shape[1]], ryh[:, c0:c0 + shape[1]]) < 1e-5
    xr = api.mirdwt_cuda(g(ryl), g(ryh), h, L).cpu().numpy()
    assert relmax(xr, w.mirdwt(ryl, ryh, h, L)) < 1e-5


def test_rdwt_batched_equals_single(api):
    h = w.make_on_filter("Daubechies", 4)
    xs = np.random.default_rng(3).standard_normal((3, 64, 64))
    yl, yh = api.mrdwt_cuda(g(xs), h, 3)
    for i in range(3):
        a, b = api.mrdwt_cuda(g(xs[i]), h, 3)
        assert torch.equal(a, yl[i]) and torch.equal(b, yh[i])


def test_rdwt_adjoints_dense(api):
    """Dense matrices of mrdwt / mirdwt on 16x16 (3 levels) from the oracle;
    the adjoint kernels must apply their transposes."""
    print("case rdwt dense adjoint", flush=True)
    h = w.make_on_filter("Daubechies", 4)
    m = n = 16
    L = 3
    N = m * n
    A = np.zeros((N + 3 * N * L, N))   # analysis matrix, rows (yl, yh)
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1
        yl, yh = w.mrdwt(e.reshape(m, n), h, L)
        A[:, i] = np.concatenate([yl.ravel(), yh.ravel()])
    S = np.zeros((N, N + 3 * N * L))   # synthesis matrix
    for i in range(N + 3 * N * L):
        e = np.zeros(N + 3 * N * L)
        e[i] = 1
        S[:, i] = w.mirdwt(e[:N].reshape(m, n), e[N:].reshape(m, 3 * n * L), h, L).ravel()
    rng = np.random.default_rng(7)
    # batch of probes: adjoint of analysis
    P = rng.standard_normal((8, N + 3 * N * L))
    xb = api.mrdwt_adjoint_cuda(g(P[:, :N].reshape(8, m, n)), g(P[:, N:].reshape(8, m, 3 * n * L)), h, L)
    ref = P @ A                      # rows: A^T p
    assert relmax(xb.cpu().numpy().reshape(8, N), ref) < 1e-5
    Q = rng.standard_normal((8, N))
    ylb, yhb = api.mirdwt_adjoint_cuda(g(Q.reshape(8, m, n)), h, L)
    got = np.concatenate([ylb.cpu().numpy().reshape(8, N), yhb.cpu().numpy().reshape(8, -1)], axis=1)
    assert relmax(got, Q @ S) < 1e-5
