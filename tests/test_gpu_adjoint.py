"""Hand-written adjoints of the NSST analysis and synthesis against dense
transposes built column by column from the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu

DCOMP = (2, 3, 3)
DSIZE = (16, 16, 8)


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def relmax(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def dense32():
    """Dense analysis A (coefficients x pixels) of the NSST on 32x32 (3 scales)
    from the oracle, built column by column, and the dense pyramid
    reconstruction P ((S+1) planes -> image) for the synthesis transpose
    S^T = F P^T, F the directional analysis (its transpose is the directional
    synthesis: tight frame, pinned in tests/test_oracle.py)."""
    L = 32
    H = o.make_dictionary(L, DCOMP, DSIZE)
    N = L * L
    cols = []
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        d = o.nsst_dec2(e.reshape(L, L), H, strict=False)
        cols.append(np.concatenate([d[0].ravel()] + [b.ravel() for b in d[1:]]))
    A = np.array(cols).T
    S = len(DCOMP)
    pcols = []
    for i in range((S + 1) * N):
        e = np.zeros((S + 1) * N)
        e[i] = 1.0
        pcols.append(o.atrousrec([e[j * N:(j + 1) * N].reshape(L, L) for j in range(S + 1)], strict=False).ravel())
    P = np.array(pcols).T
    nb = [1] + [2 ** c for c in DCOMP]
    return L, A, P, H, nb


def synthesis_transpose(y, L, P, H):
    """S^T y = (lowpass, F applied to each pyramid band) of P^T y."""
    S = len(H)
    z = (P.T @ y).reshape(S + 1, L, L)
    out = [z[0].ravel()]
    for j in range(S):
        Z = np.fft.fft2(z[j + 1])
        out.append(np.real(np.fft.ifft2(H[j] * Z[None])).ravel())
    return np.concatenate(out)


def to_cells(v, L, nb):
    out, off = [], 0
    for j, k in enumerate(nb):
        blk = v[off:off + k * L * L]
        out.append(g(blk.reshape(L, L) if j == 0 else blk.reshape(k, L, L)))
        off += k * L * L
    return out


def from_cells(cells):
    return np.concatenate([c.cpu().numpy().ravel() for c in cells])


def test_forward_matches_dense(api, dense32):
    L, A, P, H, nb = dense32
    shear = api.prepare_shearing_filters_cuda(o_filters(), L)
    x = np.random.default_rng(0).standard_normal((L, L))
    C = api.shear_trans_cuda(g(x), "maxflat", shear)
    assert relmax(from_cells(C), A @ x.ravel()) < 1e-5
    shear.close()


def o_filters():
    return [o.shearing_filters_myer(n1, lev) for lev, n1 in zip(DCOMP, DSIZE)]


def test_analysis_adjoint_dense(api, dense32):
    print("case analysis adjoint dense 32x32", flush=True)
    L, A, P, H, nb = dense32
    shear = api.prepare_shearing_filters_cuda(o_filters(), L)
    rng = np.random.default_rng(1)
    M = A.shape[0]
    # random probes and unit basis vectors (columns of A^T = rows of A)
    probes = [rng.standard_normal(M) for _ in range(8)]
    for i in rng.choice(M, 24, replace=False):
        e = np.zeros(M)
        e[i] = 1.0
        probes.append(e)
    for p in probes:
        got = api.shear_trans_adjoint_cuda(to_cells(p, L, nb), shear).cpu().numpy().ravel()
        ref = A.T @ p
        assert np.abs(got - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-3)
    shear.close()


def test_synthesis_adjoint_dense(api, dense32):
    print("case synthesis adjoint dense 32x32", flush=True)
    L, A, P, H, nb = dense32
    shear = api.prepare_shearing_filters_cuda(o_filters(), L)
    rng = np.random.default_rng(2)
    N = L * L
    probes = [rng.standard_normal(N) for _ in range(8)]
    for i in rng.choice(N, 24, replace=False):
        e = np.zeros(N)
        e[i] = 1.0
        probes.append(e)
    for p in probes:
        got = from_cells(api.inverse_shear_adjoint_cuda(g(p.reshape(L, L)), shear))
        ref = synthesis_transpose(p, L, P, H)
        assert np.abs(got - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-3)
    shear.close()


@pytest.mark.parametrize("L,dct", [(64, False), (512, False), (64, True), (512, True)])
def test_adjoint_identity_default_config(api, L, dct):
    """<A x, c> = <x, A^T c> and <S c, y> = <c, S^T y> with the 4-scale
    [8 8 16 16] dictionary (size-independent check), with the pyramid levels
    and with the DCT-domain forms (whose transposes run the other side's
    masks)."""
    print("case adjoint identity L=%d dct=%s" % (L, dct), flush=True)
    shear = api.make_dictionary(L)
    api.set_plan_options(shear, dct_front=dct, dct_back=dct)
    rng = np.random.default_rng(L)
    x = rng.standard_normal((L, L))
    c = [rng.standard_normal((L, L))] + [rng.standard_normal((d, L, L)) for d in shear.ndir]
    Ax = from_cells(api.shear_trans_cuda(g(x), "maxflat", shear)).astype(np.float64)
    Atc = api.shear_trans_adjoint_cuda([g(a) for a in c], shear).cpu().numpy().astype(np.float64)
    cf = np.concatenate([a.ravel() for a in c])
    lhs, rhs = Ax @ cf, np.sum(x * Atc)
    assert abs(lhs - rhs) <= 1e-4 * np.sqrt(np.sum(Ax * Ax) * np.sum(cf * cf))
    Sc = api.inverse_shear_cuda([g(a) for a in c], "maxflat", shear).cpu().numpy().astype(np.float64)
    Sty = from_cells(api.inverse_shear_adjoint_cuda(g(x), shear)).astype(np.float64)
    lhs, rhs = np.sum(Sc * x), Sty @ cf
    assert abs(lhs - rhs) <= 1e-4 * np.sqrt(np.sum(Sc * Sc) * np.sum(x * x))
    shear.close()


def test_batched_adjoints_equal_single(api):
    L = 64
    shear = api.make_dictionary(L)
    b = api.Batch(shear, 3)
    rng = np.random.default_rng(9)
    cb = g(rng.standard_normal((3, shear.nbands, L, L)))
    xb = torch.empty((3, L, L), device="cuda")
    b.shear_trans_adjoint(cb, xb)
    yb = b.alloc_coeffs()
    b.inverse_shear_adjoint(xb, yb)
    for i in range(3):
        cells = [cb[i, 0].contiguous()] + [cb[i, 1 + sum(shear.ndir[:j]):1 + sum(shear.ndir[:j + 1])].contiguous()
                                           for j in range(shear.nscales)]
        xs = api.shear_trans_adjoint_cuda(cells, shear)
        assert torch.equal(xs, xb[i])
        ys = api.inverse_shear_adjoint_cuda(xs, shear)
        assert torch.equal(torch.cat([ys[0][None]] + list(ys[1:])), yb[i])
    b.close()
    shear.close()


@pytest.mark.parametrize("L,nimg", [(64, 1), (256, 3), (512, 2)])
def test_dct_domain_adjoints_equal_level_adjoints(api, L, nimg):
    """A^T and S^T through the DCT domain against the transposed pyramid
    levels (pinned by the dense 32x32 tests and the identities above)."""
    print("case dct adjoints L=%d nimg=%d" % (L, nimg), flush=True)
    shear = api.make_dictionary(L)
    rng = np.random.default_rng(L + 7)
    cb = g(rng.standard_normal((nimg, shear.nbands, L, L)))
    xb = g(rng.standard_normal((nimg, L, L)))
    out = {}
    for dct in (False, True):
        b = api.Batch(shear, nimg)
        api.set_plan_options(shear, b, dct_front=dct, dct_back=dct)
        xa = torch.empty((nimg, L, L), device="cuda")
        b.shear_trans_adjoint(cb, xa)
        ya = b.alloc_coeffs()
        b.inverse_shear_adjoint(xb, ya)
        torch.cuda.synchronize()
        out[dct] = (xa, ya)
        b.close()
    ex = float((out[True][0] - out[False][0]).abs().max() / out[False][0].abs().max())
    ey = float((out[True][1] - out[False][1]).abs().max() / out[False][1].abs().max())
    print("A^T rel %.2e S^T rel %.2e" % (ex, ey), flush=True)
    assert ex <= 1e-5 and ey <= 1e-5
    shear.close()
