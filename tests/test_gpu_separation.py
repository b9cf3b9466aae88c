"""Unrolled separation (learnable soft thresholds): the hand-written adjoint
gradients against FP64 central finite differences of the oracle."""
import numpy as np
import pytest
import torch

from oracle import nsst as o
from oracle import rdwt as w
from oracle import separation as osep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods(gpu):
    from shearcuda_amd import api, separation
    return api, separation


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_soft_bands_forward_backward(mods):
    api, sep = mods
    rng = np.random.default_rng(0)
    B, nb, L = 2, 5, 16
    x = rng.standard_normal((B, nb, L, L)).astype(np.float32)
    th = np.abs(rng.standard_normal(nb)).astype(np.float32)
    gy = rng.standard_normal(x.shape).astype(np.float32)
    y = sep.soft_bands(g(x), g(th), B, nb * L * L, nb * L * L, L * L).cpu().numpy()
    t = th[None, :, None, None]
    assert np.array_equal(y, (np.sign(x) * np.maximum(np.abs(x) - t, 0)).astype(np.float32))
    gth = torch.zeros(nb, device="cuda")
    gx = sep.soft_bands_backward(g(x), g(gy), g(th), gth, B, nb * L * L, nb * L * L, L * L).cpu().numpy()
    keep = np.abs(x) > t
    assert np.array_equal(gx, gy * keep)
    ref = np.sum(-np.sign(x) * keep * gy, axis=(0, 2, 3))
    assert np.allclose(gth.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    # wavelet yh layout: rows of 3nL, bands of n
    m, n, lev = 8, 8, 2
    yh = rng.standard_normal((B, m, 3 * n * lev)).astype(np.float32)
    thw = np.abs(rng.standard_normal(1 + 3 * lev)).astype(np.float32)
    out = sep.soft_bands(g(yh), g(thw), B, m * 3 * n * lev, 3 * n * lev, n, 1).cpu().numpy()
    tb = np.repeat(thw[1:], n)[None, None, :]
    assert np.array_equal(out, (np.sign(yh) * np.maximum(np.abs(yh) - tb, 0)).astype(np.float32))


def test_unrolled_gradients_match_fp64_finite_differences(mods):
    """64x64 crop, 4-scale NSST [8 8 16 16] + 3-level D4, 10 unrolled steps."""
    print("case unrolled FD 64x64", flush=True)
    api, sep = mods
    from shearcuda_amd.synth import separation_input
    L, sigma, levels, iters = 64, 10.0, 3, 10
    _, curves, x = separation_input(L, seed=5, sigma=sigma)
    H = o.make_dictionary(L)
    h = w.make_on_filter("Daubechies", 4)
    noise = np.random.default_rng(11).standard_normal((L, L))
    s_sh = o.nsst_thresholds(1.0, (1.0, 1.0, 1.0), o.nsst_scalars(H, noise, strict=False))
    yl, yh = w.mrdwt(noise, h, levels)
    mad = lambda v: np.median(np.abs(v - np.median(v))) / 0.6745
    s_w = np.array([mad(yl)] + [mad(yh[:, k * L:(k + 1) * L]) for k in range(3 * levels)])
    th_sh = 3 * sigma * s_sh
    th_w = 3 * sigma * s_w

    def loss64(xx, tsh, tw):
        c, _ = osep.unrolled(xx, H, tsh, h, levels, tw, iters)
        return np.mean((c - curves) ** 2)

    # GPU forward/backward through the adjoint kernels
    shear = api.make_dictionary(L)
    plan = api.Batch(shear, 1)
    xg = g(x[None]).requires_grad_(True)
    tsg = g(th_sh).requires_grad_(True)
    twg = g(th_w).requires_grad_(True)
    c, _ = sep.unrolled_separation(xg, tsg, twg, plan, h, levels, iters)
    loss = torch.mean((c[0] - g(curves)) ** 2)
    loss.backward()
    ref_loss = loss64(x, th_sh, th_w)
    print("loss gpu %.6f oracle %.6f" % (loss.item(), ref_loss), flush=True)
    assert abs(loss.item() - ref_loss) <= 1e-4 * ref_loss
    gx = xg.grad[0].cpu().numpy()
    gs = tsg.grad.cpu().numpy()
    gw = twg.grad.cpu().numpy()
    rng = np.random.default_rng(3)
    pix = [tuple(p) for p in rng.integers(0, L, size=(5, 2))] + [(int(np.argmax(np.abs(gx)) // L),
                                                                  int(np.argmax(np.abs(gx)) % L))]
    eps = 1e-3
    for (r, cc) in pix:
        xp, xm = x.copy(), x.copy()
        xp[r, cc] += eps
        xm[r, cc] -= eps
        fd = (loss64(xp, th_sh, th_w) - loss64(xm, th_sh, th_w)) / (2 * eps)
        print("pixel", r, cc, "gpu %.6e fd %.6e" % (gx[r, cc], fd), flush=True)
        assert abs(gx[r, cc] - fd) <= 2e-3 * np.abs(gx).max() + 1e-7
    for k in [0, 3, 17, 40]:
        tp, tm = th_sh.copy(), th_sh.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (loss64(x, tp, th_w) - loss64(x, tm, th_w)) / (2 * eps)
        print("th_sh", k, "gpu %.6e fd %.6e" % (gs[k], fd), flush=True)
        assert abs(gs[k] - fd) <= 2e-3 * np.abs(gs).max() + 1e-7
    for k in [0, 1, 5, 9]:
        tp, tm = th_w.copy(), th_w.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (loss64(x, th_sh, tp) - loss64(x, th_sh, tm)) / (2 * eps)
        print("th_w", k, "gpu %.6e fd %.6e" % (gw[k], fd), flush=True)
        assert abs(gw[k] - fd) <= 2e-3 * np.abs(gw).max() + 1e-7
    plan.close()
    shear.close()


def test_hard_bands_forward_backward(gpu):
    print("case hard_bands", flush=True)
    from shearcuda_amd import separation as sep
    rng = np.random.default_rng(4)
    nimg, nb, plane = 3, 5, 64
    x = torch.tensor(rng.standard_normal((nimg, nb * plane)), dtype=torch.float32, device="cuda")
    gy = torch.tensor(rng.standard_normal((nimg, nb * plane)), dtype=torch.float32, device="cuda")
    th = torch.tensor(np.linspace(0.2, 1.2, nb), dtype=torch.float32, device="cuda")
    y = sep.hard_bands(x, th, nimg, nb * plane, nb * plane, plane)
    gx = sep.hard_bands_backward(x, gy, th, nimg, nb * plane, nb * plane, plane)
    t = th.repeat_interleave(plane)[None, :]
    keep = x.abs() > t
    assert torch.equal(y, torch.where(keep, x, torch.zeros_like(x)))
    assert torch.equal(gx, torch.where(keep, gy, torch.zeros_like(gy)))
