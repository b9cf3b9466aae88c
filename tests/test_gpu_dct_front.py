"""The DCT-domain decomposition front end against the pyramid front end and
the FP64 oracle.

With the half-sample symmetric extension of the a trous pyramid
(atrousdec.m with symext.m, folded with period 2L) and the symmetric maxflat
filters, every pyramid level is a diagonal mask on the 2D DCT-II
coefficients.  The front end computes the lowpass and the scale spectra of all
levels in three passes from those masks (shx_set_dec_front).  Both front ends
must give the same coefficients to FP32 rounding, and the DCT front end's
coefficients must meet the oracle bar of the pyramid front end (max and RMS
relative error <= 1e-5 per band).
"""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    e = a - b
    return np.abs(e).max() / np.abs(b).max(), np.sqrt(np.mean(e * e) / np.mean(b * b))


def _cells(C):
    return [c.cpu().numpy() for c in C]


@pytest.mark.parametrize("L,dcomp,dsize", [
    (64, None, None), (128, None, None), (256, None, None), (512, None, None), (1024, None, None),
    (256, (2, 3), (32, 16)),                       # two levels
    (256, (3, 3, 4, 4, 4), (32, 32, 16, 16, 16)),  # five levels (dilation 16)
    (128, (1,), (16,)),                            # one level
])
def test_fronts_agree(gpu, L, dcomp, dsize):
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses
    print("case dct front vs pyramid L=%d dcomp=%s" % (L, dcomp), flush=True)
    shear = api.make_dictionary(L) if dcomp is None else api.make_dictionary(L, dcomp=dcomp, dsize=dsize)
    try:
        _, x = noisy_ellipses(L, sigma=20.0, seed=21)
        xg = _gpu(x)
        assert not api.get_dec_front(shear)
        ref = _cells(api.shear_trans_cuda(xg, "maxflat", shear))
        api.set_dec_front(shear, True)
        assert api.get_dec_front(shear)
        got = _cells(api.shear_trans_cuda(xg, "maxflat", shear))
        worst = 0.0
        for j, (g, r) in enumerate(zip(got, ref)):
            g3 = g.reshape(-1, L, L)
            r3 = r.reshape(-1, L, L)
            for k in range(r3.shape[0]):
                mx, rms = _rel(g3[k], r3[k])
                worst = max(worst, mx)
                assert mx <= TOL and rms <= TOL, (L, j, k, mx, rms)
        print("L=%d worst max-rel diff between the fronts %.2e" % (L, worst), flush=True)
    finally:
        shear.close()


@pytest.mark.parametrize("L,shift", [(256, "none"), (512, "none"), (256, "center")])
def test_dct_front_matches_oracle(gpu, L, shift):
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses
    print("case dct front vs oracle L=%d shift=%s" % (L, shift), flush=True)
    _, x = noisy_ellipses(L, sigma=20.0, seed=3)
    shear = api.make_dictionary(L, shift=shift)
    try:
        api.set_dec_front(shear, True)
        C = _cells(api.shear_trans_cuda(_gpu(x), "maxflat", shear))
        ref = o.nsst_dec2(x, o.make_dictionary(L, shift=shift))
        mx, rms = _rel(C[0], ref[0])
        print("lowpass max-rel %.3e rms-rel %.3e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
        for j in range(1, len(ref)):
            for k in range(ref[j].shape[0]):
                mx, rms = _rel(C[j][k], ref[j][k])
                assert mx <= TOL and rms <= TOL, (j, k, mx, rms)
    finally:
        shear.close()


def test_denoise_and_batch_with_dct_front(gpu):
    """Denoise with the DCT front end: the output is the reconstruction of its
    own coefficients under its own keep decisions (|c| > th), the coefficients
    agree with the pyramid front end's, and the batched concurrent plan
    reproduces the plain plan bit for bit (same kernels per image)."""
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    L, n = 256, 3
    print("case dct front denoise + batch L=%d n=%d" % (L, n), flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    try:
        sc = api.nsst_scalars_cuda(shear, _gpu(white_noise(L)))
        _, x = noisy_ellipses(L, sigma=20.0, seed=12)
        xg = _gpu(x)
        _, cref = api.denoise_cuda(xg, shear, 20.0, sc)
        cref = _cells(cref)
        api.set_dec_front(shear, True)
        got, C = api.denoise_cuda(xg, shear, 20.0, sc)
        C = _cells(C)
        for j, (g, r) in enumerate(zip(C, cref)):
            mx, rms = _rel(g, r)
            assert mx <= TOL and rms <= TOL, (j, mx, rms)
        th = api.band_thresholds(shear, 20.0, sc).astype(np.float32)
        masked, k = [], 0
        for c in C:
            c3 = c.reshape(-1, L, L)
            m = np.stack([np.where(np.abs(c3[i]) > th[k + i], c3[i], 0.0) for i in range(c3.shape[0])])
            k += c3.shape[0]
            masked.append(_gpu(m.reshape(c.shape)))
        api.set_dec_front(shear, False)
        rec = api.inverse_shear_cuda(masked, "maxflat", shear).cpu().numpy()
        mx, rms = _rel(got.cpu().numpy(), rec)
        print("plain denoise vs masked reconstruction max-rel %.2e rms-rel %.2e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
        api.set_dec_front(shear, True)
        sig = np.array([10.0, 20.0, 30.0])
        xs = _gpu(np.stack([noisy_ellipses(L, sigma=float(s), seed=50 + i)[1] for i, s in enumerate(sig)]))
        thb = api.batch_thresholds(shear, sig, sc)
        api.set_dec_front(shear, True, batch=b)
        assert api.get_dec_front(shear, batch=b)
        out = torch.empty_like(xs)
        cb = b.alloc_coeffs()
        b.denoise(xs, cb, thb, out)
        torch.cuda.synchronize()
        for i in range(n):
            pd, Ct = api.denoise_cuda(xs[i].contiguous(), shear, float(sig[i]), sc)
            assert np.array_equal(out[i].cpu().numpy(), pd.cpu().numpy()), i
            flat = torch.cat([Ct[0][None]] + list(Ct[1:])).cpu().numpy()
            assert np.array_equal(cb[i].cpu().numpy(), flat), i
        # the batched forward transform through the batch API
        c2 = b.alloc_coeffs()
        b.shear_trans(xs, c2)
        api.set_dec_front(shear, False, batch=b)
        c3 = b.alloc_coeffs()
        b.shear_trans(xs, c3)
        mx, rms = _rel(c2.cpu().numpy(), c3.cpu().numpy())
        print("batched shear_trans max-rel %.2e rms-rel %.2e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
    finally:
        b.close()
        shear.close()


def test_dct_front_needs_l64(gpu):
    from shearcuda_amd import api
    from shearcuda_amd._lib import ShxError
    shear = api.make_dictionary(32)
    try:
        with pytest.raises(ShxError):
            api.set_dec_front(shear, True)
        assert not api.get_dec_front(shear)
    finally:
        shear.close()
