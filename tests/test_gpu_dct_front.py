"""The DCT-domain pyramid (decomposition front end and reconstruction back
end) against the level-by-level pyramid and the FP64 oracle.

With the half-sample symmetric extension of the a trous pyramid
(atrousdec.m / atrousrec.m with symext.m, folded with period 2L) and the
symmetric maxflat filters, every pyramid level is a diagonal mask on the 2D
DCT-II coefficients.  The front end computes the lowpass and the scale
spectra of all levels in three passes from those masks, the back end the
image from the scale partials and the lowpass (shx_set_plan_options).  Both
forms must agree to FP32 rounding, and the DCT forms must meet the oracle bar
of the pyramid forms (max and RMS relative error <= 1e-5 per band / image).
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


def _rand_cells(api, shear, L, seed):
    rng = np.random.default_rng(seed)
    cells = []
    for c in shear.alloc_coeffs():
        cells.append(_gpu(rng.standard_normal(tuple(c.shape)) * (30.0 if not cells else 3.0)))
    return cells


@pytest.mark.parametrize("L,dcomp,dsize", [
    (64, None, None), (128, None, None), (256, None, None), (512, None, None), (1024, None, None),
    (256, (2, 3), (32, 16)),
    (256, (3, 3, 4, 4, 4), (32, 32, 16, 16, 16)),
    (128, (1,), (16,)),
])
def test_backs_agree(gpu, L, dcomp, dsize):
    """inverse_shear with the DCT back end against the scale row pass +
    pyramid levels, on inconsistent random coefficients."""
    from shearcuda_amd import api
    print("case dct back vs pyramid L=%d dcomp=%s" % (L, dcomp), flush=True)
    shear = api.make_dictionary(L) if dcomp is None else api.make_dictionary(L, dcomp=dcomp, dsize=dsize)
    try:
        cells = _rand_cells(api, shear, L, 31)
        assert not api.get_plan_options(shear)["dct_back"]
        ref = api.inverse_shear_cuda(cells, "maxflat", shear).cpu().numpy()
        api.set_plan_options(shear, dct_back=True)
        assert api.get_plan_options(shear) == {"dct_front": False, "dct_back": True}
        got = api.inverse_shear_cuda(cells, "maxflat", shear).cpu().numpy()
        mx, rms = _rel(got, ref)
        print("L=%d inverse max-rel %.2e rms-rel %.2e" % (L, mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
    finally:
        shear.close()


@pytest.mark.parametrize("L,shift", [(256, "none"), (512, "none"), (256, "center")])
def test_dct_back_matches_oracle(gpu, L, shift):
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses
    print("case dct back vs oracle L=%d shift=%s" % (L, shift), flush=True)
    _, x = noisy_ellipses(L, sigma=20.0, seed=4)
    H = o.make_dictionary(L, shift=shift)
    ref_c = o.nsst_dec2(x, H)
    rng = np.random.default_rng(9)
    ref_c = [c + rng.standard_normal(c.shape) * 2.0 for c in ref_c]
    shear = api.make_dictionary(L, shift=shift)
    try:
        api.set_plan_options(shear, dct_front=True, dct_back=True)
        got = api.inverse_shear_cuda([_gpu(c) for c in ref_c], "maxflat", shear).cpu().numpy()
        mx, rms = _rel(got, o.nsst_rec2(ref_c, H))
        print("inverse max-rel %.3e rms-rel %.3e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
    finally:
        shear.close()


def test_dct_both_denoise_and_batch(gpu):
    """Denoise with both DCT forms: against its own coefficients reconstructed
    under its keep decisions by the pyramid path, and the batched concurrent
    plan against the plain plan bit for bit."""
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    L, n = 256, 3
    print("case dct front + back denoise L=%d" % L, flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    try:
        sc = api.nsst_scalars_cuda(shear, _gpu(white_noise(L)))
        _, x = noisy_ellipses(L, sigma=20.0, seed=13)
        xg = _gpu(x)
        api.set_plan_options(shear, dct_front=True, dct_back=True)
        got, C = api.denoise_cuda(xg, shear, 20.0, sc)
        C = _cells(C)
        th = api.band_thresholds(shear, 20.0, sc).astype(np.float32)
        masked, k = [], 0
        for c in C:
            c3 = c.reshape(-1, L, L)
            m = np.stack([np.where(np.abs(c3[i]) > th[k + i], c3[i], 0.0) for i in range(c3.shape[0])])
            k += c3.shape[0]
            masked.append(_gpu(m.reshape(c.shape)))
        api.set_plan_options(shear, dct_front=False, dct_back=False)
        rec = api.inverse_shear_cuda(masked, "maxflat", shear).cpu().numpy()
        mx, rms = _rel(got.cpu().numpy(), rec)
        print("denoise vs masked pyramid reconstruction max-rel %.2e rms-rel %.2e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
        api.set_plan_options(shear, dct_front=True, dct_back=True)
        api.set_plan_options(shear, b, dct_front=True, dct_back=True)
        sig = np.array([10.0, 20.0, 30.0])
        xs = _gpu(np.stack([noisy_ellipses(L, sigma=float(s), seed=80 + i)[1] for i, s in enumerate(sig)]))
        thb = api.batch_thresholds(shear, sig, sc)
        out = torch.empty_like(xs)
        b.denoise(xs, None, thb, out)
        rt = torch.empty_like(xs)
        cb = b.alloc_coeffs()
        b.shear_trans(xs, cb)
        b.inverse_shear(cb, rt)
        torch.cuda.synchronize()
        for i in range(n):
            pd, _ = api.denoise_cuda(xs[i].contiguous(), shear, float(sig[i]), sc, coeffs=False)
            assert np.array_equal(out[i].cpu().numpy(), pd.cpu().numpy()), i
            mx, rms = _rel(rt[i].cpu().numpy(), xs[i].cpu().numpy())
            assert mx <= TOL and rms <= TOL, (i, mx, rms)
    finally:
        b.close()
        shear.close()


def test_dct_back_thresholds(gpu):
    """Thresholded reconstruction (host thresholds, and the batch's device
    table) with a nonzero lowpass threshold: the DCT back end applies the
    lowpass and band keep rules as the level-by-level path does."""
    from shearcuda_amd import api
    L, n = 256, 2
    print("case dct back thresholds L=%d" % L, flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    try:
        cells = _rand_cells(api, shear, L, 41)
        th = np.full(shear.nbands, 2.0)
        th[0] = 25.0  # lowpass threshold that removes part of the lowpass
        ref = api.inverse_shear_thresh_cuda(cells, th, shear).cpu().numpy()
        api.set_plan_options(shear, dct_back=True)
        got = api.inverse_shear_thresh_cuda(cells, th, shear).cpu().numpy()
        mx, rms = _rel(got, ref)
        print("host thresholds max-rel %.2e rms-rel %.2e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
        stack = torch.stack([torch.cat([c.reshape(-1, L, L) for c in _rand_cells(api, shear, L, 50 + i)])
                             for i in range(n)]).contiguous()
        thb = torch.tensor(np.stack([th * (1 + 0.5 * i) for i in range(n)]), dtype=torch.float32, device="cuda")
        outs = []
        for dct in (False, True):
            api.set_plan_options(shear, b, dct_back=dct)
            o = torch.empty((n, L, L), device="cuda")
            b.inverse_shear(stack, o, thb)
            outs.append(o.cpu().numpy())
        mx, rms = _rel(outs[1], outs[0])
        print("device thresholds max-rel %.2e rms-rel %.2e" % (mx, rms), flush=True)
        assert mx <= TOL and rms <= TOL
    finally:
        b.close()
        shear.close()
