"""GPU parity of the 2D NSST path against the FP64 oracle (through the C ABI).

Accuracy bar (FP32 GPU vs FP64 oracle): per band, max|err|/max|ref| and
RMS(err)/RMS(ref) <= 1e-5 for the linear stages.  Hard-thresholded outputs are
compared under the GPU's own keep-mask applied to the oracle's coefficients,
and the number of keep/zero decisions that differ is reported and bounded.
"""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_errs(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    e = got - ref
    return np.abs(e).max() / np.abs(ref).max(), np.sqrt(np.mean(e * e) / np.mean(ref * ref))


def assert_close(got, ref, tol=TOL, what=""):
    mx, rms = rel_errs(got, ref)
    print("%s max-rel %.3e rms-rel %.3e" % (what, mx, rms), flush=True)
    assert mx <= tol and rms <= tol, (what, mx, rms)


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


_DICTS = {}


def oracle_dict(L, shift="none"):
    key = (L, shift)
    if key not in _DICTS:
        _DICTS[key] = o.make_dictionary(L, shift=shift)
    return _DICTS[key]


def image(L, seed=0, sigma=20.0):
    from shearcuda_amd.synth import noisy_ellipses
    return noisy_ellipses(L, sigma=sigma, seed=seed)


def to_gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("L,nlev", [(256, 4), (512, 4), (64, 3), (32, 2)])
def test_atrousdec_matches_oracle(api, L, nlev):
    print("case atrousdec L=%d nlev=%d" % (L, nlev), flush=True)
    _, x = image(L, seed=L)
    y = api.atrousdec_cuda(to_gpu(x), "maxflat", nlev)
    torch.cuda.synchronize()
    ref = o.atrousdec(x, "maxflat", nlev, strict=False)
    for j, (g, r) in enumerate(zip(y, ref)):
        assert_close(g.cpu().numpy(), r, what="L=%d band %d" % (L, j))


@pytest.mark.parametrize("L,nlev", [(256, 4), (64, 3)])
def test_atrousrec_matches_oracle(api, L, nlev):
    print("case atrousrec L=%d nlev=%d" % (L, nlev), flush=True)
    rng = np.random.default_rng(L)
    y = [rng.standard_normal((L, L)) * (10 if j else 100) for j in range(nlev + 1)]
    got = api.atrousrec_cuda([to_gpu(a) for a in y], "maxflat").cpu().numpy()
    assert_close(got, o.atrousrec(y, strict=False), what="rec L=%d" % L)


@pytest.mark.parametrize("L,shift", [(256, "none"), (512, "none"), (256, "center")])
def test_shear_trans_matches_oracle(api, L, shift):
    print("case shear_trans L=%d shift=%s" % (L, shift), flush=True)
    _, x = image(L, seed=3)
    shear = api.make_dictionary(L, shift=shift)
    C = api.shear_trans_cuda(to_gpu(x), "maxflat", shear)
    torch.cuda.synchronize()
    ref = o.nsst_dec2(x, oracle_dict(L, shift))
    assert_close(C[0].cpu().numpy(), ref[0], what="lowpass")
    for j in range(1, len(ref)):
        g = C[j].cpu().numpy()
        for k in range(ref[j].shape[0]):
            assert_close(g[k], ref[j][k], what="scale %d dir %d" % (j, k))
    shear.close()


@pytest.mark.parametrize("L", [256, 512])
def test_inverse_shear_matches_oracle(api, L):
    print("case inverse_shear L=%d" % L, flush=True)
    _, x = image(L, seed=4)
    H = oracle_dict(L)
    ref_c = o.nsst_dec2(x, H)
    rng = np.random.default_rng(9)
    # perturb the bands so the input is not a consistent transform
    ref_c = [c + rng.standard_normal(c.shape) * 2.0 for c in ref_c]
    shear = api.make_dictionary(L)
    got = api.inverse_shear_cuda([to_gpu(c) for c in ref_c], "maxflat", shear).cpu().numpy()
    assert_close(got, o.nsst_rec2(ref_c, H), what="inverse L=%d" % L)
    shear.close()


@pytest.mark.parametrize("L", [256, 1024])
def test_round_trip(api, L):
    print("case round trip L=%d" % L, flush=True)
    _, x = image(L, seed=5)
    shear = api.make_dictionary(L)
    xg = to_gpu(x)
    C = api.shear_trans_cuda(xg, "maxflat", shear)
    xr = api.inverse_shear_cuda(C, "maxflat", shear)
    assert_close(xr.cpu().numpy(), x, what="round trip L=%d" % L)
    shear.close()


def test_denoise_matches_oracle_under_keep_mask(api):
    L, sigma = 256, 20.0
    print("case denoise L=%d" % L, flush=True)
    from shearcuda_amd.synth import psnr, white_noise
    clean, x = image(L, seed=11, sigma=sigma)
    H = oracle_dict(L)
    noise = white_noise(L)
    sc = o.nsst_scalars(H, noise)
    th = o.nsst_thresholds(sigma, (0, 3, 4), sc)
    shear = api.make_dictionary(L)
    out, C = api.denoise_cuda(to_gpu(x), shear, sigma, sc)
    out = out.cpu().numpy()
    # oracle coefficients thresholded with the GPU's keep decisions
    ref_c = o.nsst_dec2(x, H)
    gpu_c = [c.cpu().numpy() for c in C]
    flips, total = 0, 0
    masked = []
    b = 0
    for j, (rc, gc) in enumerate(zip(ref_c, gpu_c)):
        rc3 = rc if rc.ndim == 3 else rc[None]
        gc3 = gc if gc.ndim == 3 else gc[None]
        mk = np.zeros_like(rc3)
        for k in range(rc3.shape[0]):
            keep_g = np.abs(gc3[k]) > np.float32(th[b])
            keep_o = np.abs(rc3[k]) > th[b]
            flips += int(np.sum(keep_g != keep_o))
            total += keep_g.size
            mk[k] = rc3[k] * keep_g
            b += 1
        masked.append(mk if rc.ndim == 3 else mk[0])
    print("threshold decision flips: %d of %d" % (flips, total), flush=True)
    assert flips <= 1e-4 * total
    assert_close(out, o.nsst_rec2(masked, H), what="denoise under GPU keep-mask")
    ref_out = o.nsst_rec2(o.nsst_HT(ref_c, sigma, (0, 3, 4), sc), H)
    p_gpu, p_ref = psnr(clean, out), psnr(clean, ref_out)
    print("PSNR noisy %.3f gpu %.3f oracle %.3f" % (psnr(clean, x), p_gpu, p_ref), flush=True)
    assert abs(p_gpu - p_ref) < 0.01
    assert p_gpu > psnr(clean, x) + 5
    shear.close()


def test_thresh_cuda_matches_oracle(api):
    print("case thresh_cuda", flush=True)
    L = 256
    _, x = image(L, seed=12)
    H = oracle_dict(L)
    ref_c = o.nsst_dec2(x, H)
    E = o.com_norm(H, L)
    sc = [0.0, 3.0, 3.0, 3.0, 4.0]
    Ct = [to_gpu(c) for c in ref_c]
    maxi = api.thresh_cuda(Ct, 20.0, 2, E, sc, 1)
    ref32 = [c.astype(np.float32).astype(np.float64) for c in ref_c]
    out, rmaxi = o.thresh(ref32, 20.0, 2, E, sc, 1)
    for j in range(len(Ct)):
        thr = np.array([np.float32(sc[j] * 20.0 * E[j, w]) for w in range(ref_c[j].shape[0] if ref_c[j].ndim == 3 else 1)])
        g = Ct[j].cpu().numpy()
        # exact: same FP32 inputs, same FP32 thresholds
        r = ref32[j] if ref32[j].ndim == 3 else ref32[j][None]
        g3 = g if g.ndim == 3 else g[None]
        for w in range(r.shape[0]):
            assert np.array_equal(g3[w], r[w] * (np.abs(r[w]) > thr[w]))
    assert maxi == pytest.approx(rmaxi, rel=1e-6)


def test_re_weight_band_matches_oracle(api):
    print("case re_weight_band", flush=True)
    L = 256
    _, x = image(L, seed=13)
    w = [1.0, 0.5, 2.0, 1.5, 0.0]
    got = api.re_weight_band_cuda(to_gpu(x), w).cpu().numpy()
    assert_close(got, o.re_weight_band(x, w), what="re_weight_band")


def test_norm_and_softth(api):
    print("case norm/softth", flush=True)
    rng = np.random.default_rng(14)
    a = rng.standard_normal((3, 64, 64)).astype(np.float32)
    t = to_gpu(a)
    assert api.norm_cuda(t, 2, 2) == pytest.approx(np.linalg.norm(a[1].astype(np.float64)), rel=1e-6)
    assert api.norm_cuda(t, 1) == pytest.approx(np.abs(a.astype(np.float64)).sum(), rel=1e-6)
    api.softth_cuda(t, 0.5)
    assert np.array_equal(t.cpu().numpy(), (np.sign(a) * np.maximum(np.abs(a) - np.float32(0.5), 0)).astype(np.float32))


def test_denoise_graph_replays_match_eager(api):
    print("case denoise graph", flush=True)
    L, sigma = 256, 20.0
    _, x = image(L, seed=21, sigma=sigma)
    shear = api.make_dictionary(L)
    xg = to_gpu(x)
    sc = api.nsst_scalars_cuda(shear, to_gpu(np.random.default_rng(3).standard_normal((L, L))))
    ref, _ = api.denoise_cuda(xg, shear, sigma, sc)
    ref = ref.cpu().numpy()
    g, Ct, out = api.denoise_graph(xg, shear, sigma, sc)
    for r in range(3):
        out.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), r
        xg.mul_(1.0)  # eager work between replays
    g.close()
    shear.close()


@pytest.mark.parametrize("L,nimg", [(256, 3), (512, 2)])
def test_batch_matches_plain(api, L, nimg):
    print("case batch L=%d nimg=%d" % (L, nimg), flush=True)
    shear = api.make_dictionary(L)
    xs = np.stack([image(L, seed=30 + i, sigma=10.0 + 5 * i)[1] for i in range(nimg)])
    xg = to_gpu(xs)
    sc = api.nsst_scalars_cuda(shear, to_gpu(np.random.default_rng(4).standard_normal((L, L))))
    sig = [10.0 + 5 * i for i in range(nimg)]
    th = api.batch_thresholds(shear, sig, sc)
    b = api.Batch(shear, nimg)
    coeff = b.alloc_coeffs()
    out = torch.empty_like(xg)
    b.denoise(xg, coeff, th, out)
    # forward and plain inverse through the batch API
    coeff2 = b.alloc_coeffs()
    b.shear_trans(xg, coeff2)
    rt = torch.empty_like(xg)
    b.inverse_shear(coeff2, rt)
    torch.cuda.synchronize()
    for i in range(nimg):
        pd, Ct = api.denoise_cuda(xg[i].contiguous(), shear, sig[i], sc)
        assert np.array_equal(out[i].cpu().numpy(), pd.cpu().numpy()), i
        flat = torch.cat([Ct[0][None]] + list(Ct[1:])).cpu().numpy()
        assert np.array_equal(coeff[i].cpu().numpy(), flat), i
        # the forward transform against the plain forward transform (the same
        # kernels; the denoise computes the coefficients in its fused row pass)
        Cf = api.shear_trans_cuda(xg[i].contiguous(), "maxflat", shear)
        flat_f = torch.cat([Cf[0][None]] + list(Cf[1:])).cpu().numpy()
        assert np.array_equal(coeff2[i].cpu().numpy(), flat_f), i
        assert_close(flat_f, flat, tol=1e-6, what="forward vs denoise coefficients %d" % i)
        assert_close(rt[i].cpu().numpy(), xs[i], what="batch round trip %d" % i)
    b.close()
    shear.close()


@pytest.mark.parametrize("parallel", [False, True])
def test_mixed_size_batches_in_one_graph(api, parallel):
    print("case mixed graph parallel=%s" % parallel, flush=True)
    plans = []
    for L, n in [(256, 2), (512, 1)]:
        shear = api.make_dictionary(L)
        xs = np.stack([image(L, seed=40 + i)[1] for i in range(n)])
        sc = api.nsst_scalars_cuda(shear, to_gpu(np.random.default_rng(5).standard_normal((L, L))))
        b = api.Batch(shear, n)
        plans.append((shear, b, to_gpu(xs), b.alloc_coeffs(), api.batch_thresholds(shear, [20.0] * n, sc),
                      torch.empty((n, L, L), device="cuda"), sc))
    if parallel:   # one graph branch per plan (no order between them)
        g = api.capture_parallel([lambda p=p: p[1].denoise(p[2], p[3], p[4], p[5]) for p in plans])
    else:
        g = api.capture(lambda: [b.denoise(x, c, t, o) for (_, b, x, c, t, o, _) in plans])
    for _ in range(3):
        for p in plans:
            p[5].zero_()
        g.launch()
        torch.cuda.synchronize()
        for shear, b, x, c, t, o, sc in plans:
            for i in range(b.nimg):
                pd, _ = api.denoise_cuda(x[i].contiguous(), shear, 20.0, sc)
                assert np.array_equal(o[i].cpu().numpy(), pd.cpu().numpy())
    g.close()
    for p in plans:
        p[1].close()
        p[0].close()


@pytest.mark.parametrize("dct", [False, True])
def test_denoise_without_coefficients(api, dct):
    """A denoise whose caller takes no coefficients (coeff NULL: the band
    coefficients are never written) returns the same image bit for bit, plain
    and batched, with either decomposition front end."""
    L, n = 256, 2
    print("case denoise without coefficients dct=%s" % dct, flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    try:
        api.set_dec_front(shear, dct)
        api.set_dec_front(shear, dct, batch=b)
        sc = api.nsst_scalars_cuda(shear, to_gpu(np.random.default_rng(6).standard_normal((L, L))))
        _, x = image(L, seed=61)
        xg = to_gpu(x)
        ref, _ = api.denoise_cuda(xg, shear, 20.0, sc)
        got, C = api.denoise_cuda(xg, shear, 20.0, sc, coeffs=False)
        assert C is None
        assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy())
        g, _, gout = api.denoise_graph(xg, shear, 20.0, sc, coeffs=False)
        for _ in range(3):
            gout.zero_()
            g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(gout.cpu().numpy(), ref.cpu().numpy())
        g.close()
        sig = [15.0, 25.0]
        xs = to_gpu(np.stack([image(L, seed=62 + i, sigma=s)[1] for i, s in enumerate(sig)]))
        th = api.batch_thresholds(shear, sig, sc)
        o1, o2 = torch.empty_like(xs), torch.empty_like(xs)
        b.denoise(xs, b.alloc_coeffs(), th, o1)
        b.denoise(xs, None, th, o2)
        torch.cuda.synchronize()
        assert np.array_equal(o1.cpu().numpy(), o2.cpu().numpy())
    finally:
        b.close()
        shear.close()
