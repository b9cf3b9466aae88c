"""End-to-end parity of the plan the benchmark times, at the sizes it times.

The timed step is the plain denoise of one noisy-ellipse image (seed 0,
sigma 20) at 512^2 (and 1024^2) with the pyramid in the DCT-II domain on
both sides, 4 directions per reconstruction column block, column tiles of 8
on both column passes and no coefficient output.  Here that plan is set
explicitly and compared with the FP64 oracle:

* per band, the decomposition (shear_trans) against nsst_dec2
  (shearlet_toolbox_1/nsst_dec2.m:37-60): max|err|/max|ref| and
  RMS(err)/RMS(ref) <= 1e-5;
* the denoised image against nsst_rec2 (nsst_rec2.m:25-35) of the oracle's
  coefficients under the GPU's keep decisions (nsst_HT.m:32-47 rule |c| > th,
  thresholds T*sigma*s in FP32 as the library forms them), <= 1e-5, with the
  decisions that differ from the oracle's own counted and bounded;
* the image without coefficient output equals the image with it, bit for
  bit (the coefficient stores are the only difference);
* a 1024^2 batch plan with the schedule of the mixed-size benchmark equals
  the plain plan with the same schedule, bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIGMA = 20.0
T = (0.0, 3.0, 4.0)


def rel_errs(got, ref):
    e = np.asarray(got, np.float64) - ref
    return np.abs(e).max() / np.abs(ref).max(), np.sqrt(np.mean(e * e) / np.mean(ref * ref))


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


def to_gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def timed_plan(api, shear, batch=None, pyr=0, group=4, ct=8):
    api.set_plan_options(shear, batch, dct_front=True, dct_back=True)
    api.set_schedule(shear, pyr, group, batch, col_tile_dec=ct, col_tile_rec=ct)
    opts = api.get_plan_options(shear, batch)
    assert opts == {"dct_front": True, "dct_back": True}


@pytest.mark.parametrize("L", [512, 1024])
def test_timed_plan_matches_oracle(api, L):
    from shearcuda_amd.synth import noisy_ellipses, psnr, white_noise
    print("case timed plan L=%d" % L, flush=True)
    clean, x = noisy_ellipses(L, sigma=SIGMA, seed=0)
    shear = api.make_dictionary(L)
    timed_plan(api, shear)
    xg = to_gpu(x)
    sc = api.nsst_scalars_cuda(shear, to_gpu(white_noise(L)))   # shared noise scalars

    out_img, none = api.denoise_cuda(xg, shear, SIGMA, sc, coeffs=False)
    out_c, Cd = api.denoise_cuda(xg, shear, SIGMA, sc)
    Cf = api.shear_trans_cuda(xg, "maxflat", shear)
    torch.cuda.synchronize()
    assert none is None
    out_img = out_img.cpu().numpy()
    assert np.array_equal(out_img, out_c.cpu().numpy())

    H = o.make_dictionary(L)
    ref_c = o.nsst_dec2(x, H)
    th = api.band_thresholds(shear, SIGMA, sc, T).astype(np.float32)
    worst = (0.0, 0.0)
    flips = total = 0
    masked = []
    b = 0
    for j, r in enumerate(ref_c):
        r3 = r if r.ndim == 3 else r[None]
        gf = Cf[j].cpu().numpy().reshape(r3.shape)
        gd = Cd[j].cpu().numpy().reshape(r3.shape)
        mk = np.empty_like(r3)
        for k in range(r3.shape[0]):
            mx, rms = rel_errs(gf[k], r3[k])
            worst = (max(worst[0], mx), max(worst[1], rms))
            assert mx <= TOL and rms <= TOL, ("band", j, k, mx, rms)
            mxd, rmsd = rel_errs(gd[k], r3[k])
            assert mxd <= TOL and rmsd <= TOL, ("denoise band", j, k, mxd, rmsd)
            keep_g = np.abs(gd[k]) > th[b]
            keep_o = np.abs(r3[k]) > np.float64(th[b])
            flips += int(np.count_nonzero(keep_g != keep_o))
            total += keep_g.size
            mk[k] = r3[k] * keep_g
            b += 1
        masked.append(mk if r.ndim == 3 else mk[0])
    print("L=%d per-band worst max-rel %.3e rms-rel %.3e" % (L, worst[0], worst[1]), flush=True)
    ref_out = o.nsst_rec2(masked, H)
    mx, rms = rel_errs(out_img, ref_out)
    print("L=%d denoise under GPU keep-mask max-rel %.3e rms-rel %.3e; decision flips %d of %d"
          % (L, mx, rms, flips, total), flush=True)
    assert mx <= TOL and rms <= TOL
    assert flips <= 1e-4 * total
    p_gpu = psnr(clean, out_img)
    p_ref = psnr(clean, o.nsst_rec2(o.nsst_HT(ref_c, SIGMA, T, sc), H))
    print("L=%d PSNR noisy %.3f gpu %.3f oracle %.3f" % (L, psnr(clean, x), p_gpu, p_ref), flush=True)
    assert abs(p_gpu - p_ref) < 0.01
    shear.close()


def test_batch_plan_1024_equals_plain(api):
    """The 1024^2 leg of the mixed-size batch (schedule: pyramid tile 1,
    8 directions per block, tiles 8/8, DCT both sides) equals the plain plan
    under the same schedule, bit for bit, for two images with their own
    noise levels."""
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    L, n = 1024, 2
    print("case batch plan L=%d" % L, flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    timed_plan(api, shear, None, pyr=1, group=8, ct=8)
    timed_plan(api, shear, b, pyr=1, group=8, ct=8)
    sig = [7.5, 33.0]
    xs = np.stack([noisy_ellipses(L, sigma=s, seed=100 + i)[1] for i, s in enumerate(sig)])
    xg = to_gpu(xs)
    sc = api.nsst_scalars_cuda(shear, to_gpu(white_noise(L)))
    th = api.batch_thresholds(shear, sig, sc)
    out = torch.empty_like(xg)
    b.denoise(xg, None, th, out)
    for i in range(n):
        pd, _ = api.denoise_cuda(xg[i].contiguous(), shear, sig[i], sc, coeffs=False)
        torch.cuda.synchronize()
        assert np.array_equal(out[i].cpu().numpy(), pd.cpu().numpy()), i
    b.close()
    shear.close()
