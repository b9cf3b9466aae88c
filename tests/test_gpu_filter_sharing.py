"""Scales with identical spatial filters share one stored set of spectra
(shearlet_toolbox_1/denoise_shearlet_example.m:26-38 uses dsize [32 32 16 16]
with dcomp [3 3 4 4], so scales 1-2 and 3-4 have the same filters).  The
shared dictionary must give the results of one that stores every scale's
spectra (made here by changing one filter value of scale 2 by one ulp, which
keeps it from being shared), and hold half the spectra."""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


@pytest.mark.parametrize("L", [256, 512])
def test_shared_spectra_equal_separate_spectra(api, L):
    from shearcuda_amd import filters
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    print("case shared filter spectra L=%d" % L, flush=True)
    w = filters.dictionary_filters()
    w_sep = [a.copy() for a in w]
    w_sep[1][3, 5, 7] = np.nextafter(w_sep[1][3, 5, 7], 1.0)
    shared = api.prepare_shearing_filters_cuda(w, L)
    sep = api.prepare_shearing_filters_cuda(w_sep, L)
    N = L // 2
    plane = 8 * (L * N + L)          # one direction's packed spectrum and beta vector
    assert sep.nbytes() - shared.nbytes() == 8 * plane   # scale 2 stored apart (4 still shares 3's)
    _, x = noisy_ellipses(L, sigma=20.0, seed=21)
    xg = torch.from_numpy(x.astype(np.float32)).cuda()
    Cs = api.shear_trans_cuda(xg, "maxflat", shared)
    Cp = api.shear_trans_cuda(xg, "maxflat", sep)
    ref = o.nsst_dec2(x, o.make_dictionary(L))
    for j in range(len(ref)):
        a, b = Cs[j].cpu().numpy(), Cp[j].cpu().numpy()
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), j
        r = ref[j]
        assert np.abs(a - r).max() <= 1e-5 * np.abs(r).max(), j
    sc = api.nsst_scalars_cuda(shared, torch.from_numpy(white_noise(L).astype(np.float32)).cuda())
    for dct in (False, True):
        for s in (shared, sep):
            api.set_plan_options(s, dct_front=dct, dct_back=dct)
        out_s, _ = api.denoise_cuda(xg, shared, 20.0, sc, coeffs=False)
        out_p, _ = api.denoise_cuda(xg, sep, 20.0, sc, coeffs=False)
        d = (out_s - out_p).abs().max().item() / out_p.abs().max().item()
        print("dct=%s shared vs separate denoise max-rel %.3e" % (dct, d), flush=True)
        assert d <= 1e-4      # allows a rare threshold decision flip from the one-ulp filter change
    # batched plan (per-scale streams with their own sub-maps) = plain, bit for bit
    b = api.Batch(shared, 2)
    api.set_plan_options(shared, b, dct_front=True, dct_back=True)   # as the plain plan above
    xs = torch.stack([xg, xg.flip(0).contiguous()])
    th = api.batch_thresholds(shared, [20.0, 10.0], sc)
    out = torch.empty_like(xs)
    b.denoise(xs, None, th, out)
    for i, sig in enumerate((20.0, 10.0)):
        pd, _ = api.denoise_cuda(xs[i].contiguous(), shared, sig, sc, coeffs=False)
        torch.cuda.synchronize()
        assert torch.equal(out[i], pd), i
    b.close()
    shared.close()
    sep.close()
