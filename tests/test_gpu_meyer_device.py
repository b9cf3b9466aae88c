"""Spatial Meyer shearing filters and whole dictionaries built on the device
(shx_meyer_filters, shx_shearing_filters_myer) against the host builders:
the FP64 filters of shearcuda_amd/filters.py (the same construction as the
oracle's shearing_filters_myer, whose pseudo-polar geometry
tests/test_oracle_meyer_geometry.py pins) and the dictionaries made from
them, and against the FP64 oracle's decomposition."""
import numpy as np
import pytest
import torch

from oracle import nsst as o

pytestmark = pytest.mark.gpu

# (n1, level): the denoise dictionary's scales, the separation demo's 80 x 80
# filters, and small and many-direction cases
CASES = [(32, 3), (32, 4), (16, 4), (16, 3), (80, 3), (80, 4), (8, 2), (64, 5), (24, 3)]


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


@pytest.mark.parametrize("n1,level", CASES)
def test_device_filters_equal_host_filters(api, n1, level):
    from shearcuda_amd import filters
    print("case meyer filters n1=%d level=%d" % (n1, level), flush=True)
    got = api.meyer_filters_cuda(n1, level).cpu().numpy()
    ref = filters.shearing_filters_myer(n1, level)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("max rel %.3e" % err, flush=True)
    assert err < 1e-13
    ref_o = o.shearing_filters_myer(n1, level)
    assert np.abs(got - ref_o).max() / np.abs(ref_o).max() < 1e-13


def test_device_filters_reject_bad_levels(api):
    from shearcuda_amd._lib import ShxError
    with pytest.raises(ShxError):
        api.meyer_filters_cuda(12, 4)      # 2 * 12 is not a multiple of 16


@pytest.mark.parametrize("L,shift", [(256, "none"), (512, "center")])
def test_device_dictionary_matches_host_dictionary_and_oracle(api, L, shift):
    from shearcuda_amd.synth import noisy_ellipses
    print("case device dictionary L=%d shift=%s" % (L, shift), flush=True)
    _, x = noisy_ellipses(L, sigma=20.0, seed=7)
    xg = torch.from_numpy(x.astype(np.float32)).cuda()
    dev = api.shearing_filters_Myer([32, 32, 16, 16], [3, 3, 4, 4], L, shift=shift)
    host = api.shearing_filters_Myer([32, 32, 16, 16], [3, 3, 4, 4], L, shift=shift, filters="host")
    Cd = api.shear_trans_cuda(xg, "maxflat", dev)
    Ch = api.shear_trans_cuda(xg, "maxflat", host)
    ref = o.nsst_dec2(x, o.make_dictionary(L, shift=shift))
    worst = 0.0
    for j in range(len(ref)):
        g, h = Cd[j].cpu().numpy(), Ch[j].cpu().numpy()
        r3 = ref[j] if ref[j].ndim == 3 else ref[j][None]
        g3 = g.reshape(r3.shape)
        # the FP64 filters differ from the host ones by a few ulps, so the
        # FP32 spectra (and the transforms) agree to FP32 rounding
        assert np.abs(g - h).max() <= 1e-6 * np.abs(h).max(), j
        for k in range(r3.shape[0]):
            e = np.abs(g3[k] - r3[k]).max() / np.abs(r3[k]).max()
            worst = max(worst, e)
            assert e <= 1e-5, (j, k, e)
    print("device dictionary worst band max-rel vs oracle %.3e" % worst, flush=True)
    rt = api.inverse_shear_cuda(Cd, "maxflat", dev).cpu().numpy()
    assert np.abs(rt - x).max() / np.abs(x).max() <= 1e-5
    dev.close()
    host.close()
