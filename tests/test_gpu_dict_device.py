"""Dictionaries whose spectra are built on the device (the default) against
dictionaries built by the FP64 host FFT, and against the FP64 oracle.

Both builders evaluate the same FP64 spectra (a direct transform of the
small-support filters on the device, an FFT on the host) and round them once
to FP32, so transforms through either dictionary agree to FP32 rounding.  The
oracle comparison pins the device builder on its own: every other GPU parity
test also runs on device-built dictionaries.
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
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("L,dcomp,dsize,shift", [
    (32, (3, 3), (16, 16), "none"),
    (64, None, None, "center"),
    (256, None, None, "none"),
    (256, (2, 3, 5), (32, 16, 16), "center"),
    (1024, None, None, "none"),
])
def test_device_and_host_spectra_agree(gpu, L, dcomp, dsize, shift):
    from shearcuda_amd import api
    print("case device spectra L=%d dcomp=%s shift=%s" % (L, dcomp, shift), flush=True)
    kw = {} if dcomp is None else {"dcomp": dcomp, "dsize": dsize}
    dev = api.make_dictionary(L, shift=shift, **kw)
    host = api.make_dictionary(L, shift=shift, spectra="host", **kw)
    try:
        x = _gpu(np.random.default_rng(L).standard_normal((L, L)) * 20.0)
        cd = api.shear_trans_cuda(x, "maxflat", dev)
        ch = api.shear_trans_cuda(x, "maxflat", host)
        worst = max(_rel(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(cd, ch))
        rd = api.inverse_shear_cuda(ch, "maxflat", dev).cpu().numpy()
        rh = api.inverse_shear_cuda(ch, "maxflat", host).cpu().numpy()
        er = _rel(rd, rh)
        print("L=%d bands max-rel %.2e, inverse max-rel %.2e" % (L, worst, er), flush=True)
        assert worst < 2e-6 and er < 2e-6
    finally:
        dev.close()
        host.close()


@pytest.mark.parametrize("L,shift", [(256, "none"), (128, "center")])
def test_device_spectra_match_oracle(gpu, L, shift):
    """Forward transform through a device-built dictionary against the FP64
    oracle (noise input, every band)."""
    from shearcuda_amd import api
    print("case device spectra vs oracle L=%d shift=%s" % (L, shift), flush=True)
    x = np.random.default_rng(7).standard_normal((L, L)) * 20.0
    H = o.make_dictionary(L, shift=shift)
    ref = o.nsst_dec2(x, H, strict=False) if L < 256 else o.nsst_dec2(x, H)
    shear = api.make_dictionary(L, shift=shift)
    try:
        got = api.shear_trans_cuda(_gpu(x), "maxflat", shear)
        for j, (g, r) in enumerate(zip(got, ref)):
            e = _rel(g.cpu().numpy(), r)
            assert e < TOL, (j, e)
    finally:
        shear.close()


def test_vanishing_spectra_rejected(gpu):
    """All-zero filters leave every bin without energy: both builders refuse
    the dictionary with the same error."""
    from shearcuda_amd import api
    from shearcuda_amd._lib import ShxError
    w = [np.zeros((4, 16, 16))]
    for where in ("device", "host"):
        with pytest.raises(ShxError, match="vanish"):
            api.prepare_shearing_filters_cuda(w, 64, spectra=where)
