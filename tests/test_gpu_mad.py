"""Device median absolute deviation (shx_band_mad, shx_nsst_scalars) against
numpy's median on the same FP32 data, bit for bit.

numpy's median is the library routine both conventions reduce to:
  fp64  values widened to double, m = median(v), MAD = median(|v - m|) / 0.6745
        (nsst_scalars.m:36-44 on a double transform);
  fp32  the same in float arithmetic (numpy's float32 median averages the two
        middle values in float32, as demo_denoise/shearcuda/shearcuda.cpp:165-181
        does), the final division in double.
Cases cover odd and even counts, one and two elements, ties around the middle,
constant planes, signed zeros, tiny values and both load paths (plane stride a
multiple of four or not).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mad64(v):
    v = np.asarray(v, dtype=np.float32).astype(np.float64).ravel()
    return np.median(np.abs(v - np.median(v))) / 0.6745


def _mad32(v):
    v = np.asarray(v, dtype=np.float32).ravel()
    m = np.median(v)
    return float(np.median(np.abs(v - m))) / 0.6745


def _planes(n, kind, nplanes, seed):
    rng = np.random.default_rng(seed)
    if kind == "normal":
        a = rng.standard_normal((nplanes, n)) * 3.0
    elif kind == "ties":
        a = np.round(rng.standard_normal((nplanes, n)) * 2.0)
    elif kind == "const":
        a = np.full((nplanes, n), -1.25)
    elif kind == "zeros":
        a = np.where(rng.random((nplanes, n)) < 0.5, -0.0, 0.0)
    elif kind == "tiny":
        a = rng.standard_normal((nplanes, n)) * 1e-30
    else:  # wide: magnitudes over many binades, both signs
        a = rng.standard_normal((nplanes, n)) * 10.0 ** rng.integers(-20, 20, (nplanes, n))
    return a.astype(np.float32)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 1001, 4096, 65539, 262144])
@pytest.mark.parametrize("kind", ["normal", "ties", "const", "zeros", "tiny", "wide"])
def test_band_mad_matches_numpy(gpu, n, kind):
    from shearcuda_amd import api
    nplanes = 3 if n < 100000 else 2
    a = _planes(n, kind, nplanes, n * 7 + len(kind))
    t = torch.from_numpy(a).cuda()
    for arith, ref in (("fp64", _mad64), ("fp32", _mad32)):
        got = api.band_mad_cuda(t, arith).cpu().numpy()
        want = np.array([ref(a[i]) for i in range(nplanes)])
        assert np.array_equal(got, want), (n, kind, arith, got, want)


def test_band_mad_unaligned_planes(gpu):
    """A stack whose planes start off 16-byte boundaries (odd plane length)
    and a view that starts one element into its allocation."""
    from shearcuda_amd import api
    a = _planes(4097, "normal", 5, 3)
    base = torch.from_numpy(np.concatenate([[0.0], a.ravel()]).astype(np.float32)).cuda()
    t = base[1:].view(5, 4097)
    got = api.band_mad_cuda(t).cpu().numpy()
    assert np.array_equal(got, np.array([_mad64(r) for r in a]))


@pytest.mark.parametrize("L", [64, 256, 512, 1024])
def test_nsst_scalars_on_device(gpu, L):
    """shx_nsst_scalars: the transform of the noise image, then exact medians,
    equal to numpy on the same coefficients; close to the FP64 oracle's."""
    from oracle import nsst as o
    from shearcuda_amd import api
    from shearcuda_amd.synth import white_noise
    print("case nsst scalars L=%d" % L, flush=True)
    shear = api.make_dictionary(L)
    try:
        w = white_noise(L)
        wg = torch.from_numpy(w.astype(np.float32)).cuda()
        C = api.shear_trans_cuda(wg, "maxflat", shear)
        cells = [C[0].cpu().numpy()[None]] + [c.cpu().numpy() for c in C[1:]]
        for arith, ref in (("fp64", _mad64), ("fp32", _mad32)):
            sc = api.nsst_scalars_cuda(shear, wg, arith)
            got = np.concatenate([[sc[0]]] + [np.asarray(s) for s in sc[1:]])
            want = np.array([ref(p) for c in cells for p in c])
            assert np.array_equal(got, want), arith
        if L == 256:
            so = o.nsst_scalars(o.make_dictionary(L), w)
            ref = np.concatenate([[so[0]]] + [np.asarray(s) for s in so[1:]])
            sc = api.nsst_scalars_cuda(shear, wg)
            got = np.concatenate([[sc[0]]] + [np.asarray(s) for s in sc[1:]])
            err = np.abs(got / ref - 1).max()
            print("L=%d device scalars vs FP64 oracle: max rel %.2e" % (L, err), flush=True)
            assert err < 1e-5
    finally:
        shear.close()


def test_wavelet_scalars_on_device(gpu):
    from oracle import rdwt as w
    from shearcuda_amd import api, separation
    L, levels = 128, 3
    h = w.make_on_filter("Daubechies", 4)
    noise = torch.from_numpy(np.random.default_rng(2).standard_normal((L, L)).astype(np.float32)).cuda()
    got = separation.wavelet_scalars(noise, h, levels)
    yl, yh = api.mrdwt_cuda(noise, h, levels)
    yl, yh = yl.cpu().numpy(), yh.cpu().numpy().reshape(L, 3 * levels, L)
    want = np.array([_mad64(yl)] + [_mad64(yh[:, k, :]) for k in range(3 * levels)])
    assert np.array_equal(got, want)


def test_scalars_3d_on_device(gpu):
    from shearcuda_amd import shearlet3d
    n = 32
    sh = shearlet3d.Shear3D(n, (8, 8, 4), 1)
    noise = torch.from_numpy(np.random.default_rng(5).standard_normal((n, n, n)).astype(np.float32)).cuda()
    est, low = shearlet3d.nsst_scalars_3d(sh, noise)
    plan = shearlet3d.Plan3D(sh, 1, group=8)
    try:
        subs = plan.pyr_dec(noise.reshape(1, n, n, n).contiguous())
        bands = plan.analysis(subs)
        want = [_mad64(b) for lvl in sh.split_bands(bands) for b in lvl[0].cpu().numpy()]
        want_low = _mad64(sh.split_subs(subs)[-1][0].cpu().numpy())
    finally:
        plan.close()
    assert np.array_equal(est, np.array(want))
    assert low == want_low
