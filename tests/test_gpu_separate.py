"""The shipped separation loop (separation.separate, shearcuda/separate.m) on
the GPU against the FP64 oracle restatement (oracle/separation.py).

Setup: the noise level, the starting threshold and the schedule must agree to
1e-5.  Iterations: each traced GPU iteration is replayed by the oracle from the
GPU's own input state (parts and delta), under the GPU's hard-threshold
decisions (its kept coefficients), as the thresholded-output rule asks; the
parts must then agree to 1e-5 of max |x| (the parts are pieces of the
re-weighted image x, which fixes their scale; early parts can be near zero),
and the number of decisions the oracle would have taken differently is
reported.
"""
import numpy as np
import pytest
import torch

from oracle import nsst as o
from oracle import separation as osep

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("L,dsize,level2,itermax", [(256, (80, 80, 80, 80), 4, 4), (128, (32, 32, 16, 16), 3, 3)])
def test_separate_matches_oracle_under_gpu_masks(gpu, L, dsize, level2, itermax):
    from shearcuda_amd import api, separation
    from shearcuda_amd.synth import separation_input
    print("case separate L=%d dsize=%s level2=%d itermax=%d" % (L, dsize, level2, itermax), flush=True)
    stop, gamma, coeff, dcomp = 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], (3, 3, 4, 4)
    strict = L >= 256
    _, _, img = separation_input(L, seed=9, sigma=10.0)
    img = img.astype(np.float32)
    shear = api.make_dictionary(L, dcomp=dcomp, dsize=dsize)
    H = o.make_dictionary(L, dcomp=dcomp, dsize=dsize)
    steps, info = [], {}

    def keep(d):
        steps.append({"delta": d["delta"], "p1_in": d["part1_in"].cpu().numpy(), "p2_in": d["part2_in"].cpu().numpy(),
                      "p1": d["part1"].cpu().numpy(), "p2": d["part2"].cpu().numpy(),
                      "keep_sh": [c.cpu().numpy() != 0 for c in d["coeff"]], "keep_w": d["yh"].cpu().numpy() != 0})

    try:
        C, P = separation.separate(torch.from_numpy(img).cuda(), level2, itermax, stop, gamma, coeff, shear,
                                   trace=keep, info=info)
    finally:
        shear.close()
    st = osep.separate_setup(img.astype(np.float64), level2, itermax, stop, coeff, H, strict=strict)
    for k in ("sigma", "deltamax", "lambda"):
        e = abs(info[k] / st[k] - 1)
        print("%s gpu %.9g oracle %.9g rel %.1e" % (k, info[k], st[k], e), flush=True)
        assert e < TOL, k
    assert len(steps) == itermax
    x = st["x"]
    scale = np.abs(x).max()
    total_flips = 0
    for i, s in enumerate(steps):
        p1, p2, flips = osep.separate_iter(x, s["p1_in"].astype(np.float64), s["p2_in"].astype(np.float64), s["delta"],
                                           st["E"], H, level2, gamma, s["keep_sh"], s["keep_w"], strict=strict)
        e1 = np.abs(s["p1"] - p1).max() / scale
        e2 = np.abs(s["p2"] - p2).max() / scale
        total_flips += flips
        print("iter %d: curve part %.2e, point part %.2e of max|x|, kept %d + %d, flips %d" %
              (i, e1, e2, sum(int(k.sum()) for k in s["keep_sh"]), int(s["keep_w"].sum()), flips), flush=True)
        assert e1 < TOL and e2 < TOL, (i, e1, e2)
    assert np.array_equal(C.cpu().numpy(), steps[-1]["p1"]) and np.array_equal(P.cpu().numpy(), steps[-1]["p2"])
    print("threshold decisions that differ from the oracle's own: %d" % total_flips, flush=True)


def test_hairline_crack_variant_matches_oracle(gpu):
    """separateHairlineCrack: a fixed starting threshold (param.deltaMax, sigma
    taken as 2.5), the Daubechies-4 wavelet, band norms passed in; same
    replay under the GPU's decisions."""
    from shearcuda_amd import api, separation
    from shearcuda_amd.synth import separation_input
    L, level2 = 128, 3
    param = {"numIter": 3, "stop": 3.0, "gamma": 2.0, "coeff": [0.1, 1.0, 1.5], "deltaMax": 300.0,
             "wavelet_name": "Daubechies", "wavelet_val": 4}
    print("case hairline crack variant L=%d" % L, flush=True)
    _, _, img = separation_input(L, seed=4, sigma=10.0)
    img = img.astype(np.float32)
    shear = api.make_dictionary(L)
    H = o.make_dictionary(L)
    steps, info = [], {}

    def keep(d):
        steps.append({"delta": d["delta"], "p1_in": d["part1_in"].cpu().numpy(), "p2_in": d["part2_in"].cpu().numpy(),
                      "p1": d["part1"].cpu().numpy(), "p2": d["part2"].cpu().numpy(),
                      "keep_sh": [c.cpu().numpy() != 0 for c in d["coeff"]], "keep_w": d["yh"].cpu().numpy() != 0})

    try:
        E = api.com_norm("maxflat", (L, L), shear)
        C, P, residual, Ct = separation.separate_hairline_crack(torch.from_numpy(img).cuda(), level2, param, shear,
                                                                E=E, trace=keep, info=info)
        Ct = [c.cpu().numpy() for c in Ct]
    finally:
        shear.close()
    wav = ("Daubechies", 4)
    st = osep.separate_setup(img.astype(np.float64), level2, 3, 3.0, param["coeff"], H, strict=False, wavelet=wav,
                             delta_max=300.0)
    assert info["sigma"] == 2.5 and info["deltamax"] == 300.0
    assert abs(info["lambda"] / st["lambda"] - 1) < 1e-12
    x = st["x"]
    scale = np.abs(x).max()
    for i, s in enumerate(steps):
        p1, p2, flips = osep.separate_iter(x, s["p1_in"].astype(np.float64), s["p2_in"].astype(np.float64), s["delta"],
                                           st["E"], H, level2, param["gamma"], s["keep_sh"], s["keep_w"],
                                           strict=False, wavelet=wav)
        e1 = np.abs(s["p1"] - p1).max() / scale
        e2 = np.abs(s["p2"] - p2).max() / scale
        print("iter %d: %.2e %.2e flips %d" % (i, e1, e2, flips), flush=True)
        assert e1 < TOL and e2 < TOL, (i, e1, e2)
    last = steps[-1]
    assert np.array_equal(C.cpu().numpy(), last["p1"]) and np.array_equal(P.cpu().numpy(), last["p2"])
    # residual of the last iteration and its thresholded cells
    res_o = x - (last["p1_in"].astype(np.float64) + last["p2_in"].astype(np.float64))
    assert np.abs(residual.cpu().numpy() - res_o).max() / scale < TOL
    assert all(np.array_equal(a != 0, b) for a, b in zip(Ct, last["keep_sh"]))


@pytest.mark.parametrize("variant", ["separate", "hairline"])
def test_native_loop_equals_python_loop(gpu, variant):
    """shx_separate (the loop inside the library) against the Python mirror:
    the same calls in the same order, so the parts are bit-identical when
    both use the same band norms; the norms it computes itself match
    com_norm's to FP32 reduction rounding."""
    from shearcuda_amd import api, separation
    from shearcuda_amd.synth import separation_input
    L = 256
    print("case native separation loop %s L=%d" % (variant, L), flush=True)
    _, _, img = separation_input(L, seed=6, sigma=10.0)
    img = torch.from_numpy(img.astype(np.float32)).cuda()
    shear = api.make_dictionary(L, dcomp=(3, 3, 4, 4), dsize=(32, 32, 32, 32))
    try:
        E = api.com_norm("maxflat", (L, L), shear)
        ip, inn = {}, {}
        if variant == "separate":
            Cp, Pp = separation.separate(img, 3, 4, 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], shear, info=ip)
            Cn, Pn = separation.separate_native(img, 3, 4, 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], shear, E=E, info=inn)
        else:
            param = {"numIter": 3, "stop": 3.0, "gamma": 2.0, "coeff": [0.1, 1.0, 1.5], "deltaMax": 300.0,
                     "wavelet_name": "Daubechies", "wavelet_val": 4}
            Cp, Pp, _, _ = separation.separate_hairline_crack(img, 3, param, shear, E=E, info=ip)
            Cn, Pn = separation.separate_native(img, 3, 3, 3.0, 2.0, param["coeff"], shear, wavelet=("Daubechies", 4),
                                                delta_max=300.0, E=E, info=inn)
        for k in ("sigma", "deltamax", "lambda"):
            assert ip[k] == inn[k], (k, ip[k], inn[k])
        assert np.array_equal(Cp.cpu().numpy(), Cn.cpu().numpy())
        assert np.array_equal(Pp.cpu().numpy(), Pn.cpu().numpy())
        print("%s: bit-identical with shared norms" % variant, flush=True)
        if variant == "separate":
            # norms computed inside the library (one FP64 reduction per band,
            # where com_norm sums FP32 partials): the norms differ by rounding,
            # so at most a rare threshold decision can differ
            Cs, Ps = separation.separate_native(img, 3, 4, 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], shear)
            e = max(np.linalg.norm(Cs.cpu().numpy() - Cp.cpu().numpy()) / np.linalg.norm(Cp.cpu().numpy()),
                    np.linalg.norm(Ps.cpu().numpy() - Pp.cpu().numpy()) / np.linalg.norm(Pp.cpu().numpy()))
            print("own norms: relative l2 diff %.2e" % e, flush=True)
            assert e < 1e-3
    finally:
        shear.close()
