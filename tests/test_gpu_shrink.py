"""Fused soft-threshold shrink (shx_batch_shrink / _backward) against the
unfused composition of the same library calls: shear_trans, soft_bands,
inverse_shear forward; inverse_shear_adjoint, soft_bands_backward,
shear_trans_adjoint backward.  The composition itself is pinned to FP64
finite differences of the oracle in test_gpu_separation.py."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def thresholds(C, per_image, seed):
    """Per-band thresholds near each band's median |c|, so that about half of
    every band survives."""
    rng = np.random.default_rng(seed)
    nimg, nb = C.shape[:2]
    med = C.abs().reshape(nimg, nb, -1).median(dim=2).values.cpu().numpy()
    if not per_image:
        med = med.mean(0)
    return torch.tensor(med * rng.uniform(0.5, 1.5, size=med.shape), dtype=torch.float32, device="cuda")


@pytest.mark.parametrize("L,nimg,per_image,dct", [(64, 2, False, False), (64, 3, True, True),
                                                  (256, 2, False, True), (256, 2, True, False),
                                                  (512, 2, False, True), (1024, 1, True, True)])
def test_shrink_matches_unfused_composition(gpu, L, nimg, per_image, dct):
    print("case shrink L=%d nimg=%d per_image=%s dct=%s" % (L, nimg, per_image, dct), flush=True)
    from shearcuda_amd import api, separation as sep
    shear = api.make_dictionary(L)
    plan = api.Batch(shear, nimg)
    api.set_plan_options(shear, plan, dct_front=dct, dct_back=dct)
    rng = np.random.default_rng(L + nimg)
    x = torch.tensor(rng.uniform(0, 255, size=(nimg, L, L)), dtype=torch.float32, device="cuda")
    nb, pl = shear.nbands, L * L
    n = nb * pl
    C2 = plan.alloc_coeffs()
    plan.shear_trans(x, C2)
    th = thresholds(C2, per_image, L)
    # forward
    out = torch.empty_like(x)
    C = plan.alloc_coeffs()
    plan.shrink(x, th, out, C)
    thr = th if per_image else th[None, :].expand(nimg, nb).contiguous()
    Cs = torch.empty_like(C2)
    for i in range(nimg):  # one image per call: per-image rows of thresholds
        sep.soft_bands(C2[i:i + 1], thr[i].contiguous(), 1, n, n, pl, out=Cs[i:i + 1])
    out2 = torch.empty_like(x)
    plan.inverse_shear(Cs, out2)
    out3 = torch.empty_like(x)
    plan.shrink(x, th, out3)  # no coefficients kept
    torch.cuda.synchronize()
    kept = float((Cs != 0).float().mean())
    print("coeff rel %.2e out rel %.2e out(no coeff) rel %.2e kept %.2f" %
          (rel(C, C2), rel(out, out2), rel(out3, out2), kept), flush=True)
    assert 0.05 < kept < 0.95
    assert rel(C, C2) <= 1e-6
    assert rel(out, out2) <= 1e-5
    assert torch.equal(out, out3)
    # backward
    gout = torch.tensor(rng.standard_normal((nimg, L, L)), dtype=torch.float32, device="cuda")
    gx = torch.empty_like(x)
    gth = torch.zeros_like(th)
    plan.shrink_backward(C, th, gout, gx, gth)
    Gs = plan.alloc_coeffs()
    plan.inverse_shear_adjoint(gout, Gs)
    gC = torch.empty_like(Gs)
    gth2 = torch.zeros((nimg, nb), dtype=torch.float32, device="cuda")
    # the mask of the coefficients the backward is given (C2 agrees with C to
    # rounding, and a coefficient at its threshold can flip between them)
    for i in range(nimg):
        gC[i] = sep.soft_bands_backward(C[i:i + 1], Gs[i:i + 1], thr[i].contiguous(), gth2[i], 1, n, n, pl)[0]
    gx2 = torch.empty_like(x)
    plan.shear_trans_adjoint(gC, gx2)
    gth_ref = gth2 if per_image else gth2.sum(0)
    torch.cuda.synchronize()
    print("gx rel %.2e gth rel %.2e" % (rel(gx, gx2), rel(gth, gth_ref)), flush=True)
    assert rel(gx, gx2) <= 1e-5
    assert rel(gth, gth_ref) <= 1e-4
    # a second backward adds into gth
    plan.shrink_backward(C, th, gout, gx, gth)
    torch.cuda.synchronize()
    assert rel(gth, 2 * gth_ref) <= 1e-4
    plan.close()
    shear.close()


def test_shrink_rejects_bad_threshold_tables(gpu):
    from shearcuda_amd import api
    shear = api.make_dictionary(64)
    plan = api.Batch(shear, 2)
    x = torch.zeros((2, 64, 64), device="cuda")
    out = torch.empty_like(x)
    with pytest.raises(ValueError):
        plan.shrink(x, torch.ones(shear.nbands - 1, device="cuda"), out)
    with pytest.raises(ValueError):
        plan.shrink(x, torch.ones((3, shear.nbands), device="cuda"), out)
    plan.close()
    shear.close()


@pytest.mark.parametrize("dct", [False, True])
def test_shrink_identities_at_full_size(gpu, dct):
    """Size-independent properties at the adjoint benchmark's size (8 x
    512^2): with thresholds just below zero the shrink is S A = identity (the
    pyramid and the directional stage reconstruct perfectly) and its backward
    is (S A)^T = identity; with thresholds above every coefficient the output
    is exactly zero and the backward passes nothing.  (At a zero threshold the
    backward's rule 1[|c| > 0] drops the few coefficients that are exactly
    zero in FP32 -- two in this stack -- so the identity needs t < 0.)"""
    print("case shrink identities dct=%s" % dct, flush=True)
    from shearcuda_amd import api
    L, nimg = 512, 8
    shear = api.make_dictionary(L)
    plan = api.Batch(shear, nimg)
    api.set_plan_options(shear, plan, dct_front=dct, dct_back=dct)
    rng = np.random.default_rng(91)
    x = torch.tensor(rng.uniform(0, 255, size=(nimg, L, L)), dtype=torch.float32, device="cuda")
    g = torch.tensor(rng.standard_normal((nimg, L, L)), dtype=torch.float32, device="cuda")
    zero = torch.full((shear.nbands,), -1e-30, dtype=torch.float32, device="cuda")
    out = torch.empty_like(x)
    C = plan.alloc_coeffs()
    plan.shrink(x, zero, out, C)
    gx = torch.empty_like(x)
    gth = torch.zeros_like(zero)
    plan.shrink_backward(C, zero, g, gx, gth)
    big = torch.full((shear.nbands,), 1e30, dtype=torch.float32, device="cuda")
    out0 = torch.empty_like(x)
    plan.shrink(x, big, out0, C)
    gx0 = torch.empty_like(x)
    gth0 = torch.zeros_like(big)
    plan.shrink_backward(C, big, g, gx0, gth0)
    torch.cuda.synchronize()
    e_fwd, e_bwd = rel(out, x), rel(gx, g)
    print("S A x vs x %.2e; (S A)^T g vs g %.2e" % (e_fwd, e_bwd), flush=True)
    assert e_fwd <= 1e-5 and e_bwd <= 1e-5
    assert torch.count_nonzero(out0) == 0 and torch.count_nonzero(gx0) == 0
    assert torch.count_nonzero(gth0) == 0
    plan.close()
    shear.close()
