"""Every schedule of a 2D plan computes the same transform.

The autotuner picks, per plan, the pyramid tile shape (3 variants), the
directions summed per reconstruction block (2, 4, 8) and the column tile width
of the decomposition and reconstruction column passes (4, 8, 16 where the size
allows; -4 and -8 at 512: the three-pass column FFT with 4 or 8 columns).  The default schedule is pinned to the FP64 oracle in
test_gpu_nsst.py; here every other schedule is compared with it: the
coefficients must agree to FP32 rounding (they do not depend on the direction
grouping) and so must the denoised image (the grouping only reorders the sum
over directions).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / np.abs(b).max()


def _tiles(api, shear):
    from shearcuda_amd._lib import ShxError
    ok = []
    for ct in (4, 8, 16, -4, -8, -2):   # negative: the three-pass engines
        try:
            api.set_schedule(shear, 0, 4, col_tile_dec=ct, col_tile_rec=ct)
            ok.append(ct)
        except ShxError:
            pass
    return ok


def _bands(cells, L):
    """Directional coefficient cells (lowpass excluded) as one (n, L, L) array."""
    return np.concatenate([t.cpu().numpy().reshape(-1, L, L) for t in cells[1:]])


@pytest.mark.parametrize("L", [128, 256, 512])
def test_every_schedule_matches_default(gpu, L):
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    print("case schedules L=%d" % L, flush=True)
    shear = api.make_dictionary(L)
    try:
        sc = api.nsst_scalars_cuda(shear, _gpu(white_noise(L)))
        _, x = noisy_ellipses(L, sigma=20.0, seed=11)
        xg = _gpu(x)
        tiles = _tiles(api, shear)
        assert 8 in tiles and (L < 128 or 4 in tiles)
        api.set_schedule(shear, 0, 4, col_tile_dec=0, col_tile_rec=0)
        ref, cref = api.denoise_cuda(xg, shear, 20.0, sc)
        ref = ref.cpu().numpy()
        cref = _bands(cref, L)
        worst_x = worst_c = 0.0
        # the three-pass engine (negative tiles) is independent of the
        # pyramid tile shape: it is crossed with the first pyramid variant
        combos = [(pv, g, ctd, ctr) for pv in range(3) for g in (2, 4, 8) for ctd in tiles for ctr in tiles
                  if pv == 0 or (ctd > 0 and ctr > 0)]
        for pv, g, ctd, ctr in combos:
            api.set_schedule(shear, pv, g, col_tile_dec=ctd, col_tile_rec=ctr)
            out, c = api.denoise_cuda(xg, shear, 20.0, sc)
            ex, ec = _rel(out.cpu().numpy(), ref), _rel(_bands(c, L), cref)
            worst_x, worst_c = max(worst_x, ex), max(worst_c, ec)
            assert ex <= TOL and ec <= TOL, (pv, g, ctd, ctr, ex, ec)
        print("L=%d tiles %s: worst rel diff image %.2e coefficients %.2e" % (L, tiles, worst_x, worst_c), flush=True)
    finally:
        shear.close()


def _flip_aware_compare(out, ref, c, cref, th):
    """Images and coefficient stacks ([n][1 + nbands][L][L], plane 0 the
    lowpass) of two schedules whose keep decisions |c| > t (th [n][1 +
    nbands]) are taken on coefficients that agree only to FP32 rounding.
    A coefficient within that rounding of its threshold can be kept by one
    schedule and dropped by the other; such a flip moves the image by at most
    the coefficient's magnitude times the peak of its synthesis atom (< 1.5).
    Checks: coefficients within TOL; every flip within the coefficients'
    rounding of its threshold; few flips; the image difference within TOL
    plus the flipped coefficients' bound.  Returns (image rel, coeff rel,
    flips)."""
    ec = _rel(c, cref)
    assert ec <= TOL, ec
    scale = np.abs(cref).max()
    t = th[:, :, None, None]
    flip = (np.abs(c) > t) != (np.abs(cref) > t)
    nflip = int(flip.sum())
    assert nflip <= 10, nflip
    if nflip:
        gap = np.abs(np.abs(cref[flip]) - np.broadcast_to(t, cref.shape)[flip])
        assert gap.max() <= 2 * TOL * scale, (gap.max(), scale)
    budget = np.zeros(out.shape[0])
    for i in range(out.shape[0]):
        budget[i] = 1.5 * np.abs(cref[i][flip[i]]).sum()
    d = np.abs(np.asarray(out, np.float64) - np.asarray(ref, np.float64)).reshape(out.shape[0], -1).max(1)
    assert np.all(d <= budget + TOL * np.abs(ref).max()), (d, budget)
    return _rel(out, ref), ec, nflip


@pytest.mark.parametrize("L,cts", [(256, (8, 4)), (512, (8, -8, -4)), (1024, (8, -2, -4))])
def test_narrow_tiles_in_batched_plan(gpu, L, cts):
    """The batched plan (image index as a grid dimension) with the 4-column
    tiles (and the three-pass engines at 512 and 1024) against the same images
    through the 8-column schedule, keep decisions compared as in
    _flip_aware_compare."""
    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    n = 3
    print("case batched tiles L=%d n=%d %s" % (L, n, cts), flush=True)
    shear = api.make_dictionary(L)
    b = api.Batch(shear, n)
    try:
        sc = api.nsst_scalars_cuda(shear, _gpu(white_noise(L)))
        sig = np.array([10.0, 20.0, 30.0])
        xs = _gpu(np.stack([noisy_ellipses(L, sigma=float(s), seed=40 + i)[1] for i, s in enumerate(sig)]))
        th = api.batch_thresholds(shear, sig, sc)
        outs, coefs = [], []
        for ct in cts:
            api.set_schedule(shear, 0, 4, batch=b, col_tile_dec=ct, col_tile_rec=ct)
            out = torch.empty_like(xs)
            C = b.alloc_coeffs()
            b.denoise(xs, C, th, out)
            outs.append(out.cpu().numpy())
            coefs.append(C.cpu().numpy())
        thn = th.cpu().numpy()
        for ct, o, c in zip(cts[1:], outs[1:], coefs[1:]):
            e, ec, nflip = _flip_aware_compare(o, outs[0], c, coefs[0], thn)
            print("batched tile %d vs 8-column rel diff %.2e coefficients %.2e, %d keep flips at the threshold"
                  % (ct, e, ec, nflip), flush=True)
    finally:
        b.close()
        shear.close()
