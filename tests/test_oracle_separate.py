"""Pins of the oracle's restatement of the shipped separation loop
(oracle/separation.py, shearcuda/separate.m) on the CPU."""
import numpy as np

from oracle import nsst as o
from oracle import separation as osep


def _case(L=64, seed=2):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((L, L)) * 10.0 + 50.0
    H = o.make_dictionary(L, dcomp=(2, 3), dsize=(16, 16))
    return img, H


def test_schedule_reaches_stop_at_last_iteration():
    """deltamax * lambda^(itermax-1) = stop * sigma (separate.m:135): the
    exponential schedule ends on the stopping threshold."""
    img, H = _case()
    itermax, stop = 6, 3.0
    st = osep.separate_setup(img, 2, itermax, stop, [0.1, 1.5, 1.5], H, strict=False)
    assert st["deltamax"] == min(st["buff"])
    end = st["deltamax"] * st["lambda"] ** (itermax - 1)
    assert abs(end / (stop * st["sigma"]) - 1) < 1e-12


def test_iteration_with_own_masks_has_no_flips():
    img, H = _case()
    st = osep.separate_setup(img, 2, 4, 3.0, [0.1, 1.5, 1.5], H, strict=False)
    x = st["x"]
    z = np.zeros_like(x)
    p1, p2, flips = osep.separate_iter(x, z, z, st["deltamax"], st["E"], H, 2, 3.0, strict=False)
    assert flips == 0
    # the first iteration at deltamax keeps nothing above the largest
    # normalised coefficient of either dictionary
    assert np.isfinite(p1).all() and np.isfinite(p2).all()


def test_forced_masks_are_counted_and_used():
    """Passing all-false masks zeroes every coefficient: the flips equal the
    decisions that were keeps, and the wavelet part reduces to the lowpass."""
    from oracle import rdwt as w
    img, H = _case()
    st = osep.separate_setup(img, 2, 4, 3.0, [0.1, 1.5, 1.5], H, strict=False)
    x = st["x"]
    z = np.zeros_like(x)
    delta = 0.5 * st["sigma"]  # low enough that both thresholds keep coefficients
    _, _, own = osep.separate_iter(x, z, z, delta, st["E"], H, 2, 3.0, strict=False)
    assert own == 0
    Ct = o.nsst_dec2(x, H, strict=False)
    keep_sh = [np.zeros(c.shape, bool) for c in Ct]
    qmf = w.make_on_filter("Symmlet", 4)
    yl, yh = w.mrdwt(x, qmf, 2)
    keep_w = np.zeros(yh.shape, bool)
    p1, p2, flips = osep.separate_iter(x, z, z, delta, st["E"], H, 2, 3.0, keep_sh, keep_w, strict=False)
    n_keep = sum(int(np.count_nonzero(np.abs(c if c.ndim == 3 else c[None])
                                      > 1.4 * delta * st["E"][s, :(c.shape[0] if c.ndim == 3 else 1), None, None]))
                 for s, c in enumerate(Ct)) + int(np.count_nonzero(np.abs(yh) > 1.4 * delta))
    assert flips == n_keep > 0
    assert np.allclose(p1, 0.0, atol=1e-12)
    assert np.allclose(p2, w.mirdwt(yl, np.zeros_like(yh), qmf, 2), atol=1e-12)
