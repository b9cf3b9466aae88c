"""CPU checks of the C ABI library and the host-side logic (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from oracle import nsst as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "shx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(shx3?_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from shearcuda_amd import _lib
    names = _declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(_lib.lib, n), n
    assert set(names) == set(_lib.EXPORTED)


def test_version_and_error_string():
    from shearcuda_amd import _lib
    assert _lib.lib.shx_version() >= 100
    assert isinstance(_lib.lib.shx_last_error(), bytes)


def test_maxflat_tables_match_oracle():
    import ctypes
    from shearcuda_amd import _lib
    for which, ref in enumerate(o.atrousfilters()):
        n = ctypes.c_int()
        _lib.check(_lib.lib.shx_maxflat_filter(which, None, ctypes.byref(n)))
        buf = (ctypes.c_double * (n.value * n.value))()
        _lib.check(_lib.lib.shx_maxflat_filter(which, buf, ctypes.byref(n)))
        got = np.frombuffer(buf, dtype=np.float64).reshape(n.value, n.value)
        assert np.array_equal(got, ref)


def test_separable_factors_exact():
    from shearcuda_amd import _lib
    assert _lib.lib.shx_maxflat_factor_error() < 1e-14


def test_bad_arguments_are_reported_without_gpu():
    import ctypes
    from shearcuda_amd import _lib
    h = ctypes.c_void_p()
    st = _lib.lib.shx_prepare_shearing_filters(300, 1, None, None, None, 0, ctypes.byref(h))
    assert st == -4 and b"not supported" in _lib.lib.shx_last_error()
    assert _lib.lib.shx_thresh(None, 0, None, None, 1.0, 2, None, 0, 0, None, 0, 0, None, None) == -1
    with pytest.raises(_lib.ShxError):
        _lib.check(_lib.lib.shx_atrousdec(None, 256, 4, None, None, None))


@pytest.mark.parametrize("n1,level", [(32, 3), (16, 4), (16, 2), (8, 2), (64, 5)])
def test_product_meyer_filters_match_oracle(n1, level):
    from shearcuda_amd.filters import shearing_filters_myer
    assert np.abs(shearing_filters_myer(n1, level) - o.shearing_filters_myer(n1, level)).max() < 1e-14


def test_band_thresholds_match_nsst_HT():
    from shearcuda_amd.api import band_thresholds

    class D:
        nscales = 3
        ndir = [2, 1, 2]
    sc = [2.0, np.array([1.0, 2.0]), np.array([3.0]), np.array([0.5, 0.25])]
    th = band_thresholds(D, 10.0, sc, (0.0, 3.0, 4.0))
    assert np.allclose(th, [0.0, 30.0, 60.0, 90.0, 20.0, 10.0])
    ref = o.nsst_thresholds(10.0, (0.0, 3.0, 4.0), sc)
    assert np.allclose(th, ref)
    # scalar tables of the wrong shape are rejected (the C ABI reads 1 + sum D
    # scalars and 3 T values from the host pointers it is given)
    for bad in (sc[:-1], [2.0, np.array([1.0]), sc[2], sc[3]], [np.array([1.0, 2.0])] + sc[1:]):
        with pytest.raises(ValueError):
            band_thresholds(D, 10.0, bad, (0.0, 3.0, 4.0))
    with pytest.raises(ValueError):
        band_thresholds(D, 10.0, sc, (0.0, 3.0))
