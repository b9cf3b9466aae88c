"""Orthonormal wavelet filters for the undecimated wavelet transform.

Host-side mirror of shearcuda/MakeONFilter.m (Haar :55-57, Daubechies 4 and 6
:111-119, Symmlet 4 :174-177), normalised to unit l2 norm as :270 does.
"""
import numpy as np

_TAPS = {
    ("Haar", 0): [1.0, 1.0],
    ("Daubechies", 4): [.482962913145, .836516303738, .224143868042, -.129409522551],
    ("Daubechies", 6): [.332670552950, .806891509311, .459877502118, -.135011020010, -.085441273882,
                        .035226291882],
    ("Symmlet", 4): [-.107148901418, -.041910965125, .703739068656, 1.136658243408, .421234534204,
                     -.140317624179, -.017824701442, .045570345896],
}


def MakeONFilter(name, par=0):
    """qmf = MakeONFilter(Type, Par) for the filters listed above."""
    key = (name, 0 if name == "Haar" else int(par))
    if key not in _TAPS:
        raise ValueError("MakeONFilter: %s %s is not provided" % (name, par))
    f = np.array(_TAPS[key], dtype=np.float64)
    return f / np.sqrt(np.sum(f * f))
