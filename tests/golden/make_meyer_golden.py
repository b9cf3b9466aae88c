"""Writes tests/golden/meyer_filters.npz: the oracle's FP64 spatial Meyer
shearing filters for the (dsize, dcomp) pairs of the 2D denoise path
(shearlet_toolbox_1/denoise_shearlet_example.m:26-29) and the small sizes the
adjoint checks use.  A regression pin: the values follow from
shearlet_toolbox_1/shearing_filters_Myer.m, which no MATLAB here can run, so
the other invariants in tests/test_oracle.py carry the parity argument."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import nsst as o  # noqa: E402

PAIRS = [(32, 3), (16, 4), (16, 2), (8, 2)]

if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "meyer_filters.npz"),
                        **{"w_%d_%d" % (n1, lev): o.shearing_filters_myer(n1, lev) for n1, lev in PAIRS})
    print("wrote meyer_filters.npz")
