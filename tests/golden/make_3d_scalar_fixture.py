"""Writes tests/golden/nsst3d_scalars_3884444.npz from the reference's own
Monte-Carlo noise-scalar fixture shearcuda3d/DenoiseDemo/nsstScalarsData3884444.mat
(DenoiseDemo/NsstScalars.m output for level 3, directions 8x8 / 4x4 / 4x4,
Pyr_mode 1.5).  The .mat file is MATLAB v5 data read with scipy.io.loadmat;
only the numbers are kept: band_est (3 cones x levels, MATLAB (l2, l1) arrays)
and low_pass_est.
"""
import os

import numpy as np
import scipy.io as sio

SRC = "/root/reference/shearcuda3d/DenoiseDemo/nsstScalarsData3884444.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nsst3d_scalars_3884444.npz")


def main():
    s = sio.loadmat(SRC)["nsstScalars"]
    be = s["bandEst"][0, 0]
    out = {"low_pass_est": np.float64(s["lowPassEst"][0, 0].ravel()[0])}
    for c in range(be.shape[0]):
        for l in range(be.shape[1]):
            out["band_est_c%d_l%d" % (c, l)] = np.asarray(be[c, l], dtype=np.float32)
    np.savez(OUT, **out)


if __name__ == "__main__":
    main()
