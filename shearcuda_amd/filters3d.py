"""3D cone filters of the shearlet volume transform (host side, FP64).

Builds the spatial filters of shearcuda3d/3DShearTrans/GetMeyerBasedFilter.m:12-151
(with GeneratePyramidSection.m:1-14, GenerateXYZ.m:12-54, PolarToRec.m:10-19):

* a pseudo-polar grid of side n (24): polar line (i, j) of a cone runs through
  the points (k, y_i(k), z_j(k)), k = 1..n, with
  y_i(k) = round(i + (k-1)(n+1-2i)/(n-1)) (and z_j alike); the three cones
  take the radial coordinate along axis 1, 2 or 3 (the axis orders of
  GeneratePyramidSection.m);
* a direction window (l2, l1) of a cone puts weight 1 on the polar points of
  its line ranges, 1/2 on the lines it shares with a neighbouring window and
  1/3 on the cube's corner lines, where several points land on one voxel the
  last in the reference's loop order (j outer, i middle, k inner) wins;
* every window is divided by the sum A of all windows, so the windows form a
  partition of unity, and the filter is real(fftshift(ifftn(fftshift(W / A))));
* the first two cones are swapped at the end (GetMeyerBasedFilter.m:148-150).

A level's filters depend only on its direction count.  Band order within a
level: cone-major, then the MATLAB linear (column-major) index over the
(l2, l1) cell, the order of ShDecThRec_cuda.cpp:128-145,260-291.
"""
import numpy as np

FILTER_SIDE = 24  # GetMeyerBasedFilter.m:18 (the filterSize argument is ignored)


def polar_grid(n):
    """(X, Y, Z) 1-based coordinates, each (n, n, n) indexed [i-1, k-1, j-1]."""
    i = np.arange(1, n + 1, dtype=np.float64)[:, None, None]
    k = np.arange(1, n + 1, dtype=np.float64)[None, :, None]
    j = np.arange(1, n + 1, dtype=np.float64)[None, None, :]
    X = np.broadcast_to(k, (n, n, n)).astype(np.int64)
    Y = np.broadcast_to(np.floor(i + (k - 1) * (n + 1 - 2 * i) / (n - 1) + 0.5), (n, n, n)).astype(np.int64)
    Z = np.broadcast_to(np.floor(j + (k - 1) * (n + 1 - 2 * j) / (n - 1) + 0.5), (n, n, n)).astype(np.int64)
    return X, Y, Z


def cone_sections(n):
    X, Y, Z = polar_grid(n)
    return [(X, Y, Z), (Y, X, Z), (Z, Y, X)]


def _window_ranges(D, n):
    """For each (l2, l1): 1-based inclusive i-range, j-range and the weight
    tables along i and j (edges 1/2) plus the corner given 1/3."""
    s = n // D
    out = {}
    for l2 in range(1, D + 1):
        for l1 in range(1, D + 1):
            i0, i1 = (l1 - 1) * s + 1, l1 * s + (0 if l1 == D else 1)
            j0, j1 = (l2 - 1) * s + 1, l2 * s + (0 if l2 == D else 1)
            corner = None
            if l2 in (1, D) and l1 in (1, D):
                corner = (i0 if l1 == 1 else i1, j0 if l2 == 1 else j1)
            out[(l2, l1)] = (i0, i1, j0, j1, corner)
    return out


def _window(section, n, i0, i1, j0, j1, corner):
    X, Y, Z = section
    jj, ii, kk = np.meshgrid(np.arange(j0, j1 + 1), np.arange(i0, i1 + 1), np.arange(1, n + 1), indexing="ij")
    jj, ii, kk = jj.ravel(), ii.ravel(), kk.ravel()
    w = np.ones(ii.shape)
    w[(ii == i0) | (ii == i1) | (jj == j0) | (jj == j1)] = 0.5
    if corner is not None:
        w[(ii == corner[0]) & (jj == corner[1])] = 1.0 / 3.0
    vox = ((X[ii - 1, kk - 1, jj - 1] - 1) * n + Y[ii - 1, kk - 1, jj - 1] - 1) * n + Z[ii - 1, kk - 1, jj - 1] - 1
    # last write wins: keep the final occurrence of every voxel
    order = np.arange(vox.size)
    last = np.full(n ** 3, -1, dtype=np.int64)
    np.maximum.at(last, vox, order)
    hit = last >= 0
    block = np.zeros(n ** 3)
    block[hit] = w[last[hit]]
    return block.reshape(n, n, n)


def cone_filters(D, n=FILTER_SIDE):
    """Spatial filters of one direction count: array (3 * D * D, n, n, n) in
    band order (cone-major, then column-major over (l2, l1))."""
    sections = cone_sections(n)
    ranges = _window_ranges(D, n)
    W = np.zeros((3, D, D, n, n, n))
    for c in range(3):
        for (l2, l1), r in ranges.items():
            W[c, l2 - 1, l1 - 1] = _window(sections[c], n, *r)
    A = W.sum(axis=(0, 1, 2))
    if np.any(A == 0):
        raise ValueError("%d directions leave frequencies uncovered; use a direction count dividing %d" % (D, n))
    W /= A
    F = np.real(np.fft.fftshift(np.fft.ifftn(np.fft.fftshift(W, axes=(3, 4, 5)), axes=(3, 4, 5)), axes=(3, 4, 5)))
    F = F[[1, 0, 2]]  # swap cones 1 and 2
    # band order: cone, then l1 major / l2 minor (MATLAB linear index of (l2, l1))
    return np.ascontiguousarray(F.transpose(0, 2, 1, 3, 4, 5).reshape(3 * D * D, n, n, n))


def get_meyer_based_filter(num_dirs, n=FILTER_SIDE):
    """Per-level filter arrays for direction counts num_dirs (finest first).
    Levels with equal counts share one array object."""
    cache = {}
    out = []
    for D in num_dirs:
        if D not in cache:
            cache[D] = cone_filters(D, n)
        out.append(cache[D])
    return out
