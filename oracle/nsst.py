"""FP64 CPU oracle for the 2D nonsubsampled shearlet transform (NSST).

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP library is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``shearcuda_amd``) never does.

It restates the reference's CPU toolbox ``shearlet_toolbox_1`` in numpy, in
FP64, one function per MATLAB file, each citing the file:line it follows:

* ``atrousfilters``      shearlet_toolbox_1/atrousfilters.m:106-175 (maxflat)
* ``symext``             shearlet_toolbox_1/symext.m:18-31
* ``upsample2df``        shearlet_toolbox_1/upsample2df.m:4-6
* ``atrousc``            shearlet_toolbox_1/atrousc.cpp:43-73,116-122
* ``conv2_valid``        MATLAB ``conv2(..,'valid')`` as used in atrousdec.m:34-35
* ``atrousdec``          shearlet_toolbox_1/atrousdec.m:19-51
* ``atrousrec``          shearlet_toolbox_1/atrousrec.m:18-41
* ``meyer_wind``         shearlet_toolbox_1/meyer_wind.m:11-19
* ``windowing``          shearlet_toolbox_1/windowing.m:11-27
* ``gen_x_y_cordinates`` shearlet_toolbox_1/gen_x_y_cordinates.m:22-111
* ``rec_from_pol``       shearlet_toolbox_1/rec_from_pol.m:16-35
* ``shearing_filters_myer`` shearlet_toolbox_1/shearing_filters_Myer.m:18-27
* ``shear_spectra``      shearlet_toolbox_1/nsst_dec2.m:47-58 (+ the circular
                         n1/2 shift of shearcuda/src_cuda/prepare_shearing_filters_cuda.cpp:254
                         with shearcuda/shear_cuda.cu:2303-2305 when ``shift='center'``)
* ``conv2p``             shearlet_toolbox_1/conv2p.m:12
* ``nsst_dec2``          shearlet_toolbox_1/nsst_dec2.m:37-60
* ``nsst_rec2``          shearlet_toolbox_1/nsst_rec2.m:25-35
* ``nsst_scalars``       shearlet_toolbox_1/nsst_scalars.m:28-45 (noise passed in)
* ``nsst_HT``            shearlet_toolbox_1/nsst_HT.m:32-47
* ``thresh``             shearcuda/thresh.m:28-55
* ``soft_threshold``     shearcuda/SoftThresh.m:11-13
* ``re_weight_band``     shearcuda/re_weight_band.m:23-29
* ``com_norm``           shearcuda/com_norm.m:25-55

How each function is pinned (no MATLAB or Octave exists here, and the
reference's ``atrousc.cpp`` needs the MATLAB ``mex.h`` header, which the image
lacks, so it is not built):

* maxflat tables: perfect-reconstruction identity conv2(h0,g0)+conv2(h1,g1)=delta
  (atrousfilters.m:16-18) to 1e-15, and 8-fold symmetry.
* symext: brute-force index checks and the closed-form half-sample reflection.
* atrousc / conv2_valid: scipy.signal.convolve2d(.., 'valid') with the
  upsampled filter.
* atrousdec/atrousrec: reconstruction error of the quasi-tight pair recorded in
  the tests (PR through the symmetric boundary) and linearity.
* Meyer filters: hand-computed entries of the window, partition of unity of the
  squared Meyer windows, realness, and committed checksums.
* nsst_dec2/nsst_rec2: tight-frame round trip rec(dec(x)) = x, the adjoint
  identity, and numpy.fft as the FFT.
* nsst_HT / thresh / soft_threshold: closed forms on small arrays.
"""
import numpy as np

# --- maxflat quadrant tables: shearlet_toolbox_1/atrousfilters.m:109-164 -----
# Numeric data that has to be identical to the reference's filter design.
_H0_QUADRANT = np.array([
    [-7.900496718847182e-07, 0.0, 1.4220894093924927e-05, 2.5281589500310983e-05, -4.9773129328737247e-05, -0.00022753430550279883, -0.00033182086219158167],
    [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    [1.4220894093924927e-05, 0.0, -0.0002559760936906487, -0.00045506861100559767, 0.0008959163279172705, 0.004095617499050379, 0.00597277551944847],
    [2.5281589500310983e-05, 0.0, -0.00045506861100559767, 0.0009765625, 0.0015927401385195919, -0.0087890625, -0.01795090623402861],
    [-4.9773129328737247e-05, 0.0, 0.0008959163279172705, 0.0015927401385195919, -0.0031357071477104465, -0.014334661246676327, -0.020904714318069645],
    [-0.00022753430550279883, 0.0, 0.004095617499050379, -0.0087890625, -0.014334661246676327, 0.0791015625, 0.16155815610625748],
    [-0.00033182086219158167, 0.0, 0.00597277551944847, -0.01795090623402861, -0.020904714318069645, 0.16155815610625748, 0.3177420190660832],
])

_H1_QUADRANT = np.array([
    [6.391587676622346e-10, 0.0, -1.7257286726880333e-08, -3.067962084778726e-08, 1.3805829381504267e-07, 5.522331752601707e-07, 3.3747582932565985e-07, -1.9328161134105974e-06, -5.6949046198705095e-06, -7.649452131381623e-06],
    [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    [-1.7257286726880333e-08, 0.0, 4.65946741625769e-07, 8.283497628902559e-07, -3.727573933006152e-06, -1.4910295732024608e-05, -9.111847391792816e-06, 5.2186035062086126e-05, 0.00015376242473650378, 0.00020653520754730382],
    [-3.067962084778726e-08, 0.0, 8.283497628902559e-07, -2.9917573832012203e-07, -6.6267981031220475e-06, 5.3851632897621965e-06, 4.049868144081346e-05, -1.884807151416769e-05, -0.00023692226948222173, -0.0003769812640795245],
    [1.3805829381504267e-07, 0.0, -3.727573933006152e-06, -6.6267981031220475e-06, 2.9820591464049215e-05, 0.00011928236585619686, 7.289477913434253e-05, -0.000417488280496689, -0.0012300993978920302, -0.0016522816603784306],
    [5.522331752601707e-07, 0.0, -1.4910295732024608e-05, 5.3851632897621965e-06, 0.00011928236585619686, -9.693293921571956e-05, -0.0007289762659346422, 0.00033926528725501844, 0.004264600850679991, 0.006785662753431441],
    [3.3747582932565985e-07, 0.0, -9.111847391792816e-06, 4.049868144081346e-05, 7.289477913434253e-05, -0.0007289762659346422, -0.001468581806076247, 0.002551416930771248, 0.01181401175635013, 0.017093222023136675],
    [-1.9328161134105974e-06, 0.0, 5.2186035062086126e-05, -1.884807151416769e-05, -0.000417488280496689, 0.00033926528725501844, 0.002551416930771248, -0.0011874285053925643, -0.01492610297737997, -0.023749819637010044],
    [-5.6949046198705095e-06, 0.0, 0.00015376242473650378, -0.00023692226948222173, -0.0012300993978920302, 0.004264600850679991, 0.01181401175635013, -0.01492610297737997, -0.0826466923977296, -0.12203257624594532],
    [-7.649452131381623e-06, 0.0, 0.00020653520754730382, -0.0003769812640795245, -0.0016522816603784306, 0.006785662753431441, 0.017093222023136675, -0.023749819637010044, -0.12203257624594532, 0.821896776039774],
])

_G0_QUADRANT = np.array([
    [-6.391587676622346e-10, 0.0, 1.7257286726880333e-08, 3.067962084778726e-08, -1.3805829381504267e-07, -5.522331752601707e-07, -3.3747582932565985e-07, 1.9328161134105974e-06, 5.6949046198705095e-06, 7.649452131381623e-06],
    [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    [1.7257286726880333e-08, 0.0, -4.65946741625769e-07, -8.283497628902559e-07, 3.727573933006152e-06, 1.4910295732024608e-05, 9.111847391792816e-06, -5.2186035062086126e-05, -0.00015376242473650378, -0.00020653520754730382],
    [3.067962084778726e-08, 0.0, -8.283497628902559e-07, -1.2809236054493144e-06, 6.6267981031220475e-06, 2.305662489808766e-05, 1.0064497559808503e-05, -8.06981871433068e-05, -0.00021814634152337594, -0.00028666046030363884],
    [-1.3805829381504267e-07, 0.0, 3.727573933006152e-06, 6.6267981031220475e-06, -2.9820591464049215e-05, -0.00011928236585619686, -7.289477913434253e-05, 0.000417488280496689, 0.0012300993978920302, 0.0016522816603784306],
    [-5.522331752601707e-07, 0.0, 1.4910295732024608e-05, 2.305662489808766e-05, -0.00011928236585619686, -0.00041501924816557786, -0.00018116095607655303, 0.0014525673685795225, 0.0039266341474207675, 0.005159888285465499],
    [-3.3747582932565985e-07, 0.0, 9.111847391792816e-06, 1.0064497559808503e-05, -7.289477913434253e-05, -0.00018116095607655303, 0.001468581806076247, 0.0006340633462679356, -0.01181401175635013, -0.021745034491193898],
    [1.9328161134105974e-06, 0.0, -5.2186035062086126e-05, -8.06981871433068e-05, 0.000417488280496689, 0.0014525673685795225, 0.0006340633462679356, -0.005083985790028328, -0.013743219515972684, -0.018059608999129246],
    [5.6949046198705095e-06, 0.0, -0.00015376242473650378, -0.00021814634152337594, 0.0012300993978920302, 0.0039266341474207675, -0.01181401175635013, -0.013743219515972684, 0.0826466923977296, 0.1638988884584603],
    [7.649452131381623e-06, 0.0, -0.00020653520754730382, -0.00028666046030363884, 0.0016522816603784306, 0.005159888285465499, -0.021745034491193898, -0.018059608999129246, 0.1638988884584603, 0.31358726209239235],
])

_G1_QUADRANT = np.array([
    [-7.900496718847182e-07, 0.0, 1.4220894093924927e-05, 2.5281589500310983e-05, -4.9773129328737247e-05, -0.00022753430550279883, -0.00033182086219158167],
    [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    [1.4220894093924927e-05, 0.0, -0.0002559760936906487, -0.00045506861100559767, 0.0008959163279172705, 0.004095617499050379, 0.00597277551944847],
    [2.5281589500310983e-05, 0.0, -0.00045506861100559767, -0.0009765625, 0.0015927401385195919, 0.0087890625, 0.01329909376597139],
    [-4.9773129328737247e-05, 0.0, 0.0008959163279172705, 0.0015927401385195919, -0.0031357071477104465, -0.014334661246676327, -0.020904714318069645],
    [-0.00022753430550279883, 0.0, 0.004095617499050379, 0.0087890625, -0.014334661246676327, -0.0791015625, -0.1196918438937425],
    [-0.00033182086219158167, 0.0, 0.00597277551944847, 0.01329909376597139, -0.020904714318069645, -0.1196918438937425, 0.8177420190660831],
])


def _mirror(q):
    """Mirror a quadrant into the full filter (atrousfilters.m:167-175)."""
    q = np.hstack([q, np.fliplr(q[:, :-1])])
    return np.vstack([q, np.flipud(q[:-1, :])])


def atrousfilters(fname="maxflat"):
    """Return (h0, h1, g0, g1) FP64 (atrousfilters.m:106-175)."""
    if fname != "maxflat":
        raise ValueError("only the 'maxflat' pyramid filters are restated")
    return (_mirror(_H0_QUADRANT), _mirror(_H1_QUADRANT),
            _mirror(_G0_QUADRANT), _mirror(_G1_QUADRANT))


def _fold(s, n):
    """Half-sample symmetric fold of index s into [0, n) with period 2n."""
    t = np.mod(s, 2 * n)
    return np.where(t < n, t, 2 * n - 1 - t)


def symext(x, p, q, shift, strict=True):
    """Symmetric extension (symext.m:18-31) for a p-by-q filter.

    ``strict=True`` reproduces MATLAB exactly, including its single reflection:
    it raises IndexError where MATLAB's indexing would go out of range.
    ``strict=False`` folds indices with period 2n instead, which equals the
    strict result whenever the strict one exists.
    """
    m, n = x.shape
    p2, q2 = p // 2, q // 2
    s1, s2 = shift
    ss = p2 - s1 + 1
    rr = q2 - s2 + 1
    if strict:
        # columns: [fliplr(x(:,1:ss)) x x(:,n:-1:n-p-s1+1)]
        lo = n - p - s1 + 1          # 1-based last column of the right part
        if ss > n or ss < 0 or (lo <= n and lo < 1):
            raise IndexError("symext: reflection exceeds the image (MATLAB error)")
        left = x[:, :ss][:, ::-1]
        right = x[:, np.arange(n, lo - 1, -1) - 1] if lo <= n else x[:, :0]
        yT = np.hstack([left, x, right])
        lo = m - q - s2 + 1
        if rr > yT.shape[0] or rr < 0 or (lo <= m and lo < 1):
            raise IndexError("symext: reflection exceeds the image (MATLAB error)")
        top = yT[:rr, :][::-1, :]
        bot = yT[np.arange(m, lo - 1, -1) - 1, :] if lo <= m else yT[:0, :]
        yT = np.vstack([top, yT, bot])
        return yT[: m + p - 1, : n + q - 1]
    rows = _fold(np.arange(m + p - 1) - rr, m)
    cols = _fold(np.arange(n + q - 1) - ss, n)
    return x[np.ix_(rows, cols)]


def upsample2df(h, power):
    """Zero-insertion upsampling of a 2D filter by 2**power (upsample2df.m:4-6)."""
    f = 2 ** power
    m, n = h.shape
    ho = np.zeros((f * m, f * n), dtype=h.dtype)
    ho[::f, ::f] = h
    return ho


def atrousc(xe, h, M):
    """Dilated 'valid' convolution of atrousc.cpp:43-73 with M = diag(M, M).

    y(r,c) = sum_{a,b} h(len-1-a, len-1-b) * xe(r + M-1 + a*M, c + M-1 + b*M),
    output size (rows - M*len + 1) x (cols - M*len + 1) (atrousc.cpp:121-122).
    """
    fr, fc = h.shape
    R = xe.shape[0] - M * fr + 1
    C = xe.shape[1] - M * fc + 1
    y = np.zeros((R, C), dtype=np.float64)
    for a in range(fr):
        for b in range(fc):
            w = h[fr - 1 - a, fc - 1 - b]
            if w == 0.0:
                continue
            r0 = M - 1 + a * M
            c0 = M - 1 + b * M
            y += w * xe[r0:r0 + R, c0:c0 + C]
    return y


def conv2_valid(xe, h):
    """MATLAB conv2(xe, h, 'valid')."""
    fr, fc = h.shape
    R = xe.shape[0] - fr + 1
    C = xe.shape[1] - fc + 1
    y = np.zeros((R, C), dtype=np.float64)
    for a in range(fr):
        for b in range(fc):
            w = h[a, b]
            if w == 0.0:
                continue
            y += w * xe[fr - 1 - a:fr - 1 - a + R, fc - 1 - b:fc - 1 - b + C]
    return y


def _ext(x, h, shift, strict):
    return symext(x, h.shape[0], h.shape[1], shift, strict)


def atrousdec(x, fname="maxflat", nlevels=4, strict=True):
    """A trous pyramid decomposition (atrousdec.m:19-51).

    Returns a list [lowpass, coarsest band, ..., finest band] of nlevels+1
    arrays the size of x.
    """
    h0, h1, _, _ = atrousfilters(fname)
    x = np.asarray(x, dtype=np.float64)
    y = [None] * (nlevels + 1)
    shift = (1, 1)
    y0 = conv2_valid(_ext(x, h0, shift, strict), h0)
    y1 = conv2_valid(_ext(x, h1, shift, strict), h1)
    y[nlevels] = y1
    x = y0
    for i in range(1, nlevels):
        s = -2 ** (i - 1) + 2
        L = 2 ** i
        y0 = atrousc(_ext(x, upsample2df(h0, i), (s, s), strict), h0, L)
        y1 = atrousc(_ext(x, upsample2df(h1, i), (s, s), strict), h1, L)
        y[nlevels - i] = y1
        x = y0
    y[0] = x
    return y


def atrousrec(y, fname="maxflat", strict=True):
    """Inverse a trous pyramid (atrousrec.m:18-41)."""
    nlevels = len(y) - 1
    _, _, g0, g1 = atrousfilters(fname)
    x = np.asarray(y[0], dtype=np.float64)
    for i in range(nlevels - 1, 0, -1):
        y1 = np.asarray(y[nlevels - i], dtype=np.float64)
        s = -2 ** (i - 1) + 2
        L = 2 ** i
        x = (atrousc(_ext(x, upsample2df(g0, i), (s, s), strict), g0, L)
             + atrousc(_ext(y1, upsample2df(g1, i), (s, s), strict), g1, L))
    shift = (1, 1)
    x = (conv2_valid(_ext(x, g0, shift, strict), g0)
         + conv2_valid(_ext(np.asarray(y[nlevels], dtype=np.float64), g1, shift, strict), g1))
    return x


# --- Meyer shearing filters ----------------------------------------------------

def meyer_wind(x):
    """Meyer window (meyer_wind.m:11-19), scalar x."""
    if -1 / 3 + 1 / 2 < x < 1 / 3 + 1 / 2:
        return 1.0
    if (1 / 3 + 1 / 2 <= x <= 2 / 3 + 1 / 2) or (-2 / 3 + 1 / 2 <= x <= 1 / 3 + 1 / 2):
        w = 3 * abs(x - 1 / 2) - 1
        z = w ** 4 * (35 - 84 * w + 70 * w ** 2 - 20 * w ** 3)
        return np.cos(np.pi / 2 * z) ** 2
    return 0.0


def windowing(x, L):
    """L-band Meyer windowing of the vector x (windowing.m:11-27).

    Returns an (N, L) array, column j = band j.
    """
    N = len(x)
    y = np.zeros((N, L))
    T = N // L
    g = np.zeros(2 * T)
    for j in range(1, 2 * T + 1):
        n = -1 * T / 2 + j - 1
        g[j - 1] = meyer_wind(n / T)
    for j in range(1, L + 1):
        index = 1
        k = -1 * T / 2
        while k <= 1.5 * T - 1:
            in_sig = int(np.floor(np.mod(k + (j - 1) * T, N))) + 1
            y[in_sig - 1, j - 1] = g[index - 1] * x[in_sig - 1]
            index += 1
            k += 1
    return y


def _matlab_round(v):
    """MATLAB round(): halves away from zero."""
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def _matlab_linspace(a, b, n):
    """MATLAB linspace(a,b,n) = [a+(0:n-2)*(b-a)/(n-1) b]."""
    out = np.empty(n)
    out[: n - 1] = a + (np.arange(n - 1) * (b - a)) / (n - 1)
    out[n - 1] = b
    return out


def gen_x_y_cordinates(n):
    """Pseudo-polar index grids (gen_x_y_cordinates.m:22-111).

    Returns 1-based integer arrays x1n, y1n, x2n, y2n (n x n) and the count
    matrix D (n x n).
    """
    n = n + 1
    x1 = np.zeros((n, n))
    y1 = np.zeros((n, n))
    x2 = np.zeros((n, n))
    y2 = np.zeros((n, n))
    for i in range(1, n + 1):
        y0, x0, x_n, y_n = 1, i, n - i + 1, n
        flag = x_n == x0
        if not flag:
            m1 = (y_n - y0) / (x_n - x0)
        xt = _matlab_linspace(x0, x_n, n)
        for j in range(1, n + 1):
            if not flag:
                yv = _matlab_round(m1 * (xt[j - 1] - x0) + y0)
                xv = _matlab_round(xt[j - 1])
                y1[i - 1, j - 1] = yv
                x1[i - 1, j - 1] = xv
                x2[i - 1, j - 1] = yv
                y2[i - 1, j - 1] = xv
            else:
                x1[i - 1, j - 1] = (n - 1) / 2 + 1
                y1[i - 1, j - 1] = j
                x2[i - 1, j - 1] = j
                y2[i - 1, j - 1] = (n - 1) / 2 + 1
    n = n - 1
    x1n = x1[:n, :n].copy()
    y1n = y1[:n, :n].copy()
    x2n = x2[1:n + 1, :n].copy()
    y2n = y2[1:n + 1, :n].copy()
    x1n = np.flipud(x1n)
    y2n[n - 1, 0] = n
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            D[int(y1n[i, j]) - 1, int(x1n[i, j]) - 1] += 1
    for i in range(n):
        for j in range(n):
            D[int(y2n[i, j]) - 1, int(x2n[i, j]) - 1] += 1
    return (x1n.astype(int), y1n.astype(int), x2n.astype(int), y2n.astype(int), D)


def rec_from_pol(l, n, x1, y1, x2, y2, D):
    """Re-assemble radial slices into a block (rec_from_pol.m:16-35, option 0)."""
    C = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            C[y1[i, j] - 1, x1[i, j] - 1] += l[i, j]
            C[y2[i, j] - 1, x2[i, j] - 1] += l[i + n, j]
    return C / D


def shearing_filters_myer(n1, level):
    """Spatial Meyer shearing filters (shearing_filters_Myer.m:18-27).

    Returns (2**level, n1, n1); w_s[k] is MATLAB's w_s(:,:,k+1).
    """
    x11, y11, x12, y12, F1 = gen_x_y_cordinates(n1)
    wf = windowing(np.ones(2 * n1), 2 ** level)
    out = np.zeros((2 ** level, n1, n1))
    for k in range(2 ** level):
        temp = np.outer(wf[:, k], np.ones(n1))
        C = rec_from_pol(temp, n1, x11, y11, x12, y12, F1)
        out[k] = np.real(np.fft.fftshift(np.fft.ifft2(np.fft.fftshift(C)))) / np.sqrt(n1)
    return out


def shear_spectra(w_s, L, shift="none"):
    """Normalised L x L shearing-filter spectra of one scale (nsst_dec2.m:47-58).

    shear_f(:,:,k) = fft2(w_s(:,:,k), L, L) / L, then divided per frequency by
    sqrt(sum_k |shear_f(:,:,k)|^2).  ``shift='center'`` first rolls the padded
    filter by -n1/2 in both axes, the convention of the MEX GPU path
    (prepare_shearing_filters_cuda.cpp:254, shear_cuda.cu:2303-2305).
    """
    D, n1, _ = w_s.shape
    H = np.zeros((D, L, L), dtype=np.complex128)
    for k in range(D):
        pad = np.zeros((L, L))
        pad[:n1, :n1] = w_s[k]
        if shift == "center":
            pad = np.roll(pad, (-(n1 // 2), -(n1 // 2)), axis=(0, 1))
        elif shift != "none":
            raise ValueError(shift)
        H[k] = np.fft.fft2(pad) / L
    d = np.sqrt(np.sum(H * np.conj(H), axis=0).real)
    return H / d


def conv2p(X, y):
    """Periodic convolution real(ifft2(X .* fft2(y))) (conv2p.m:12)."""
    return np.real(np.fft.ifft2(X * np.fft.fft2(y)))


def make_dictionary(L, dcomp=(3, 3, 4, 4), dsize=(32, 32, 16, 16), shift="none"):
    """Spectra of every scale, coarse to fine (nsst_dec2.m:47-58)."""
    return [shear_spectra(shearing_filters_myer(n1, lev), L, shift)
            for lev, n1 in zip(dcomp, dsize)]


def nsst_dec2(x, shear_f, lpfilt="maxflat", strict=True):
    """Forward NSST (nsst_dec2.m:37-60) with precomputed spectra ``shear_f``.

    Returns [lowpass (L,L), scale 1 (D1,L,L), ..., scale S (DS,L,L)].
    """
    level = len(shear_f)
    y = atrousdec(x, lpfilt, level, strict)
    dst = [y[0]]
    for i in range(level):
        Y = np.fft.fft2(y[i + 1])
        dst.append(np.real(np.fft.ifft2(shear_f[i] * Y[None])))
    return dst


def nsst_rec2(dst, shear_f, lpfilt="maxflat", strict=True):
    """Inverse NSST (nsst_rec2.m:25-35)."""
    level = len(dst) - 1
    y = [np.asarray(dst[0], dtype=np.float64)]
    for i in range(level):
        C = np.fft.fft2(np.asarray(dst[i + 1], dtype=np.float64))
        y.append(np.real(np.sum(np.fft.ifft2(np.conj(shear_f[i]) * C), axis=0)))
    return np.real(atrousrec(y, lpfilt, strict))


def _mad(v):
    v = v.ravel()
    return np.median(np.abs(v - np.median(v))) / 0.6745


def nsst_scalars(shear_f, noise, lpfilt="maxflat", strict=True):
    """Per-band noise scalars by MAD/.6745 (nsst_scalars.m:28-45).

    ``noise`` is the unit-variance white-noise image (the reference draws it
    with MATLAB's randn; here it is an input).  Returns [s_low, array(D1), ...].
    """
    level = len(shear_f)
    y = atrousdec(noise, lpfilt, level, strict)
    out = [_mad(y[0])]
    for i in range(level):
        Y = np.fft.fft2(y[i + 1])
        out.append(np.array([_mad(np.real(np.fft.ifft2(shear_f[i][k] * Y)))
                             for k in range(shear_f[i].shape[0])]))
    return out


def nsst_HT(dst, sigma, Tscalars, dst_scalars):
    """Hard thresholding, keep iff |c| > T*sigma*s (nsst_HT.m:32-47)."""
    level = len(dst) - 1
    out = [dst[0] * (np.abs(dst[0]) > Tscalars[0] * sigma * dst_scalars[0])]
    for i in range(level):
        T = Tscalars[1] if i < level - 1 else Tscalars[2]
        th = T * sigma * np.asarray(dst_scalars[i + 1])[:, None, None]
        out.append(dst[i + 1] * (np.abs(dst[i + 1]) > th))
    return out


def nsst_thresholds(sigma, Tscalars, dst_scalars):
    """The per-band thresholds nsst_HT applies, flattened low-pass first."""
    level = len(dst_scalars) - 1
    th = [Tscalars[0] * sigma * dst_scalars[0]]
    for i in range(level):
        T = Tscalars[1] if i < level - 1 else Tscalars[2]
        th.extend(T * sigma * np.asarray(dst_scalars[i + 1]))
    return np.array(th, dtype=np.float64)


def thresh(Ct, lam, option, E, sc, opt):
    """Hard thresholding of shearcuda/thresh.m:28-55.

    Ct: list of arrays; scale s is (D_s, L, L) or (L, L).  E is (S+1, maxD).
    Returns (coeff, maxi).
    """
    Ct = [np.array(c, dtype=np.float64, copy=True) for c in Ct]
    temp = []
    if option == 2:
        for s in range(len(Ct)):
            c = Ct[s] if Ct[s].ndim == 3 else Ct[s][None]
            for w in range(c.shape[0]):
                c[w] = c[w] * (np.abs(c[w]) > sc[s] * lam * E[s, w])
                if opt == 1:
                    temp.append(np.max(np.abs(c[w] / E[s, w])))
            Ct[s] = c if Ct[s].ndim == 3 else c[0]
    else:
        Ct[1] = Ct[1] * (np.abs(Ct[1]) > lam)
        if opt == 1:
            temp.append(np.max(np.abs(Ct[1])))
    maxi = max(temp) if opt == 1 else 0.0
    return Ct, maxi


def soft_threshold(x, t):
    """sign(x) .* max(|x| - t, 0) (shearcuda/SoftThresh.m:11-13)."""
    return np.sign(x) * np.maximum(np.abs(x) - t, 0.0)


def re_weight_band(P, coeff, strict=True):
    """atrousrec(coeff(j) * atrousdec(P)) (shearcuda/re_weight_band.m:23-29)."""
    y = atrousdec(P, "maxflat", len(coeff) - 1, strict)
    y = [c * yy for c, yy in zip(coeff, y)]
    return atrousrec(y, "maxflat", strict)


def com_norm(shear_f, L, lpfilt="maxflat", strict=True):
    """Exact band norms E(s,w) from an impulse (shearcuda/com_norm.m:38-53).

    Returns an (S+1, maxD) array (zero-padded rows, like MATLAB's E).
    """
    F = np.ones((L, L))
    X = np.fft.fftshift(np.real(np.fft.ifft2(F))) * np.sqrt(F.size)
    C = nsst_dec2(X, shear_f, lpfilt, strict)
    maxD = max(s.shape[0] for s in shear_f)
    E = np.zeros((len(C), maxD))
    for s, c in enumerate(C):
        c = c if c.ndim == 3 else c[None]
        for w in range(c.shape[0]):
            E[s, w] = np.sqrt(np.sum(c[w] * c[w]) / c[w].size)
    return E
