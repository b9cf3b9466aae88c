"""Host-side mirror of ShearCuda's MATLAB entry points on GPU tensors.

Each function keeps the reference's name, argument meaning and error text and
calls the C ABI (include/shx.h); device memory and streams come from PyTorch
(tensors on ``cuda``), which is plumbing only.  Inputs must already be FP32
tensors on the GPU: unlike the reference MEX files, which forward CPU arrays to
the MATLAB implementation (e.g. shearcuda/src_cuda/shear_trans_cuda.cpp:94-98),
there is no CPU path here and CPU inputs raise TypeError.

Arrays are row-major: element [r, c] is MATLAB element (r+1, c+1).  Scale j
of a coefficient list is a (D_j, L, L) tensor (MATLAB's L x L x D_j cell
element); cell 0 is the (L, L) lowpass.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr_array
from .filters import DEFAULT_DCOMP, DEFAULT_DSIZE, DEFAULT_TSCALARS, dictionary_filters, shearing_filters_myer

SHIFT_NONE = 0
SHIFT_CENTER = 1
_SHIFTS = {"none": SHIFT_NONE, "center": SHIFT_CENTER}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu(x, name="x"):
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("%s must be a FP32 tensor on the GPU (no CPU path)" % name)
    if x.dtype != torch.float32:
        raise TypeError("%s: Signal should be GPUsingle (float32), got %s" % (name, x.dtype))
    if not x.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _workspace(nfloats, device):
    """Scratch of the C calls that take an optional workspace, from torch's
    caching allocator: passing one spares the library a hipMalloc, a stream
    synchronisation and a hipFree per call."""
    return torch.empty(int(nfloats), device=device, dtype=torch.float32)


def _require_shape(x, shape, name="x"):
    """_require_gpu plus an exact shape; the C ABI trusts buffer sizes, so
    every tensor is checked before anything is launched."""
    _require_gpu(x, name)
    if tuple(x.shape) != tuple(shape):
        raise ValueError("%s must have shape %s, got %s" % (name, tuple(shape), tuple(x.shape)))


def _require_cells(Ct, shear, name="Ct"):
    """Coefficient cells of a dictionary: Ct{1} (L, L), Ct{j+1} (D_j, L, L)."""
    if len(Ct) != shear.nscales + 1:
        raise ValueError("Incorrect number of cell elements")
    L = shear.L
    _require_shape(Ct[0], (L, L), "%s{1}" % name)
    for j, d in enumerate(shear.ndir):
        _require_shape(Ct[j + 1], (d, L, L), "%s{%d}" % (name, j + 2))


def _flat_scalars(shear, scalars):
    """[s_low, array(D_1), ..., array(D_S)] -> 1 + sum(D) FP64 values."""
    if len(scalars) != shear.nscales + 1:
        raise ValueError("scalars must hold the lowpass and one array per scale (%d entries)" % (shear.nscales + 1))
    if np.size(scalars[0]) != 1:
        raise ValueError("scalars[0] (lowpass) must be one value")
    parts = [np.asarray(scalars[0], dtype=np.float64).reshape(1)]
    for j, d in enumerate(shear.ndir):
        a = np.asarray(scalars[j + 1], dtype=np.float64).ravel()
        if a.size != d:
            raise ValueError("scalars[%d] must hold %d values, got %d" % (j + 1, d, a.size))
        parts.append(a)
    return np.ascontiguousarray(np.concatenate(parts))


def _tscalars(Tscalars):
    T = np.ascontiguousarray(np.asarray(Tscalars, dtype=np.float64).ravel())
    if T.size != 3:
        raise ValueError("Tscalars must hold 3 values (lowpass, coarser scales, finest scale)")
    return T


class ShearDict:
    """A prepared shearlet dictionary (the ``shear`` struct of
    prepare_shearing_filters_cuda.cpp:234-248): FP32 spectra on the GPU plus
    workspaces.  Free it with clear_shearing_filters_cuda() or close()."""

    def __init__(self, handle, L, ndir, dsize, shift, w_s, device):
        self._h = handle
        self.L = L
        self.ndir = list(ndir)
        self.dsize = list(dsize)
        self.shift = shift
        self.w_s = w_s
        self.device = device
        self.atrousfilter = "maxflat"

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("Invalid shearlet dictionary (already cleared)")
        return self._h

    @property
    def nscales(self):
        return len(self.ndir)

    @property
    def nbands(self):
        return 1 + sum(self.ndir)

    def nbytes(self):
        return int(lib.shx_dict_bytes(self.handle))

    def close(self):
        if self._h is not None:
            lib.shx_clear_shearing_filters(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_coeffs(self):
        """Empty coefficient cells for this dictionary (Ct of shear_trans_cuda)."""
        L = self.L
        dev = self.device
        return [torch.empty((L, L), device=dev, dtype=torch.float32)] + [
            torch.empty((d, L, L), device=dev, dtype=torch.float32) for d in self.ndir]


def prepare_shearing_filters_cuda(w_s, L, shift="none", spectra="device"):
    """shear = prepare_shearing_filters_cuda(w_s, L)
    (shearcuda/src_cuda/prepare_shearing_filters_cuda.cpp:61-249).

    w_s: list over scales of (D_j, n1_j, n1_j) FP64 arrays.
    shift: 'none' (toolbox convention, nsst_dec2.m:50) or 'center' (the MEX GPU
    path's circular n1/2 shift, prepare_shearing_filters_cuda.cpp:254).
    spectra: 'device' (FP64 direct transform on the GPU) or 'host' (FP64 FFT
    on the host, uploaded)."""
    if shift not in _SHIFTS:
        raise ValueError("shift must be 'none' or 'center'")
    if spectra not in ("device", "host"):
        raise ValueError("spectra must be 'device' or 'host'")
    w_s = [np.ascontiguousarray(np.asarray(w, dtype=np.float64)) for w in w_s]
    for w in w_s:
        if w.ndim != 3 or w.shape[1] != w.shape[2]:
            raise ValueError("each w_s{j} must be (D, n1, n1)")
    S = len(w_s)
    ndir = (ctypes.c_int * S)(*[w.shape[0] for w in w_s])
    dsize = (ctypes.c_int * S)(*[w.shape[1] for w in w_s])
    ptrs = (ctypes.POINTER(ctypes.c_double) * S)(
        *[w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for w in w_s])
    handle = ctypes.c_void_p()
    torch.cuda.init()
    check(lib.shx_prepare_shearing_filters_ex(int(L), S, ndir, dsize, ptrs, _SHIFTS[shift],
                                              1 if spectra == "host" else 0, ctypes.byref(handle)))
    return ShearDict(handle, int(L), list(ndir), list(dsize), shift, w_s, torch.device("cuda", torch.cuda.current_device()))


def shearing_filters_Myer(m, num, L, dataType="GPUsingle", shift="none", filters="device"):
    """shear = shearing_filters_Myer(dsize, dcomp, L, 'GPUsingle')
    (shearcuda/shearing_filters_Myer.m:29-81, GPU branch).

    filters: 'device' builds the spatial Meyer filters and their spectra on
    the GPU (shx_shearing_filters_myer); 'host' builds the filters in numpy
    (filters.py) and the spectra on the GPU."""
    if dataType not in ("GPUsingle",):
        raise ValueError("only 'GPUsingle' dictionaries exist on this path")
    if shift not in _SHIFTS:
        raise ValueError("shift must be 'none' or 'center'")
    if filters == "host":
        return prepare_shearing_filters_cuda(dictionary_filters(num, m), L, shift)
    if filters != "device":
        raise ValueError("filters must be 'device' or 'host'")
    if len(m) != len(num):
        raise ValueError("dcomp and dsize must have the same length")
    S = len(m)
    dcomp = (ctypes.c_int * S)(*[int(v) for v in num])
    dsize = (ctypes.c_int * S)(*[int(v) for v in m])
    handle = ctypes.c_void_p()
    torch.cuda.init()
    check(lib.shx_shearing_filters_myer(int(L), S, dcomp, dsize, _SHIFTS[shift], 0, ctypes.byref(handle)))
    return ShearDict(handle, int(L), [1 << int(v) for v in num], [int(v) for v in m], shift, None,
                     torch.device("cuda", torch.cuda.current_device()))


def meyer_filters_cuda(n1, level):
    """The 2^level spatial Meyer shearing filters of side n1, built on the GPU
    in FP64 (shx_meyer_filters): a (2^level, n1, n1) float64 device tensor,
    element [k, r, c] = MATLAB w_s(r+1, c+1, k+1)."""
    w = torch.empty((1 << int(level), int(n1), int(n1)), device="cuda", dtype=torch.float64)
    check(lib.shx_meyer_filters(int(n1), int(level), ctypes.c_void_p(w.data_ptr()), _stream()))
    return w


def clear_shearing_filters_cuda(shear):
    """clear_shearing_filters_cuda(shear) (clear_shearing_filters_cuda.cpp:51-96)."""
    shear.close()


def _check_pfilt(pfilt):
    if pfilt != "maxflat":
        raise ValueError("only the 'maxflat' pyramid filters are supported on this path")


def shear_trans_cuda(x, pfilt, shear, Ct=None):
    """Ct = shear_trans_cuda(x, 'maxflat', shear [, Ct])
    (shearcuda/src_cuda/shear_trans_cuda.cpp:56-217)."""
    _check_pfilt(pfilt)
    _require_gpu(x)
    L = shear.L
    if x.dim() != 2:
        raise ValueError("Input data should be 2-dimensional")
    if x.shape[0] != x.shape[1]:
        raise ValueError("Input data should be a square")
    if x.shape[0] != L:
        raise ValueError("Input data size does not match dictionary size")
    if Ct is None:
        Ct = shear.alloc_coeffs()
    _require_cells(Ct, shear)
    check(lib.shx_shear_trans(shear.handle, ctypes.c_void_p(x.data_ptr()),
                              ptr_array([c.data_ptr() for c in Ct]), _stream()))
    return Ct


def inverse_shear_cuda(Ct, pfilt, shear, out=None):
    """x = inverse_shear_cuda(Ct, 'maxflat', shear)
    (shearcuda/src_cuda/inverse_shear_cuda.cpp:57-219)."""
    _check_pfilt(pfilt)
    _require_cells(Ct, shear)
    if out is None:
        out = torch.empty((shear.L, shear.L), device=Ct[0].device, dtype=torch.float32)
    _require_shape(out, (shear.L, shear.L), "out")
    check(lib.shx_inverse_shear(shear.handle, ptr_array([c.data_ptr() for c in Ct]),
                                ctypes.c_void_p(out.data_ptr()), _stream()))
    return out


def shear_trans_adjoint_cuda(Ct_bar, shear, out=None):
    """Transpose of shear_trans_cuda: coefficient gradients -> image gradient."""
    _require_cells(Ct_bar, shear, "Ct_bar")
    if out is None:
        out = torch.empty((shear.L, shear.L), device=Ct_bar[0].device, dtype=torch.float32)
    _require_shape(out, (shear.L, shear.L), "out")
    check(lib.shx_shear_trans_adjoint(shear.handle, ptr_array([c.data_ptr() for c in Ct_bar]),
                                      ctypes.c_void_p(out.data_ptr()), _stream()))
    return out


def inverse_shear_adjoint_cuda(x_bar, shear, Ct_bar=None):
    """Transpose of inverse_shear_cuda: image gradient -> coefficient gradients."""
    _require_shape(x_bar, (shear.L, shear.L), "x_bar")
    if Ct_bar is None:
        Ct_bar = shear.alloc_coeffs()
    _require_cells(Ct_bar, shear, "Ct_bar")
    check(lib.shx_inverse_shear_adjoint(shear.handle, ctypes.c_void_p(x_bar.data_ptr()),
                                        ptr_array([c.data_ptr() for c in Ct_bar]), _stream()))
    return Ct_bar


def inverse_shear_thresh_cuda(Ct, th, shear, out=None):
    """Inverse NSST of hard-thresholded bands (thresholds fused into the
    reconstruction's first read; Ct is not modified).  th: 1 + sum(D_j)
    thresholds, lowpass first."""
    _require_cells(Ct, shear)
    th = np.ascontiguousarray(np.asarray(th, dtype=np.float32).ravel())
    if th.size != shear.nbands:
        raise ValueError("need %d thresholds" % shear.nbands)
    if out is None:
        out = torch.empty((shear.L, shear.L), device=Ct[0].device, dtype=torch.float32)
    _require_shape(out, (shear.L, shear.L), "out")
    check(lib.shx_inverse_shear_thresh(shear.handle, ptr_array([c.data_ptr() for c in Ct]),
                                       th.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       ctypes.c_void_p(out.data_ptr()), _stream()))
    return out


def band_thresholds(shear, sigma, scalars, Tscalars=DEFAULT_TSCALARS):
    """nsst_HT thresholds T*sigma*s per band (shearlet_toolbox_1/nsst_HT.m:32-47),
    lowpass first.  scalars: [s_low, array(D_1), ..., array(D_S)]."""
    S = shear.nscales
    sc = _flat_scalars(shear, scalars)
    T = _tscalars(Tscalars)
    th = [T[0] * sigma * sc[0]]
    k = 1
    for j, d in enumerate(shear.ndir):
        th.extend((T[1] if j < S - 1 else T[2]) * sigma * sc[k:k + d])
        k += d
    return np.array(th, dtype=np.float64)


def denoise_cuda(x, shear, sigma, scalars, Tscalars=DEFAULT_TSCALARS, Ct=None, out=None, coeffs=True):
    """Plain NSST hard-threshold denoise: shear_trans + nsst_HT + inverse with
    the threshold fused into the inverse.  Returns (x_hat, Ct); with
    coeffs=False the coefficients are neither returned nor written
    (x_hat, None)."""
    L = shear.L
    _require_shape(x, (L, L))
    if not coeffs:
        Ct = None
    elif Ct is None:
        Ct = shear.alloc_coeffs()
    if Ct is not None:
        _require_cells(Ct, shear)
    if out is None:
        out = torch.empty_like(x)
    _require_shape(out, (L, L), "out")
    sc = _flat_scalars(shear, scalars)
    T = _tscalars(Tscalars)
    cells = ptr_array([c.data_ptr() for c in Ct]) if Ct is not None else None
    check(lib.shx_denoise(shear.handle, ctypes.c_void_p(x.data_ptr()), cells, float(sigma), T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                          sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_void_p(out.data_ptr()),
                          _stream()))
    return out, Ct


def thresh_cuda(Ct, lam, option, E, sc, opt):
    """maxi = thresh_cuda(Ct, lambda, option, E, sc, opt), in place
    (shearcuda/src_cuda/thresh_cuda.cpp:57-217)."""
    for i, c in enumerate(Ct):
        _require_gpu(c, "Ct{%d}" % (i + 1))
        if c.dim() not in (2, 3):
            raise ValueError("Ct{%d} must be a plane or a stack of planes" % (i + 1))
    E = np.asarray(E, dtype=np.float64)
    if E.ndim == 1:
        E = E[:, None]
    if E.ndim != 2:
        raise ValueError("E must be a matrix")
    sc = np.ascontiguousarray(np.asarray(sc, dtype=np.float64).ravel())
    Ecm = np.asfortranarray(E)
    planes = (ctypes.c_int * len(Ct))(*[c.shape[0] if c.dim() == 3 else 1 for c in Ct])
    elems = (ctypes.c_int64 * len(Ct))(*[c.shape[-1] * c.shape[-2] for c in Ct])
    maxi = ctypes.c_double(0.0)
    Eflat = np.ascontiguousarray(Ecm.ravel(order="F"))
    check(lib.shx_thresh(ptr_array([c.data_ptr() for c in Ct]), len(Ct), planes, elems, float(lam), int(option),
                         Eflat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), E.shape[0], E.shape[1],
                         sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), sc.size, int(opt), ctypes.byref(maxi),
                         _stream()))
    return maxi.value


def atrousdec_cuda(x, pfilt, nlevels):
    """y = atrousdec(x, 'maxflat', nlevels) on the GPU
    (shearcuda/src_cuda/atrousdec.cpp:41-131): [lowpass, coarse ... fine]."""
    _check_pfilt(pfilt)
    _require_gpu(x)
    if x.dim() != 2 or x.shape[0] != x.shape[1]:
        raise ValueError("Input data should be a square")
    L = x.shape[0]
    y = [torch.empty_like(x) for _ in range(nlevels + 1)]
    work = _workspace(2 * L * L, x.device)
    check(lib.shx_atrousdec(ctypes.c_void_p(x.data_ptr()), L, int(nlevels), ptr_array([t.data_ptr() for t in y]),
                            _vp(work), _stream()))
    return y


def atrousrec_cuda(y, pfilt):
    """x = atrousrec(y, 'maxflat') on the GPU (shearcuda/src_cuda/atrousrec.cpp:42-123)."""
    _check_pfilt(pfilt)
    if len(y) < 2:
        raise ValueError("y must hold the lowpass and at least one level")
    _require_gpu(y[0], "y{1}")
    if y[0].dim() != 2 or y[0].shape[0] != y[0].shape[1]:
        raise ValueError("Input data should be a square")
    for i, t in enumerate(y):
        _require_shape(t, y[0].shape, "y{%d}" % (i + 1))
    out = torch.empty_like(y[0])
    work = _workspace(2 * y[0].numel(), out.device)
    check(lib.shx_atrousrec(ptr_array([t.data_ptr() for t in y]), y[0].shape[0], len(y) - 1,
                            ctypes.c_void_p(out.data_ptr()), _vp(work), _stream()))
    return out


def convolution_cuda(y, shear, Ct=None):
    """Directional bands of given pyramid levels y = [lowpass, coarse ... fine]
    (Shearcuda::convolution, demo_denoise/shearcuda/shearcuda.cpp:549-624);
    Ct[0] is a copy of y[0]."""
    if len(y) != shear.nscales + 1:
        raise ValueError("y must hold the lowpass and one level per scale")
    for i, t in enumerate(y):
        _require_shape(t, (shear.L, shear.L), "y{%d}" % (i + 1))
    Ct = shear.alloc_coeffs() if Ct is None else Ct
    _require_cells(Ct, shear)
    check(lib.shx_convolution(shear.handle, ptr_array([t.data_ptr() for t in y]),
                              ptr_array([c.data_ptr() for c in Ct]), _stream()))
    return Ct


def applyfilters_cuda(Ct, shear):
    """Conjugate-filtered band sums per scale (Shearcuda::applyfilters,
    demo_denoise/shearcuda/shearcuda.cpp:719-800): [None, level 1 ... S]."""
    _require_cells(Ct, shear)
    y = [None] + [torch.empty((shear.L, shear.L), device=Ct[0].device) for _ in range(shear.nscales)]
    check(lib.shx_applyfilters(shear.handle, ptr_array([c.data_ptr() for c in Ct]),
                               ptr_array([0] + [t.data_ptr() for t in y[1:]]), _stream()))
    return y


def add_noise_cuda(x, sigma, seed=1234):
    """x += N(0, sigma^2) in place (Shearcuda::addNoise,
    demo_denoise/shearcuda/shearcuda.cpp:140-163; counter-based generator)."""
    _require_gpu(x)
    check(lib.shx_add_noise(ctypes.c_void_p(x.data_ptr()), x.numel(), float(sigma), int(seed), _stream()))
    return x


def maxflat_filters():
    """The four maxflat a trous filters (h0, h1, g0, g1) of the library, FP64
    (shearlet_toolbox_1/atrousfilters.m:106-175)."""
    out = []
    for which in range(4):
        n = ctypes.c_int()
        check(lib.shx_maxflat_filter(which, None, ctypes.byref(n)))
        h = np.empty((n.value, n.value))
        check(lib.shx_maxflat_filter(which, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
        out.append(h)
    return out


def save_shearing_filters(path, dsize=DEFAULT_DSIZE, dcomp=DEFAULT_DCOMP, dataType="single"):
    """Write the SHFM0002 filter file of
    demo_denoise/makeFilters/save_shearing_filters.m:43-85 (read by the C++
    demo's ShearConfig): header, scale count, data type (0 single, 2 double),
    per scale size, direction count and the column-major Meyer filters, then
    the four maxflat filters."""
    if dataType not in ("single", "double"):
        raise ValueError("dataType must be 'single' or 'double'")
    dt = np.float32 if dataType == "single" else np.float64
    with open(path, "wb") as f:
        f.write(b"SHFM0002")
        f.write(np.array([len(dcomp), 0 if dataType == "single" else 2], dtype=np.uint16).tobytes())
        for n1, lev in zip(dsize, dcomp):
            w = shearing_filters_myer(n1, lev)
            f.write(np.array([n1, w.shape[0]], dtype=np.uint16).tobytes())
            for k in range(w.shape[0]):
                f.write(np.ascontiguousarray(w[k].T, dtype=dt).tobytes())
        for h in maxflat_filters():
            f.write(np.array([h.shape[0]], dtype=np.uint16).tobytes())
            f.write(np.ascontiguousarray(h.T, dtype=dt).tobytes())


def re_weight_band_cuda(x, coeff, shear=None):
    """y = re_weight_band_cuda(x, coeff [, shear])
    (shearcuda/src_cuda/re_weight_band_cuda.cpp:59-183); always 'maxflat'."""
    _require_gpu(x)
    if x.dim() != 2 or x.shape[0] != x.shape[1]:
        raise ValueError("Input data should be a square")
    coeff = np.ascontiguousarray(np.asarray(coeff, dtype=np.float64).ravel())
    if coeff.size <= 1:
        raise ValueError("There should be at least one band")
    out = torch.empty_like(x)
    work = _workspace((coeff.size + 2) * x.numel(), x.device)
    check(lib.shx_re_weight_band(ctypes.c_void_p(x.data_ptr()), x.shape[0],
                                 coeff.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), coeff.size,
                                 ctypes.c_void_p(out.data_ptr()), _vp(work), _stream()))
    return out


def norm_cuda(x, p, idx=None):
    """norm_cuda(x, p, idx): (sum |x(:,:,idx)|^p)^(1/p)
    (shearcuda/src_cuda/norm_cuda.cpp:57-215); idx is 1-based like MATLAB."""
    _require_gpu(x)
    t = x if idx is None else x[idx - 1]
    t = t.contiguous()
    s = ctypes.c_double(0.0)
    check(lib.shx_norm_pow(ctypes.c_void_p(t.data_ptr()), t.numel(), float(p), ctypes.byref(s), _stream()))
    return s.value ** (1.0 / p)


def softth_cuda(x, t):
    """softth_cuda(x, t): in-place sign(x).*max(|x|-t,0) (softth_cuda.cpp:49-112)."""
    _require_gpu(x)
    check(lib.shx_softth(ctypes.c_void_p(x.data_ptr()), x.numel(), float(t), _stream()))
    return x


def com_norm(pfilt, n, shear):
    """E = com_norm(pfilt, n, shear): exact band norms from the transform of a
    scaled impulse (shearcuda/com_norm.m:25-55).  Returns (S+1, maxD) FP64."""
    _check_pfilt(pfilt)
    L = int(n[0]) if np.ndim(n) else int(n)
    X = np.zeros((L, L), dtype=np.float32)
    X[L // 2, L // 2] = L   # fftshift(real(ifft2(ones))) * sqrt(L*L)
    C = shear_trans_cuda(torch.from_numpy(X).to(shear.device), pfilt, shear)
    E = np.zeros((len(C), max(shear.ndir)))
    for s, c in enumerate(C):
        c = c if c.dim() == 3 else c[None]
        for w in range(c.shape[0]):
            E[s, w] = norm_cuda(c, 2, w + 1) / np.sqrt(c.shape[-1] * c.shape[-2])
    return E


MAD_ARITH = {"fp64": 0, "fp32": 1}


def band_mad_cuda(planes, arith="fp64"):
    """Median absolute deviation / 0.6745 of each plane of a contiguous device
    stack (first dimension indexes the planes), computed on the device with
    exact medians (shx_band_mad).  arith "fp64": values widened to double, as
    numpy / MATLAB on the double transform (nsst_scalars.m:36-44); "fp32":
    float arithmetic as the demo (demo_denoise/shearcuda/shearcuda.cpp:246-252).
    Returns a float64 device tensor with one entry per plane."""
    _require_gpu(planes, "planes")
    if arith not in MAD_ARITH:
        raise ValueError("arith must be 'fp64' or 'fp32'")
    if planes.dtype != torch.float32 or not planes.is_contiguous():
        raise ValueError("planes must be a contiguous float32 tensor")
    n = planes.shape[0] if planes.dim() > 0 else 1
    pe = planes.numel() // max(n, 1)
    out = torch.empty(n, device=planes.device, dtype=torch.float64)
    if n:
        check(lib.shx_band_mad(ctypes.c_void_p(planes.data_ptr()), pe, pe, n, MAD_ARITH[arith],
                               ctypes.c_void_p(out.data_ptr()), _stream()))
    return out


def nsst_scalars_cuda(shear, noise, arith="fp64"):
    """Per-band MAD/0.6745 noise scalars of the transform of a unit white-noise
    image (shearlet_toolbox_1/nsst_scalars.m:28-45), all on the device: the
    transform, then exact medians per band (shx_nsst_scalars).  arith "fp32"
    gives demo_denoise/shearcuda/shearcuda.cpp:241-268's float medians.
    Returns [s_low, array(D_1), ..., array(D_S)]."""
    _require_gpu(noise, "noise")
    if arith not in MAD_ARITH:
        raise ValueError("arith must be 'fp64' or 'fp32'")
    if tuple(noise.shape) != (shear.L, shear.L) or noise.dtype != torch.float32 or not noise.is_contiguous():
        raise ValueError("noise must be a contiguous %dx%d float32 image" % (shear.L, shear.L))
    sc = np.zeros(shear.nbands, dtype=np.float64)
    check(lib.shx_nsst_scalars(shear.handle, ctypes.c_void_p(noise.data_ptr()), MAD_ARITH[arith],
                               sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _stream()))
    out, k = [float(sc[0])], 1
    for d in shear.ndir:
        out.append(sc[k:k + d].copy())
        k += d
    return out


def make_dictionary(L, dcomp=DEFAULT_DCOMP, dsize=DEFAULT_DSIZE, shift="none", spectra="device"):
    """Shortcut: spatial Meyer filters + prepare_shearing_filters_cuda."""
    return prepare_shearing_filters_cuda(dictionary_filters(dcomp, dsize), L, shift, spectra)


# ---- HIP graphs ----------------------------------------------------------------

class Graph:
    """A recorded sequence of shx launches (hipGraph); launch() replays it on
    the current stream with the recorded pointers and thresholds."""

    def __init__(self, handle, keep=()):
        self._h = handle
        self._keep = keep  # tensors the graph reads or writes

    def launch(self):
        check(lib.shx_graph_launch(self._h, _stream()))

    def close(self):
        if self._h is not None:
            lib.shx_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture(fn, keep=()):
    """Record the shx calls made by fn() into a Graph (on a side stream)."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        check(lib.shx_capture_begin(ctypes.c_void_p(s.cuda_stream)))
        err = None
        try:
            fn()
        except Exception as e:  # finish the capture before re-raising
            err = e
        h = ctypes.c_void_p()
        st = lib.shx_capture_end(ctypes.c_void_p(s.cuda_stream), ctypes.byref(h))
    torch.cuda.current_stream().wait_stream(s)
    if err is not None:
        if st == 0:
            lib.shx_graph_destroy(h)
        raise err
    check(st)
    return Graph(h, keep)


def capture_parallel(fns, keep=()):
    """Record independent groups of shx calls into one Graph whose groups have
    no order between them (e.g. the batch plans of a mixed-size batch): each
    fn is recorded into a graph of its own, and shx_graph_parallel puts
    copies of those graphs side by side as child nodes of one graph, so the
    runtime may run them concurrently.  No stream is forked while a capture
    is open."""
    parts = []
    try:
        for fn in fns:
            parts.append(capture(fn))
        arr = (ctypes.c_void_p * len(parts))(*[p._h.value for p in parts])
        h = ctypes.c_void_p()
        check(lib.shx_graph_parallel(arr, len(parts), ctypes.byref(h)))
    finally:
        for p in parts:
            p.close()
    return Graph(h, keep)


def denoise_graph(x, shear, sigma, scalars, Tscalars=DEFAULT_TSCALARS, Ct=None, out=None, coeffs=True):
    """Graph of one plain denoise step (replay with .launch()); coeffs=False:
    the image only (no coefficients written)."""
    if Ct is None and coeffs:
        Ct = shear.alloc_coeffs()
    if out is None:
        out = torch.empty_like(x)
    g = capture(lambda: denoise_cuda(x, shear, sigma, scalars, Tscalars, Ct=Ct, out=out, coeffs=coeffs),
                keep=(x, Ct, out))
    return g, (Ct if coeffs else None), out


# ---- batched form -----------------------------------------------------------------

class Batch:
    """nimg images of the dictionary's size through every kernel at once.
    Buffers: x/out (nimg, L, L); coefficients (nimg, 1 + sum D, L, L) with the
    lowpass first; thresholds (nimg, 1 + sum D) FP32 on the device."""

    def __init__(self, shear, nimg):
        self.shear = shear
        self.nimg = int(nimg)
        h = ctypes.c_void_p()
        check(lib.shx_batch_create(shear.handle, self.nimg, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h is not None:
            lib.shx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_coeffs(self):
        L = self.shear.L
        return torch.empty((self.nimg, self.shear.nbands, L, L), device=self.shear.device, dtype=torch.float32)

    def _img(self, t, name):
        L = self.shear.L
        _require_shape(t, (self.nimg, L, L), name)

    def _coef(self, t, name):
        L = self.shear.L
        _require_shape(t, (self.nimg, self.shear.nbands, L, L), name)

    def _th(self, th):
        _require_shape(th, (self.nimg, self.shear.nbands), "th")

    def shear_trans(self, x, coeff):
        self._img(x, "x")
        self._coef(coeff, "coeff")
        check(lib.shx_batch_shear_trans(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(coeff.data_ptr()),
                                        _stream()))
        return coeff

    def inverse_shear(self, coeff, out, th=None):
        self._coef(coeff, "coeff")
        self._img(out, "out")
        if th is not None:
            self._th(th)
        thp = ctypes.c_void_p(th.data_ptr()) if th is not None else None
        check(lib.shx_batch_inverse_shear(self._h, ctypes.c_void_p(coeff.data_ptr()), thp,
                                          ctypes.c_void_p(out.data_ptr()), _stream()))
        return out

    def shear_trans_adjoint(self, coeff_bar, x_bar):
        self._coef(coeff_bar, "coeff_bar")
        self._img(x_bar, "x_bar")
        check(lib.shx_batch_shear_trans_adjoint(self._h, ctypes.c_void_p(coeff_bar.data_ptr()),
                                                ctypes.c_void_p(x_bar.data_ptr()), _stream()))
        return x_bar

    def inverse_shear_adjoint(self, x_bar, coeff_bar):
        self._img(x_bar, "x_bar")
        self._coef(coeff_bar, "coeff_bar")
        check(lib.shx_batch_inverse_shear_adjoint(self._h, ctypes.c_void_p(x_bar.data_ptr()),
                                                  ctypes.c_void_p(coeff_bar.data_ptr()), _stream()))
        return coeff_bar

    def _shrink_th(self, th):
        """Row stride of a soft-threshold table: (nbands,) shared by every
        image (stride 0) or (nimg, nbands)."""
        if th.dim() == 1:
            _require_shape(th, (self.shear.nbands,), "th")
            return 0
        self._th(th)
        return self.shear.nbands

    def shrink(self, x, th, out, coeff=None):
        """out = S(soft(A x, th)) with the soft threshold fused into the
        reconstruction's reads (shx_batch_shrink); coeff, if given, receives
        A x for shrink_backward."""
        self._img(x, "x")
        self._img(out, "out")
        if coeff is not None:
            self._coef(coeff, "coeff")
        stride = self._shrink_th(th)
        cptr = ctypes.c_void_p(coeff.data_ptr()) if coeff is not None else None
        check(lib.shx_batch_shrink(self._h, ctypes.c_void_p(x.data_ptr()), cptr, ctypes.c_void_p(th.data_ptr()),
                                   stride, ctypes.c_void_p(out.data_ptr()), _stream()))
        return out

    def shrink_backward(self, coeff, th, g, gx, gth):
        """gx = A^T(1[|c| > th] . S^T g); gth += the threshold gradient
        (shx_batch_shrink_backward; gth has th's shape and is not cleared)."""
        self._coef(coeff, "coeff")
        self._img(g, "g")
        self._img(gx, "gx")
        stride = self._shrink_th(th)
        _require_shape(gth, tuple(th.shape), "gth")
        check(lib.shx_batch_shrink_backward(self._h, ctypes.c_void_p(coeff.data_ptr()),
                                            ctypes.c_void_p(th.data_ptr()), stride, ctypes.c_void_p(g.data_ptr()),
                                            ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(gth.data_ptr()),
                                            _stream()))
        return gx

    def denoise(self, x, coeff, th, out):
        """coeff None: the coefficients are neither returned nor written."""
        self._img(x, "x")
        self._img(out, "out")
        if coeff is not None:
            self._coef(coeff, "coeff")
        self._th(th)
        cptr = ctypes.c_void_p(coeff.data_ptr()) if coeff is not None else None
        check(lib.shx_batch_denoise(self._h, ctypes.c_void_p(x.data_ptr()), cptr,
                                    ctypes.c_void_p(th.data_ptr()), ctypes.c_void_p(out.data_ptr()), _stream()))
        return out


def batch_thresholds(shear, sigmas, scalars, Tscalars=DEFAULT_TSCALARS):
    """(nimg, 1 + sum D) FP32 device table of nsst_HT thresholds, one row per
    image noise level."""
    rows = [band_thresholds(shear, float(s), scalars, Tscalars) for s in np.atleast_1d(sigmas)]
    return torch.from_numpy(np.array(rows, dtype=np.float32)).to(shear.device)


def autotune(shear, batch=None, reps=5):
    """Pick the pyramid tile shape, the reconstruction's directions per column
    block, the column tile widths and the decomposition front end for a plan
    by timing them (outside any graph capture).  Returns the chosen schedule
    as a dict."""
    bh = batch._h if batch is not None else None
    chosen = (ctypes.c_int * 4)()
    check(lib.shx_autotune(shear.handle, bh, int(reps), chosen, _stream()))
    opts = get_plan_options(shear, batch)
    return {"pyramid_variant": chosen[0], "group": chosen[1], "col_tile_dec": chosen[2], "col_tile_rec": chosen[3],
            "dct_front": opts["dct_front"], "dct_back": opts["dct_back"]}


def set_schedule(shear, pyramid_variant, group, batch=None, col_tile_dec=0, col_tile_rec=0):
    check(lib.shx_set_schedule(shear.handle, batch._h if batch is not None else None, int(pyramid_variant),
                               int(group), int(col_tile_dec), int(col_tile_rec)))


def get_plan_options(shear, batch=None):
    """{"dct_front": bool, "dct_back": bool} of a plan (shx_get_plan_options)."""
    f, k = ctypes.c_int(), ctypes.c_int()
    check(lib.shx_get_plan_options(shear.handle, batch._h if batch is not None else None, ctypes.byref(f),
                                   ctypes.byref(k)))
    return {"dct_front": bool(f.value), "dct_back": bool(k.value)}


def set_plan_options(shear, batch=None, dct_front=None, dct_back=None):
    """Plan options (None keeps the current value): dct_front, the
    decomposition's pyramid and scale spectra in the DCT-II domain; dct_back,
    the reconstruction's scale sum and pyramid in the DCT-II domain (L >= 64)."""
    cur = get_plan_options(shear, batch)
    f = cur["dct_front"] if dct_front is None else bool(dct_front)
    k = cur["dct_back"] if dct_back is None else bool(dct_back)
    check(lib.shx_set_plan_options(shear.handle, batch._h if batch is not None else None, int(f), int(k)))


def set_dec_front(shear, dct, batch=None):
    """Decomposition front end of a plan (set_plan_options(dct_front=...))."""
    set_plan_options(shear, batch, dct_front=dct)


def get_dec_front(shear, batch=None):
    return get_plan_options(shear, batch)["dct_front"]


# ---- undecimated wavelet transform -----------------------------------------------

def _rdwt_dims(x):
    if x.dim() == 2:
        return 1, x.shape[0], x.shape[1]
    if x.dim() == 3:
        return x.shape[0], x.shape[1], x.shape[2]
    raise ValueError("expected (m, n) or (nimg, m, n)")


def _h_arg(h):
    h = np.ascontiguousarray(np.asarray(h, dtype=np.float64).ravel())
    return h, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def mrdwt_cuda(x, h, L):
    """[yl, yh] = mrdwt_cuda(x, h, L) (shearcuda/src_cuda/mrdwt_cuda.cpp:58-253);
    x is (m, n) or a batch (nimg, m, n); yh is (..., m, 3*n*L)."""
    _require_gpu(x)
    nimg, m, n = _rdwt_dims(x)
    h, hp = _h_arg(h)
    yl = torch.empty_like(x)
    yh = torch.empty(x.shape[:-1] + (3 * n * L,), device=x.device, dtype=torch.float32)
    work = _workspace(2 * m * n * nimg, x.device)
    check(lib.shx_mrdwt(ctypes.c_void_p(x.data_ptr()), m, n, hp, h.size, int(L), ctypes.c_void_p(yl.data_ptr()),
                        ctypes.c_void_p(yh.data_ptr()), _vp(work), nimg, _stream()))
    return yl, yh


def mirdwt_cuda(yl, yh, h, L):
    """x = mirdwt_cuda(yl, yh, h, L) (shearcuda/src_cuda/mirdwt_cuda.cpp:58-255)."""
    _require_gpu(yl, "yl")
    nimg, m, n = _rdwt_dims(yl)
    _require_shape(yh, yl.shape[:-1] + (3 * n * L,), "yh")
    h, hp = _h_arg(h)
    x = torch.empty_like(yl)
    work = _workspace(2 * m * n * nimg, yl.device)
    check(lib.shx_mirdwt(ctypes.c_void_p(yl.data_ptr()), ctypes.c_void_p(yh.data_ptr()), m, n, hp, h.size, int(L),
                         ctypes.c_void_p(x.data_ptr()), _vp(work), nimg, _stream()))
    return x


def mrdwt_adjoint_cuda(yl_bar, yh_bar, h, L):
    """Transpose of mrdwt: (yl_bar, yh_bar) -> x_bar."""
    _require_gpu(yl_bar, "yl_bar")
    nimg, m, n = _rdwt_dims(yl_bar)
    _require_shape(yh_bar, yl_bar.shape[:-1] + (3 * n * L,), "yh_bar")
    h, hp = _h_arg(h)
    xb = torch.empty_like(yl_bar)
    work = _workspace(2 * m * n * nimg, yl_bar.device)
    check(lib.shx_mrdwt_adjoint(ctypes.c_void_p(yl_bar.data_ptr()), ctypes.c_void_p(yh_bar.data_ptr()), m, n, hp,
                                h.size, int(L), ctypes.c_void_p(xb.data_ptr()), _vp(work), nimg, _stream()))
    return xb


def mirdwt_adjoint_cuda(x_bar, h, L):
    """Transpose of mirdwt: x_bar -> (yl_bar, yh_bar)."""
    _require_gpu(x_bar, "x_bar")
    nimg, m, n = _rdwt_dims(x_bar)
    h, hp = _h_arg(h)
    yl = torch.empty_like(x_bar)
    yh = torch.empty(x_bar.shape[:-1] + (3 * n * L,), device=x_bar.device, dtype=torch.float32)
    work = _workspace(2 * m * n * nimg, x_bar.device)
    check(lib.shx_mirdwt_adjoint(ctypes.c_void_p(x_bar.data_ptr()), m, n, hp, h.size, int(L),
                                 ctypes.c_void_p(yl.data_ptr()), ctypes.c_void_p(yh.data_ptr()), _vp(work), nimg,
                                 _stream()))
    return yl, yh
