"""Host-side mirror of the reference's 3D shearlet volume path on GPU tensors.

Mirrors shearcuda3d/: GetMeyerBasedFilter (filters3d.py), DoPyrDec / DoPyrRec
(3DBP/PyrNDDec_mm.m, PyrNDRec_mm.m), ShDec / ShRec, ShDecThRec_cuda
(3DShearTrans/src_cuda/ShDecThRec_cuda.cpp), NsstScalars and the denoise flow of
DenoiseDemo/main.m.  Every transform runs in the HIP library through the C ABI
(include/shx.h, shx3_*); volumes are FP32 tensors on the GPU (no CPU path).

Layouts: a batch of volumes is a (nvol, n, n, n) tensor; a subband block is
(nvol, subs_elems) with the band-pass subbands finest first, then the lowpass;
a band block is (nvol, bands_elems), level-major, cone-major, then the MATLAB
linear index of the (l2, l1) direction cell.
"""
import ctypes

import numpy as np
import torch

from ._lib import check, lib
from .filters3d import FILTER_SIDE, get_meyer_based_filter

# DenoiseDemo/main.m uses level 3 with 8x8 / 4x4 / 4x4 directions; the
# volume workload here keeps 8x8 on the two finer levels and 4x4 on the coarser.
DEFAULT_NUM_DIRS = (8, 8, 4)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu(x, name):
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("%s must be a FP32 tensor on the GPU (no CPU path)" % name)
    if x.dtype != torch.float32:
        raise TypeError("%s: data should be GPUsingle (float32), got %s" % (name, x.dtype))
    if not x.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Shear3D:
    """3D dictionary: band-pass configuration plus the cone filter spectra on
    the GPU (the F cell array of GetMeyerBasedFilter, after deepCopyToGpu)."""

    def __init__(self, n, num_dirs=DEFAULT_NUM_DIRS, pyr_mode=1, filters=None):
        self.n = int(n)
        self.num_dirs = tuple(int(d) for d in num_dirs)
        self.nlevels = len(self.num_dirs)
        self.pyr_mode = pyr_mode
        if filters is None:
            filters = get_meyer_based_filter(self.num_dirs)
        self.fsize = filters[0].shape[-1]
        host = []
        seen = {}
        for f in filters:
            key = id(f)
            if key not in seen:
                seen[key] = np.ascontiguousarray(f, dtype=np.float32)
            host.append(seen[key])
        self._host = host
        ptrs = (ctypes.c_void_p * self.nlevels)(*[a.ctypes.data for a in host])
        nd = (ctypes.c_int * self.nlevels)(*self.num_dirs)
        h = ctypes.c_void_p()
        check(lib.shx3_prepare(self.n, self.nlevels, float(pyr_mode), nd, self.fsize, ptrs, ctypes.byref(h)))
        self._h = h
        sides = (ctypes.c_int * (self.nlevels + 1))()
        fft = (ctypes.c_int * self.nlevels)()
        first = (ctypes.c_int * (self.nlevels + 1))()
        check(lib.shx3_info(self._h, sides, fft, first))
        self.sides = list(sides)
        self.fft_sides = list(fft)
        self.band_first = list(first)
        self.nbands = self.band_first[-1]
        self.subs_elems = int(lib.shx3_subs_elems(self._h))
        self.bands_elems = int(lib.shx3_bands_elems(self._h))

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("3D dictionary already released")
        return self._h

    def nbytes(self):
        return int(lib.shx3_dict_bytes(self.handle))

    def close(self):
        if self._h is not None:
            lib.shx3_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- layout helpers -------------------------------------------------------
    def split_subs(self, subs):
        """Views of one subband block (nvol, subs_elems): list of (nvol, s, s, s)."""
        out, off = [], 0
        for s in self.sides:
            e = s ** 3
            out.append(subs[:, off:off + e].view(subs.shape[0], s, s, s))
            off += e
        return out

    def split_bands(self, bands):
        """Views of a band block: per level a (nvol, nb, s, s, s) tensor."""
        out, off = [], 0
        for l in range(self.nlevels):
            s = self.sides[l]
            nb = self.band_first[l + 1] - self.band_first[l]
            e = nb * s ** 3
            out.append(bands[:, off:off + e].view(bands.shape[0], nb, s, s, s))
            off += e
        return out


class Plan3D:
    """Workspaces for batches of nvol volumes (group = bands per launch)."""

    def __init__(self, shear, nvol, group=8):
        self.shear = shear
        self.nvol = int(nvol)
        self.group = int(group)
        h = ctypes.c_void_p()
        check(lib.shx3_plan_create(shear.handle, self.nvol, self.group, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("3D plan already released")
        return self._h

    def nbytes(self):
        return int(lib.shx3_plan_bytes(self.handle))

    def set_schedule(self, group=None, plane_cols=-1):
        """Bands per launch (at most the plan's group) and the fused plane
        kernel's columns per block (0 automatic, 32, 64; -1, the default,
        separate row and column passes)."""
        check(lib.shx3_set_schedule(self.handle, int(self.group if group is None else group), int(plane_cols)))

    def close(self):
        if self._h is not None:
            lib.shx3_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vols(self, x, name):
        _require_gpu(x, name)
        n = self.shear.n
        if tuple(x.shape) != (self.nvol, n, n, n):
            raise ValueError("%s must be (%d, %d, %d, %d), got %s" % (name, self.nvol, n, n, n, tuple(x.shape)))

    def _block(self, t, elems, name):
        _require_gpu(t, name)
        if tuple(t.shape) != (self.nvol, elems):
            raise ValueError("%s must be (%d, %d), got %s" % (name, self.nvol, elems, tuple(t.shape)))

    def _th(self, th):
        _require_gpu(th, "th")
        if tuple(th.shape) != (self.nvol, self.shear.nbands):
            raise ValueError("th must be (%d, %d) per-band thresholds" % (self.nvol, self.shear.nbands))

    def new_subs(self):
        return torch.zeros(self.nvol, self.shear.subs_elems, device="cuda")

    def new_bands(self):
        return torch.empty(self.nvol, self.shear.bands_elems, device="cuda")

    def pyr_dec(self, x, subs=None):
        """DoPyrDec (3DShearTrans/DoPyrDec.m:1-6) of every volume."""
        self._vols(x, "x")
        subs = self.new_subs() if subs is None else subs
        self._block(subs, self.shear.subs_elems, "subs")
        check(lib.shx3_pyr_dec(self.handle, _ptr(x), _ptr(subs), _stream()))
        return subs

    def pyr_rec(self, subs, out=None):
        """DoPyrRec (3DShearTrans/DoPyrRec.m:1-4)."""
        self._block(subs, self.shear.subs_elems, "subs")
        n = self.shear.n
        out = torch.empty(self.nvol, n, n, n, device="cuda") if out is None else out
        self._vols(out, "out")
        check(lib.shx3_pyr_rec(self.handle, _ptr(subs), _ptr(out), _stream()))
        return out

    def dec_thresh_rec(self, subs, th, rec=None):
        """ShDecThRec_cuda: thresholded band sums of every level (level slots
        of rec; its lowpass slot is left as given)."""
        self._block(subs, self.shear.subs_elems, "subs")
        self._th(th)
        rec = self.new_subs() if rec is None else rec
        self._block(rec, self.shear.subs_elems, "rec")
        check(lib.shx3_dec_thresh_rec(self.handle, _ptr(subs), _ptr(th), _ptr(rec), _stream()))
        return rec

    def analysis(self, subs, bands=None):
        """ShDec for the three cones: every band of every level."""
        self._block(subs, self.shear.subs_elems, "subs")
        bands = self.new_bands() if bands is None else bands
        self._block(bands, self.shear.bands_elems, "bands")
        check(lib.shx3_analysis(self.handle, _ptr(subs), _ptr(bands), _stream()))
        return bands

    def analysis_adjoint(self, bands, subs_bar=None):
        """Transpose of analysis (level slots of subs_bar)."""
        self._block(bands, self.shear.bands_elems, "bands")
        subs_bar = self.new_subs() if subs_bar is None else subs_bar
        self._block(subs_bar, self.shear.subs_elems, "subs_bar")
        check(lib.shx3_analysis_adjoint(self.handle, _ptr(bands), _ptr(subs_bar), _stream()))
        return subs_bar

    def denoise(self, x, th, out=None):
        """DenoiseDemo/main.m:100-150 on every volume."""
        self._vols(x, "x")
        self._th(th)
        n = self.shear.n
        out = torch.empty(self.nvol, n, n, n, device="cuda") if out is None else out
        self._vols(out, "out")
        check(lib.shx3_denoise(self.handle, _ptr(x), _ptr(th), _ptr(out), _stream()))
        return out


def sh_rec(bands_level):
    """ShRec (3DShearTrans/ShRec.m:11-23): plain sum of a level's bands."""
    return bands_level.sum(dim=1)


def nsst_scalars_3d(shear, noise):
    """NsstScalars (DenoiseDemo/NsstScalars.m:1-19) with the noise volume passed
    in: (bandEst in band order, lowPassEst).  Bands and their exact medians
    (MAD/0.6745, shx_band_mad) are computed on the device."""
    from . import api
    plan = Plan3D(shear, 1, group=8)
    try:
        subs = plan.pyr_dec(noise.reshape(1, *noise.shape).contiguous())
        bands = plan.analysis(subs)
        per_level = shear.split_bands(bands)
        est = [api.band_mad_cuda(lvl[0].contiguous()).cpu().numpy() for lvl in per_level]
        low = float(api.band_mad_cuda(shear.split_subs(subs)[-1][0].reshape(1, -1).contiguous()).cpu()[0])
    finally:
        plan.close()
    return np.concatenate(est), low


def thresholds_3d(shear, band_est, sigmas, T):
    """Per-volume, per-band hard thresholds T[l] * sigma * bandEst
    (DenoiseDemo/ThresholdShCoeff.m:7, ShDecThRec_cuda.cpp:172,321) as a
    (nvol, nbands) device tensor."""
    sig = np.atleast_1d(np.asarray(sigmas, dtype=np.float64))
    lev = np.zeros(shear.nbands)
    for l in range(shear.nlevels):
        lev[shear.band_first[l]:shear.band_first[l + 1]] = T[l]
    th = sig[:, None] * (lev * np.asarray(band_est, dtype=np.float64))[None, :]
    return torch.tensor(th, dtype=torch.float32, device="cuda")


def denoise_3d(shear, xn, band_est, sigma, T=(3.0, 3.0, 3.0), plan=None):
    """Denoise a batch (nvol, n, n, n) with one sigma per volume."""
    nvol = xn.shape[0]
    own = plan is None
    plan = Plan3D(shear, nvol) if own else plan
    try:
        th = thresholds_3d(shear, band_est, np.broadcast_to(np.asarray(sigma, dtype=np.float64), (nvol,)), T)
        return plan.denoise(xn, th)
    finally:
        if own:
            plan.close()
