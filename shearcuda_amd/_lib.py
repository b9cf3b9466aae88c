"""ctypes binding of the in-tree HIP library ``libshx.so`` (C ABI: include/shx.h).

There is no fallback: if the library is missing the import fails, naming the
build command.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libshx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "shearcuda_amd: HIP library %s is not built; run `python -m shearcuda_amd.build` "
        "(or __graft_entry__.build())" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

c_int = ctypes.c_int
c_double = ctypes.c_double
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_int64 = ctypes.c_int64
c_void_p = ctypes.c_void_p
P = ctypes.POINTER

# (name, restype, argtypes)
_SIGNATURES = [
    ("shx_last_error", ctypes.c_char_p, []),
    ("shx_version", c_int, []),
    ("shx_prepare_shearing_filters", c_int, [c_int, c_int, P(c_int), P(c_int), P(P(c_double)), c_int, P(c_void_p)]),
    ("shx_prepare_shearing_filters_ex", c_int, [c_int, c_int, P(c_int), P(c_int), P(P(c_double)), c_int, c_int,
                                                P(c_void_p)]),
    ("shx_shearing_filters_myer", c_int, [c_int, c_int, P(c_int), P(c_int), c_int, c_int, P(c_void_p)]),
    ("shx_meyer_filters", c_int, [c_int, c_int, c_void_p, c_void_p]),
    ("shx_clear_shearing_filters", c_int, [c_void_p]),
    ("shx_dict_info", c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    ("shx_dict_bytes", c_size_t, [c_void_p]),
    ("shx_shear_trans", c_int, [c_void_p, c_void_p, P(c_void_p), c_void_p]),
    ("shx_inverse_shear", c_int, [c_void_p, P(c_void_p), c_void_p, c_void_p]),
    ("shx_inverse_shear_thresh", c_int, [c_void_p, P(c_void_p), P(c_float), c_void_p, c_void_p]),
    ("shx_denoise", c_int, [c_void_p, c_void_p, P(c_void_p), c_double, P(c_double), P(c_double), c_void_p,
                            c_void_p]),
    ("shx_denoise_profile", c_int, [c_void_p, c_void_p, P(c_void_p), c_double, P(c_double), P(c_double), c_void_p,
                                    c_int, c_int, P(c_double), P(c_double), ctypes.c_char_p, c_int, P(c_int),
                                    c_void_p]),
    ("shx_batch_create", c_int, [c_void_p, c_int, P(c_void_p)]),
    ("shx_batch_destroy", c_int, [c_void_p]),
    ("shx_batch_bytes", c_size_t, [c_void_p]),
    ("shx_batch_shear_trans", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx_batch_inverse_shear", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx_batch_denoise", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx_shear_trans_adjoint", c_int, [c_void_p, P(c_void_p), c_void_p, c_void_p]),
    ("shx_inverse_shear_adjoint", c_int, [c_void_p, c_void_p, P(c_void_p), c_void_p]),
    ("shx_batch_shear_trans_adjoint", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx_batch_inverse_shear_adjoint", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx_batch_shrink", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("shx_batch_shrink_backward", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                          c_void_p]),
    ("shx_autotune", c_int, [c_void_p, c_void_p, c_int, P(c_int), c_void_p]),
    ("shx_set_schedule", c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int]),
    ("shx_set_plan_options", c_int, [c_void_p, c_void_p, c_int, c_int]),
    ("shx_get_plan_options", c_int, [c_void_p, c_void_p, P(c_int), P(c_int)]),
    ("shx_capture_begin", c_int, [c_void_p]),
    ("shx_graph_parallel", c_int, [P(c_void_p), c_int, P(c_void_p)]),
    ("shx_capture_end", c_int, [c_void_p, P(c_void_p)]),
    ("shx_graph_launch", c_int, [c_void_p, c_void_p]),
    ("shx_graph_destroy", c_int, [c_void_p]),
    ("shx_denoise_graph", c_int, [c_void_p, c_void_p, P(c_void_p), c_double, P(c_double), P(c_double), c_void_p,
                                  P(c_void_p)]),
    ("shx_thresh", c_int, [P(c_void_p), c_int, P(c_int), P(c_int64), c_double, c_int, P(c_double), c_int, c_int,
                           P(c_double), c_int, c_int, P(c_double), c_void_p]),
    ("shx_atrousdec", c_int, [c_void_p, c_int, c_int, P(c_void_p), c_void_p, c_void_p]),
    ("shx_atrousdec_noalloc", c_int, [c_void_p, c_int, c_int, P(c_void_p), c_void_p]),
    ("shx_atrousrec", c_int, [P(c_void_p), c_int, c_int, c_void_p, c_void_p, c_void_p]),
    ("shx_re_weight_band", c_int, [c_void_p, c_int, P(c_double), c_int, c_void_p, c_void_p, c_void_p]),
    ("shx_mrdwt", c_int, [c_void_p, c_int, c_int, P(c_double), c_int, c_int, c_void_p, c_void_p, c_void_p, c_int,
                          c_void_p]),
    ("shx_mirdwt", c_int, [c_void_p, c_void_p, c_int, c_int, P(c_double), c_int, c_int, c_void_p, c_void_p, c_int,
                           c_void_p]),
    ("shx_mrdwt_adjoint", c_int, [c_void_p, c_void_p, c_int, c_int, P(c_double), c_int, c_int, c_void_p, c_void_p,
                                  c_int, c_void_p]),
    ("shx_mirdwt_adjoint", c_int, [c_void_p, c_int, c_int, P(c_double), c_int, c_int, c_void_p, c_void_p, c_void_p,
                                   c_int, c_void_p]),
    ("shx_soft_bands", c_int, [c_void_p, c_void_p, c_int, c_int64, c_int64, c_int64, c_int, c_void_p, c_void_p]),
    ("shx_soft_bands_backward", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int64, c_int64, c_int64,
                                        c_int, c_void_p, c_void_p]),
    ("shx_hard_bands", c_int, [c_void_p, c_void_p, c_int, c_int64, c_int64, c_int64, c_int, c_void_p, c_void_p]),
    ("shx_hard_bands_backward", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64, c_int64, c_int64, c_int,
                                        c_void_p, c_void_p]),
    ("shx_norm_pow", c_int, [c_void_p, c_int64, c_double, P(c_double), c_void_p]),
    ("shx_softth", c_int, [c_void_p, c_int64, c_double, c_void_p]),
    ("shx_band_mad", c_int, [c_void_p, ctypes.c_longlong, ctypes.c_longlong, c_int, c_int, c_void_p, c_void_p]),
    ("shx_nsst_scalars", c_int, [c_void_p, c_void_p, c_int, P(c_double), c_void_p]),
    ("shx_separate", c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_double, P(c_double), c_int, P(c_double),
                             c_int, c_double, P(c_double), c_int, c_int, c_void_p, c_void_p, P(c_double), c_void_p]),
    ("shx_maxflat_factor_error", c_double, []),
    ("shx_maxflat_filter", c_int, [c_int, P(c_double), P(c_int)]),
    ("shx_norm_pow_diff", c_int, [c_void_p, c_void_p, c_int64, c_double, P(c_double), c_void_p]),
    ("shx_add_noise", c_int, [c_void_p, c_int64, c_double, ctypes.c_uint64, c_void_p]),
    ("shx_convolution", c_int, [c_void_p, P(c_void_p), P(c_void_p), c_void_p]),
    ("shx_applyfilters", c_int, [c_void_p, P(c_void_p), P(c_void_p), c_void_p]),
    ("shx_convert", c_int, [c_void_p, c_int, c_void_p, c_int, c_int64, c_void_p]),
    ("shx3_prepare", c_int, [c_int, c_int, c_double, P(c_int), c_int, P(c_void_p), P(c_void_p)]),
    ("shx3_destroy", c_int, [c_void_p]),
    ("shx3_info", c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    ("shx3_subs_elems", ctypes.c_longlong, [c_void_p]),
    ("shx3_bands_elems", ctypes.c_longlong, [c_void_p]),
    ("shx3_dict_bytes", c_size_t, [c_void_p]),
    ("shx3_plan_create", c_int, [c_void_p, c_int, c_int, P(c_void_p)]),
    ("shx3_plan_destroy", c_int, [c_void_p]),
    ("shx3_plan_bytes", c_size_t, [c_void_p]),
    ("shx3_set_schedule", c_int, [c_void_p, c_int, c_int]),
    ("shx3_pyr_dec", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx3_pyr_rec", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx3_dec_thresh_rec", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx3_analysis", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx3_analysis_adjoint", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("shx3_denoise", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
]

EXPORTED = [s[0] for s in _SIGNATURES]

for _name, _res, _args in _SIGNATURES:
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class ShxError(RuntimeError):
    """A non-zero status from the C ABI; the message is shx_last_error()."""

    def __init__(self, status, msg):
        super().__init__("shx error %d: %s" % (status, msg))
        self.status = status


def check(status):
    if status != 0:
        raise ShxError(status, lib.shx_last_error().decode(errors="replace"))
    return status


def ptr_array(ptrs):
    arr = (c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr
