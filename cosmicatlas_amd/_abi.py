"""ctypes declarations of the C ABI in include/cosmo_mesh.h (libcosmo_mesh.so).

The library is built in-tree (cosmicatlas_amd/lib/libcosmo_mesh.so) by
``__graft_entry__.build()``.  There is no fallback: if the library is missing,
``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcosmo_mesh.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cosmo_mesh.h")

CM_OK, CM_ERR_INVALID, CM_ERR_UNSUPPORTED, CM_ERR_NOMEM, CM_ERR_DEVICE, CM_ERR_BUDGET, \
    CM_ERR_TIMEOUT, CM_ERR_CANCELLED = range(8)
CM_F32, CM_F64 = 0, 1
CM_NGP, CM_CIC, CM_TSC, CM_PCS = 0, 1, 2, 3
CM_AOS, CM_SOA = 0, 1
CM_MAS_AUTO, CM_MAS_PLAIN, CM_MAS_TILED = 0, 1, 2


class MasDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("L", C.c_double),
        ("xmin", C.c_double * 3),
        ("layout", C.c_int),
        ("pos_dtype", C.c_int),
        ("scheme", C.c_int),
        ("interlace_shift", C.c_double),
        ("algo", C.c_int),
    ]


class PkOpts(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("L", C.c_double),
        ("nshells", C.c_int),
        ("kmin", C.c_double),
        ("dk", C.c_double),
        ("scheme", C.c_int),
        ("mas_correction", C.c_int),
        ("shot_noise", C.c_double),
        ("normal_power", C.c_double),
        ("float_shell_rule", C.c_int),
        ("log_binning", C.c_int),
        ("dlog", C.c_double),
        ("window_set", C.c_int),
    ]


class TwebOpts(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("L", C.c_double),
        ("R_smooth", C.c_double),
        ("lambda_th", C.c_double),
        ("margin_tol", C.c_double),
        ("delta_scale", C.c_double),
        ("hessian", C.c_int),
    ]


class TwebStats(C.Structure):
    _fields_ = [("counts", C.c_int64 * 4), ("n_marginal", C.c_int64), ("margin_tol", C.c_double)]


class MeasureOpts(C.Structure):
    _fields_ = [
        ("mas", MasDesc),
        ("nshells", C.c_int),
        ("kmin", C.c_double),
        ("dk", C.c_double),
        ("mas_correction", C.c_int),
        ("sn_correction", C.c_int),
        ("float_shell_rule", C.c_int),
        ("do_tweb", C.c_int),
        ("R_smooth", C.c_double),
        ("lambda_th", C.c_double),
        ("margin_tol", C.c_double),
        ("fused", C.c_int),
        ("interlace", C.c_int),
        ("log_binning", C.c_int),
        ("dlog", C.c_double),
    ]


class MeasureInfo(C.Structure):
    _fields_ = [
        ("n_obj", C.c_double),
        ("shot_noise", C.c_double),
        ("normal_power", C.c_double),
        ("tweb", TwebStats),
    ]


class Tune(C.Structure):
    _fields_ = [("xnb", C.c_int), ("xe", C.c_int), ("tz", C.c_int), ("tseg", C.c_int), ("dep_wg", C.c_int)]


class ZeldovichOpts(C.Structure):
    _fields_ = [
        ("NL", C.c_int),
        ("L", C.c_double),
        ("A", C.c_double),
        ("ns", C.c_double),
        ("seed", C.c_uint64),
    ]


class FkpOpts(C.Structure):
    _fields_ = [
        ("nbar", C.c_double),
        ("P_est", C.c_double),
        ("fkp_weight", C.c_int),
        ("window", C.c_int),
        ("nbar_g", C.c_void_p),
        ("nbar_r", C.c_void_p),
    ]


class BkdOpts(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("L", C.c_double),
        ("nbins", C.c_int),
        ("kmin", C.c_double),
        ("dk", C.c_double),
        ("scheme", C.c_int),
        ("mas_correction", C.c_int),
        ("sn_correction", C.c_int),
        ("n_obj", C.c_double),
    ]


class BkOpts(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("L", C.c_double),
        ("nshells", C.c_int),
        ("kmin", C.c_double),
        ("dk", C.c_double),
        ("shell_min", C.c_int),
        ("scheme", C.c_int),
        ("mas_correction", C.c_int),
        ("sn_correction", C.c_int),
        ("n_obj", C.c_double),
    ]


class Request(C.Structure):
    _fields_ = [
        ("opts", MeasureOpts),
        ("pos", C.c_void_p),
        ("w", C.c_void_p),
        ("n", C.c_size_t),
        ("gen", ZeldovichOpts),
        ("pk0", C.c_void_p),
        ("pk2", C.c_void_p),
        ("pk4", C.c_void_p),
        ("nmodes", C.c_void_p),
        ("labels", C.c_void_p),
        ("info", C.POINTER(MeasureInfo)),
    ]


class PackedMock(C.Structure):
    _fields_ = [
        ("opts", MeasureOpts),
        ("pos", C.c_void_p),
        ("w", C.c_void_p),
        ("n", C.c_size_t),
        ("pk0", C.c_void_p),
        ("pk2", C.c_void_p),
        ("pk4", C.c_void_p),
        ("nmodes", C.c_void_p),
        ("labels", C.c_void_p),
        ("info", C.POINTER(MeasureInfo)),
    ]


_LIB = None


def _declare(L):
    P = C.c_void_p
    st = C.c_int
    sig = {
        "cm_ctx_create": (st, [C.c_int, C.c_size_t, C.POINTER(P)]),
        "cm_ctx_destroy": (None, [P]),
        "cm_last_error": (C.c_char_p, [P]),
        "cm_status_string": (C.c_char_p, [C.c_int]),
        "cm_set_stream": (st, [P, P]),
        "cm_get_stream": (P, [P]),
        "cm_set_graphs": (st, [P, C.c_int]),
        "cm_graph_counts": (st, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "cm_autotune": (st, [P, C.POINTER(MeasureOpts), P, P, C.c_size_t, P, C.POINTER(Tune),
                             C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "cm_set_tune": (st, [P, C.c_int, C.POINTER(Tune)]),
        "cm_get_tune": (st, [P, C.c_int, C.POINTER(Tune)]),
        "cm_synchronize": (st, [P]),
        "cm_dev_alloc": (st, [P, C.c_size_t, C.POINTER(P)]),
        "cm_dev_free": (st, [P, P]),
        "cm_memset": (st, [P, P, C.c_int, C.c_size_t]),
        "cm_copy_h2d": (st, [P, P, P, C.c_size_t]),
        "cm_copy_d2h": (st, [P, P, P, C.c_size_t]),
        "cm_profile_enable": (st, [P, C.c_int]),
        "cm_kernel_time": (st, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "cm_profile_reset": (st, [P]),
        "cm_mesh_assign": (st, [P, C.POINTER(MasDesc), P, P, C.c_size_t, P, C.c_int, C.c_int]),
        "cm_mesh_sum": (st, [P, P, C.c_int, C.c_size_t, C.POINTER(C.c_double)]),
        "cm_mesh_affine": (st, [P, P, C.c_int, C.c_size_t, C.c_double, C.c_double]),
        "cm_fft_supported": (C.c_int, [C.c_int]),
        "cm_fft_r2c": (st, [P, C.c_int, C.c_int, P, P, C.c_double]),
        "cm_fft_c2r": (st, [P, C.c_int, C.c_int, P, P]),
        "cm_pk_shells": (st, [P, C.POINTER(PkOpts), C.c_int, P, P, P, P, P]),
        "cm_tweb": (st, [P, C.POINTER(TwebOpts), P, P, P, C.POINTER(TwebStats)]),
        "cm_poisson": (st, [P, C.POINTER(TwebOpts), P, P]),
        "cm_tweb_from_phi": (st, [P, C.POINTER(TwebOpts), P, P, P, C.POINTER(TwebStats)]),
        "cm_tweb_f64": (st, [P, C.POINTER(TwebOpts), P, P, P, P, C.POINTER(TwebStats)]),
        "cm_pk_shells_noalloc": (st, [P, C.POINTER(PkOpts), C.c_int, P, P, P, P, P]),
        "cm_measure": (st, [P, C.POINTER(MeasureOpts), P, P, C.c_size_t, P, P, P, P, P,
                            C.POINTER(MeasureInfo)]),
        "cm_zeldovich": (st, [P, C.POINTER(ZeldovichOpts), P]),
        "cm_tweb_tclass": (st, [P, P, C.c_size_t, C.POINTER(C.c_int), C.c_int, P]),
        "cm_bispectrum": (st, [P, C.POINTER(BkOpts), P, C.c_int, P, P, P, P, C.POINTER(C.c_int)]),
        "cm_bispectrum_direct": (st, [P, C.POINTER(BkdOpts), P, P, P, P, P, P, P, P]),
        "cm_measure_fkp": (st, [P, C.POINTER(MeasureOpts), C.POINTER(FkpOpts), P, P, C.c_size_t, P, P, C.c_size_t,
                                P, P, P, P, P, C.POINTER(MeasureInfo)]),
        "cm_request_footprint": (C.c_size_t, [C.POINTER(Request)]),
        "cm_server_create": (st, [C.c_int, C.c_int, C.c_size_t, C.POINTER(P)]),
        "cm_server_submit": (st, [P, C.POINTER(Request), C.POINTER(C.c_int64)]),
        "cm_server_wait": (st, [P, C.c_int64, C.c_double, C.POINTER(C.c_double), C.c_char_p, C.c_size_t]),
        "cm_server_destroy": (st, [P]),
        "cm_server_stats": (st, [P, C.POINTER(C.c_int64), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_size_t), C.POINTER(C.c_int64)]),
        "cm_measure_batch": (st, [P, C.POINTER(Request), C.c_int, P, P]),
        "cm_server_last_error": (st, [P, C.c_char_p, C.c_size_t]),
        "cm_measure_packed": (st, [P, P, C.c_int, P]),
        "cm_uniform": (st, [P, C.c_size_t, C.c_double, C.c_uint64, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def load() -> C.CDLL:
    """Load libcosmo_mesh.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)")
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


def header_functions(path: str = HEADER):
    """Names of the functions declared in the C header."""
    import re
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_[a-z0-9_]+)\s*\(", txt)))
