"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP path is compared
against; it is imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` and never by the product package.

Every wrapper names the reference function it restates (paths relative to the
COSMICATLAS source tree); the C file carries the line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NGP, CIC, TSC, PCS = 0, 1, 2, 3


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
    ]


def build(force: bool = False) -> None:
    """Compile liboracle.so / libcpuref.so with the committed Makefile."""
    so = os.path.join(HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        so = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        P = C.c_void_p
        L.or_mas.restype = C.c_size_t
        L.or_mas.argtypes = [C.c_int, C.c_double, P, P, C.c_int, C.c_int, C.c_size_t, P,
                             C.c_int, C.c_double, P]
        L.or_mas_kernel.restype = C.c_double
        L.or_mas_kernel.argtypes = [C.c_int, C.c_double]
        L.or_fft_r2c.argtypes = [C.c_int, P, P]
        L.or_fft_c2r.argtypes = [C.c_int, P, P]
        L.or_fft_1d.argtypes = [C.c_int, P, P, C.c_int]
        L.or_pk_shells.argtypes = [P, C.POINTER(PkOpts), P, P, P, P]
        L.or_window_shells.argtypes = [P, C.POINTER(PkOpts), P, P]
        L.or_shell_index.restype = C.c_int
        L.or_shell_index.argtypes = [C.POINTER(PkOpts), C.c_int, C.c_int, C.c_int]
        L.or_smooth_poisson.argtypes = [C.c_int, C.c_double, C.c_double, P]
        L.or_gradfindif.argtypes = [C.c_int, C.c_double, P, P, C.c_int]
        L.or_eigen3.argtypes = [P, P]
        L.or_tweb.argtypes = [C.c_int, C.c_double, P, C.c_double, C.c_double, P, P, P, P]
        L.or_tweb_from_phi.argtypes = [C.c_int, C.c_double, P, C.c_double, P, P, P]
        L.or_tweb_spectral.argtypes = [C.c_int, C.c_double, P, C.c_double, C.c_double, P, P, P]
        L.or_bk_direct.argtypes = [C.c_int, P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, P, P, P, P, P]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def mas(pos, N, L, scheme, weights=None, xmin=(0.0, 0.0, 0.0), soa=False, interlace_shift=0.0,
        mesh=None):
    """grid_assignment_{NGP,CIC,TSC,PCS} over all particles (Miscelanious.cpp:2597-2868).

    ``pos``: float32 or float64 array, AoS (n,3) or SoA (3,n) when ``soa``.
    Returns the float64 N^3 mesh (accumulated into ``mesh`` if given)."""
    pos = np.ascontiguousarray(pos)
    f32 = pos.dtype == np.float32
    n = pos.shape[1] if soa else pos.shape[0]
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=pos.dtype)
    if mesh is None:
        mesh = np.zeros((N, N, N), np.float64)
    xm = np.asarray(xmin, np.float64)
    lib().or_mas(N, L, _p(xm), _p(pos), int(f32), int(soa), n, _p(weights), scheme,
                 float(interlace_shift), _p(mesh))
    return mesh


def mas_kernel(scheme, s):
    return lib().or_mas_kernel(scheme, float(s))


def fft_r2c(x):
    x = np.ascontiguousarray(x, np.float64)
    N = x.shape[0]
    out = np.empty((N, N, N // 2 + 1), np.complex128)
    lib().or_fft_r2c(N, _p(x), _p(out))
    return out


def fft_c2r(h, N):
    h = np.ascontiguousarray(h, np.complex128)
    out = np.empty((N, N, N), np.float64)
    lib().or_fft_c2r(N, _p(h), _p(out))
    return out


def fft_1d(x, sign=-1):
    x = np.ascontiguousarray(x, np.complex128)
    out = np.empty_like(x)
    lib().or_fft_1d(len(x), _p(x), _p(out), sign)
    return out


def pk_opts(N, L, nshells, kmin, dk, scheme, shot_noise, normal_power, mas_correction=True,
            float_shell_rule=False, log_binning=False, dlog=0.0):
    return PkOpts(N, L, nshells, kmin, dk, scheme, int(mas_correction), shot_noise,
                  normal_power, int(float_shell_rule), int(log_binning), float(dlog))


def pk_shells(dk, opts: PkOpts):
    """power_spectrum_fkp box branch (FftwFunctions.cpp:958-1657)."""
    dk = np.ascontiguousarray(dk, np.complex128)
    ns = opts.nshells
    p0, p2, p4 = (np.zeros(ns) for _ in range(3))
    nm = np.zeros(ns, np.int64)
    lib().or_pk_shells(_p(dk), C.byref(opts), _p(p0), _p(p2), _p(p4), _p(nm))
    return p0, p2, p4, nm


def window_shells(dk, opts: PkOpts):
    """Window shells of power_spectrum_fkp's random branch (zero mode plus the
    three non-first octant copies; FftwFunctions.cpp:1248-1262, 1354-1568)."""
    dk = np.ascontiguousarray(dk, np.complex128)
    pw = np.zeros(opts.nshells)
    nm = np.zeros(opts.nshells, np.int64)
    lib().or_window_shells(_p(dk), C.byref(opts), _p(pw), _p(nm))
    return pw, nm


def shell_index(opts: PkOpts, i, j, k):
    return lib().or_shell_index(C.byref(opts), i, j, k)


def smooth_poisson(dk, N, L, R):
    dk = np.ascontiguousarray(dk, np.complex128).copy()
    lib().or_smooth_poisson(N, L, R, _p(dk))
    return dk


def gradfindif(x, L, dim):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib().or_gradfindif(x.shape[0], L, _p(x), _p(out), dim)
    return out


def eigen3(m9):
    m = np.ascontiguousarray(m9, np.float64)
    ev = np.zeros(3)
    lib().or_eigen3(_p(m), _p(ev))
    return ev


def tweb(delta, L, R, lambda_th, want_lambdas=True, want_phi=False):
    """Cwclass::get_CWC with Gaussian pre-smoothing (Cwclass.cpp:235-894)."""
    delta = np.ascontiguousarray(delta, np.float64)
    N = delta.shape[0]
    labels = np.zeros((N, N, N), np.int8)
    lam = np.zeros((3, N, N, N)) if want_lambdas else None
    phi = np.zeros((N, N, N)) if want_phi else None
    counts = np.zeros(4, np.int64)
    lib().or_tweb(N, L, _p(delta), R, lambda_th, _p(labels), _p(lam), _p(phi), _p(counts))
    return labels, lam, phi, counts


def tweb_spectral(delta, L, R, lambda_th):
    """T-web with the spectral tidal tensor T_ab = c2r(W k_a k_b / k^2 delta_k)
    (EigenValuesTweb _USE_GFFT_EIGENV_ branch, Miscelanious.cpp:1703-1718;
    calc_LapPhiv, NumericalMethods.cpp:2605-2679): (labels, lambdas, counts)."""
    delta = np.ascontiguousarray(delta, np.float64)
    N = delta.shape[0]
    labels = np.zeros((N, N, N), np.int8)
    lam = np.zeros((3, N, N, N))
    counts = np.zeros(4, np.int64)
    lib().or_tweb_spectral(N, L, _p(delta), R, lambda_th, _p(labels), _p(lam), _p(counts))
    return labels, lam, counts


def tweb_from_phi(phi, L, lambda_th):
    phi = np.ascontiguousarray(phi, np.float64)
    N = phi.shape[0]
    labels = np.zeros((N, N, N), np.int8)
    lam = np.zeros((3, N, N, N))
    counts = np.zeros(4, np.int64)
    lib().or_tweb_from_phi(N, L, _p(phi), lambda_th, _p(labels), _p(lam), _p(counts))
    return labels, lam, counts


def bk_direct(dk, Dmin, ndel, nb, scheme=0, mas_correction=False):
    """Brute-force bispectrum sums over ordered pairs of the (N+1)^3 grid
    (FftwFunctions::bispectrum_fkp, FftwFunctions.cpp:3105-3297).  dk: complex
    half spectrum N x N x (N/2+1).  Returns (bsum, snsum, cnt) of shape
    (nb, nb, nb) filled at a <= b <= c, and (psum, pcnt) of shape (nb,)."""
    dk = np.ascontiguousarray(dk, np.complex128)
    N = dk.shape[0]
    bsum = np.zeros((nb, nb, nb))
    snsum = np.zeros((nb, nb, nb))
    cnt = np.zeros((nb, nb, nb), np.int64)
    psum = np.zeros(nb)
    pcnt = np.zeros(nb, np.int64)
    lib().or_bk_direct(N, _p(dk), float(Dmin), float(ndel), nb, int(scheme), int(mas_correction), _p(bsum),
                       _p(snsum), _p(cnt), _p(psum), _p(pcnt))
    return bsum, snsum, cnt, psum, pcnt
