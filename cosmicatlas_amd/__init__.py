"""cosmicatlas_amd -- MI355X (gfx950) mesh-measurement path of COSMICATLAS/BAM.

Python front end of libcosmo_mesh.so (C ABI in include/cosmo_mesh.h).  Device
buffers are torch tensors on ``cuda`` (used only as HIP memory); every compute
call goes through the HIP library.  The C++ host classes that keep the
reference's interface (Params, FftwFunctions, PowerSpectrumF, Cwclass, the
``-p`` CLI) live in cosmicatlas_amd/host and call the same C ABI.

The entry points mirror the reference functions they replace:

=====================================  ==========================================
``Context.mesh_assign``                grid_assignment_* / getDensity_*
                                       (Source/Miscelanious.cpp:2597-2868, 65-815)
``Context.fft_r2c`` / ``fft_c2r``      do_fftw_r2c / do_fftw_c2r
                                       (Source/NumericalMethods.cpp:1718-1856)
``Context.pk_shells``                  FftwFunctions::power_spectrum_fkp
                                       (Source/FftwFunctions.cpp:958-1657)
``Context.poisson``                    PoissonSolver (Source/Miscelanious.cpp:1189)
``Context.tweb`` / ``tweb_from_phi``   Cwclass::get_CWC / EigenValuesTweb
                                       (Source/Cwclass.cpp:235-894)
``Context.measure``                    PowerSpectrumF::compute_power_spectrum +
                                       Cwclass T-web on the same mesh
=====================================  ==========================================
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from ._abi import (CM_AOS, CM_CIC, CM_F32, CM_F64, CM_MAS_AUTO, CM_MAS_PLAIN, CM_MAS_TILED,  # noqa: F401
                   CM_NGP, CM_PCS, CM_SOA, CM_TSC)

SCHEMES = {"NGP": CM_NGP, "CIC": CM_CIC, "TSC": CM_TSC, "PCS": CM_PCS}
ALGOS = {"auto": CM_MAS_AUTO, "plain": CM_MAS_PLAIN, "tiled": CM_MAS_TILED}


class CosmoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def _torch():
    import torch
    return torch


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _scheme(s):
    return SCHEMES[s.upper()] if isinstance(s, str) else int(s)


def fft_supported(N: int) -> bool:
    return bool(_abi.load().cm_fft_supported(int(N)))


class Context:
    """One library context (device, stream, resident buffers)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, use_torch_stream: bool = True):
        self.lib = _abi.load()
        self.device = device
        h = C.c_void_p()
        st = self.lib.cm_ctx_create(device, hbm_budget, C.byref(h))
        if st != 0:
            raise CosmoError(st, f"cm_ctx_create(device={device}) failed: "
                             f"{self.lib.cm_status_string(st).decode()}")
        self.h = h
        if use_torch_stream:
            torch = _torch()
            s = torch.cuda.current_stream(device).cuda_stream
            self._check(self.lib.cm_set_stream(self.h, C.c_void_p(s)))

    # ------------------------------------------------------------------ utils
    def close(self):
        if getattr(self, "h", None):
            self.lib.cm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != 0:
            raise CosmoError(st, self.lib.cm_last_error(self.h).decode())

    def synchronize(self):
        self._check(self.lib.cm_synchronize(self.h))

    def set_graphs(self, on: bool = True):
        """hipGraph replay of repeated fused measurements (on by default)."""
        self._check(self.lib.cm_set_graphs(self.h, int(on)))

    def graph_counts(self):
        """(graphs captured, graph launches) of this context."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.lib.cm_graph_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, pos, N, L, scheme, nshells, kmin, dk_bin, weights=None, do_tweb=False, R_smooth=0.0,
                 lambda_th=0.0, margin_tol=0.0, labels=None, **kw):
        """Tune the fused path's launch shapes for meshes of N on this device
        (process-wide); returns the shapes and the kernel ms per mock before and
        after."""
        torch = _torch()
        o = _measure_opts(N, L, scheme, nshells, kmin, dk_bin, do_tweb=do_tweb, R_smooth=R_smooth,
                          lambda_th=lambda_th, margin_tol=margin_tol, **kw)
        if do_tweb and labels is None:
            labels = torch.empty((N, N, N), dtype=torch.int8, device=pos.device)
        t = _abi.Tune()
        before, after = C.c_double(), C.c_double()
        self._check(self.lib.cm_autotune(self.h, C.byref(o), _ptr(pos), _ptr(weights), pos.shape[0], _ptr(labels),
                                         C.byref(t), C.byref(before), C.byref(after)))
        return {"tune": {f: getattr(t, f) for f, _ in _abi.Tune._fields_}, "ms_before": before.value,
                "ms_after": after.value}

    def set_tune(self, N, tune=None):
        """Set (None: clear) the process-wide launch shapes for meshes of N."""
        if tune is None:
            self._check(self.lib.cm_set_tune(self.h, int(N), None))
            return
        t = _abi.Tune(*[int(tune.get(f, 0)) for f, _ in _abi.Tune._fields_])
        self._check(self.lib.cm_set_tune(self.h, int(N), C.byref(t)))

    def get_tune(self, N):
        t = _abi.Tune()
        self._check(self.lib.cm_get_tune(self.h, int(N), C.byref(t)))
        return {f: getattr(t, f) for f, _ in _abi.Tune._fields_}

    def profile(self, on: bool = True):
        self._check(self.lib.cm_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._check(self.lib.cm_profile_reset(self.h))

    def kernel_time(self, name: str):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(self.lib.cm_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # --------------------------------------------------------- mass assignment
    def mesh_assign(self, pos, N, L, scheme, weights=None, xmin=(0.0, 0.0, 0.0), soa=False,
                    interlace_shift=0.0, algo="auto", mesh=None, mesh_dtype=None, accumulate=False):
        """Deposit particles on an N^3 mesh (device tensors).  ``pos`` is (n,3)
        (AoS) or (3,n) with ``soa``; float32 or float64."""
        torch = _torch()
        assert pos.is_cuda and pos.is_contiguous()
        n = pos.shape[1] if soa else pos.shape[0]
        pdt = CM_F32 if pos.dtype == torch.float32 else CM_F64
        if mesh is None:
            mdt = mesh_dtype or torch.float32
            mesh = torch.empty((N, N, N), dtype=mdt, device=pos.device)
        mcode = CM_F32 if mesh.dtype == torch.float32 else CM_F64
        if weights is not None:
            assert weights.dtype == pos.dtype and weights.is_contiguous() and weights.numel() == n
        d = _abi.MasDesc(N, float(L), (C.c_double * 3)(*xmin), CM_SOA if soa else CM_AOS, pdt,
                         _scheme(scheme), float(interlace_shift), ALGOS[algo])
        self._check(self.lib.cm_mesh_assign(self.h, C.byref(d), _ptr(pos), _ptr(weights), n,
                                            _ptr(mesh), mcode, int(accumulate)))
        return mesh

    def mesh_sum(self, mesh):
        torch = _torch()
        out = C.c_double()
        code = CM_F32 if mesh.dtype == torch.float32 else CM_F64
        self._check(self.lib.cm_mesh_sum(self.h, _ptr(mesh), code, mesh.numel(), C.byref(out)))
        return out.value

    def mesh_affine(self, mesh, scale, shift):
        torch = _torch()
        code = CM_F32 if mesh.dtype == torch.float32 else CM_F64
        self._check(self.lib.cm_mesh_affine(self.h, _ptr(mesh), code, mesh.numel(), float(scale),
                                            float(shift)))
        return mesh

    # -------------------------------------------------------------------- FFT
    def fft_r2c(self, x, subtract=0.0, out=None):
        """Unnormalised forward r2c of an (N,N,N) real tensor -> (N,N,N/2+1) complex."""
        torch = _torch()
        N = x.shape[0]
        cdt = torch.complex64 if x.dtype == torch.float32 else torch.complex128
        if out is None:
            out = torch.empty((N, N, N // 2 + 1), dtype=cdt, device=x.device)
        code = CM_F32 if x.dtype == torch.float32 else CM_F64
        self._check(self.lib.cm_fft_r2c(self.h, N, code, _ptr(x), _ptr(out), float(subtract)))
        return out

    def fft_c2r(self, h, N, out=None):
        """Inverse of a Hermitian half spectrum, divided by N^3 (h is overwritten)."""
        torch = _torch()
        rdt = torch.float32 if h.dtype == torch.complex64 else torch.float64
        if out is None:
            out = torch.empty((N, N, N), dtype=rdt, device=h.device)
        code = CM_F32 if rdt == torch.float32 else CM_F64
        self._check(self.lib.cm_fft_c2r(self.h, N, code, _ptr(h), _ptr(out)))
        return out

    # ------------------------------------------------------------------- P(k)
    def pk_shells(self, dk, N, L, nshells, kmin, dk_bin, scheme, shot_noise, normal_power,
                  mas_correction=True, float_shell_rule=False, log_binning=False, dlog=0.0, window_set=False):
        torch = _torch()
        code = CM_F32 if dk.dtype == torch.complex64 else CM_F64
        o = _abi.PkOpts(N, float(L), nshells, float(kmin), float(dk_bin), _scheme(scheme),
                        int(mas_correction), float(shot_noise), float(normal_power),
                        int(float_shell_rule), int(log_binning), float(dlog), int(window_set))
        p0, p2, p4 = (np.zeros(nshells) for _ in range(3))
        nm = np.zeros(nshells, np.int64)
        self._check(self.lib.cm_pk_shells(self.h, C.byref(o), code, _ptr(dk),
                                          p0.ctypes.data_as(C.c_void_p), p2.ctypes.data_as(C.c_void_p),
                                          p4.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p)))
        return p0, p2, p4, nm

    # ------------------------------------------------------------------ T-web
    def _tweb_opts(self, N, L, R, lambda_th, margin_tol, delta_scale, hessian="fd"):
        if hessian not in ("fd", "spectral"):
            raise ValueError(f"hessian must be 'fd' or 'spectral', not {hessian!r}")
        return _abi.TwebOpts(N, float(L), float(R), float(lambda_th), float(margin_tol),
                             float(delta_scale), 1 if hessian == "spectral" else 0)

    def poisson(self, delta, L, R=0.0, delta_scale=1.0):
        torch = _torch()
        N = delta.shape[0]
        phi = torch.empty_like(delta)
        o = self._tweb_opts(N, L, R, 0.0, 0.0, delta_scale)
        self._check(self.lib.cm_poisson(self.h, C.byref(o), _ptr(delta), _ptr(phi)))
        return phi

    def tweb(self, delta, L, R, lambda_th, margin_tol=0.0, delta_scale=1.0, want_lambdas=False, hessian="fd"):
        """T-web of a float32 delta: (labels, lambdas or None, stats).  hessian:
        "fd" (the 4th-order FD stencil applied twice to phi) or "spectral"
        (T_ab = c2r(k_a k_b / k^2 W(k) delta_k))."""
        torch = _torch()
        N = delta.shape[0]
        labels = torch.empty((N, N, N), dtype=torch.int8, device=delta.device)
        lam = torch.empty((3, N, N, N), dtype=torch.float32, device=delta.device) if want_lambdas else None
        st = _abi.TwebStats()
        o = self._tweb_opts(N, L, R, lambda_th, margin_tol, delta_scale, hessian)
        self._check(self.lib.cm_tweb(self.h, C.byref(o), _ptr(delta), _ptr(labels), _ptr(lam), C.byref(st)))
        return labels, lam, _tweb_stats(st)

    def tweb_f64(self, delta, L, R, lambda_th, margin_tol=0.0, delta_scale=1.0, want_lambdas=False,
                 want_phi=False, hessian="fd"):
        """FP64 T-web of a float64 delta: (labels, lambdas or None, phi or None, stats).
        want_phi needs hessian="fd" (the spectral Hessian forms no potential)."""
        torch = _torch()
        N = delta.shape[0]
        labels = torch.empty((N, N, N), dtype=torch.int8, device=delta.device)
        lam = torch.empty((3, N, N, N), dtype=torch.float64, device=delta.device) if want_lambdas else None
        phi = torch.empty((N, N, N), dtype=torch.float64, device=delta.device) if want_phi else None
        st = _abi.TwebStats()
        o = self._tweb_opts(N, L, R, lambda_th, margin_tol, delta_scale, hessian)
        self._check(self.lib.cm_tweb_f64(self.h, C.byref(o), _ptr(delta), _ptr(labels), _ptr(lam), _ptr(phi),
                                         C.byref(st)))
        return labels, lam, phi, _tweb_stats(st)

    def tweb_from_phi(self, phi, L, lambda_th, margin_tol=0.0, want_lambdas=False):
        torch = _torch()
        N = phi.shape[0]
        labels = torch.empty((N, N, N), dtype=torch.int8, device=phi.device)
        lam = torch.empty((3, N, N, N), dtype=torch.float32, device=phi.device) if want_lambdas else None
        st = _abi.TwebStats()
        o = self._tweb_opts(N, L, 0.0, lambda_th, margin_tol, 1.0)
        self._check(self.lib.cm_tweb_from_phi(self.h, C.byref(o), _ptr(phi), _ptr(labels), _ptr(lam),
                                              C.byref(st)))
        return labels, lam, _tweb_stats(st)

    def measure_fkp(self, pos, rand, N, L, scheme, nshells, kmin, dk_bin, weights=None, rweights=None, nbar=0.0,
                    P_est=0.0, fkp_weight=False, window=False, mas_correction=True, sn_correction=True,
                    float_shell_rule=False, nbar_g=None, nbar_r=None, log_binning=False, dlog=0.0):
        """FKP estimator with a random catalogue ``rand`` ((nr,3) float32);
        ``nbar_g`` / ``nbar_r`` (device float32, one per object) give each
        object's mean density in place of the constant ``nbar``."""
        o = _measure_opts(N, L, scheme, nshells, kmin, dk_bin, mas_correction=mas_correction,
                          sn_correction=sn_correction, float_shell_rule=float_shell_rule, log_binning=log_binning,
                          dlog=dlog)
        torch = _torch()
        for a, m in ((nbar_g, pos), (nbar_r, rand)):
            if a is not None and (a.dtype != torch.float32 or not a.is_contiguous() or a.numel() != m.shape[0]):
                raise ValueError("nbar_g / nbar_r: contiguous float32, one value per object")
        f = _abi.FkpOpts(float(nbar), float(P_est), int(fkp_weight), int(window),                          nbar_g.data_ptr() if nbar_g is not None else None,
                         nbar_r.data_ptr() if nbar_r is not None else None)
        p0, p2, p4, pw = (np.zeros(nshells) for _ in range(4))
        nm = np.zeros(nshells, np.int64)
        info = _abi.MeasureInfo()
        self._check(self.lib.cm_measure_fkp(self.h, C.byref(o), C.byref(f), _ptr(pos), _ptr(weights), pos.shape[0],
                                            _ptr(rand), _ptr(rweights), rand.shape[0],
                                            p0.ctypes.data_as(C.c_void_p), p2.ctypes.data_as(C.c_void_p),
                                            p4.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p),
                                            pw.ctypes.data_as(C.c_void_p) if window else None, C.byref(info)))
        return _fkp_result(p0, p2, p4, nm, pw if window else None, info)

    def bispectrum(self, dk, N, L, nshells, kmin, dk_bin, scheme, n_obj, shell_min=0, mas_correction=True,
                   sn_correction=True):
        """Bispectrum of the triangle bins from a complex64 half spectrum; returns
        dict with ijk (nbins,3), bk, sn, ntri."""
        o = _abi.BkOpts(N, float(L), nshells, float(kmin), float(dk_bin), shell_min, _scheme(scheme),
                        int(mas_correction), int(sn_correction), float(n_obj))
        cap = 8 * 256
        ijk = np.zeros((cap, 3), np.int32)
        bk, sn, nt = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        nb = C.c_int()
        self._check(self.lib.cm_bispectrum(self.h, C.byref(o), _ptr(dk), cap, ijk.ctypes.data_as(C.c_void_p),
                                           bk.ctypes.data_as(C.c_void_p), sn.ctypes.data_as(C.c_void_p),
                                           nt.ctypes.data_as(C.c_void_p), C.byref(nb)))
        n = nb.value
        return {"ijk": ijk[:n], "bk": bk[:n], "sn": sn[:n], "ntri": nt[:n]}

    def bispectrum_direct(self, dk, N, L, nbins, kmin, dk_bin, scheme, n_obj, mas_correction=True,
                          sn_correction=True):
        """Brute-force bispectrum over ordered pairs of the (N+1)^3 mode grid
        (FftwFunctions::get_bispectrum_fkp); returns dict of (nbins,)*3 arrays bk,
        sn, ntri, bsum, snsum (filled at a <= b <= c) and (nbins,) arrays pk, nk."""
        o = _abi.BkdOpts(N, float(L), nbins, float(kmin), float(dk_bin), _scheme(scheme), int(mas_correction),
                         int(sn_correction), float(n_obj))
        sh = (nbins, nbins, nbins)
        bk, sn, bs, ss = np.zeros(sh), np.zeros(sh), np.zeros(sh), np.zeros(sh)
        nt = np.zeros(sh, np.int64)
        pk, nk = np.zeros(nbins), np.zeros(nbins, np.int64)
        a = [x.ctypes.data_as(C.c_void_p) for x in (bk, sn, nt, pk, nk, bs, ss)]
        self._check(self.lib.cm_bispectrum_direct(self.h, C.byref(o), _ptr(dk), *a))
        return {"bk": bk, "sn": sn, "ntri": nt, "pk": pk, "nk": nk, "bsum": bs, "snsum": ss}

    def tweb_tclass(self, labels, cwt_used, out=None):
        """Cwclass::get_Tclassification for every cell: int32 tensor like labels."""
        torch = _torch()
        if out is None:
            out = torch.empty(labels.shape, dtype=torch.int32, device=labels.device)
        codes = (C.c_int * len(cwt_used))(*[int(c) for c in cwt_used])
        self._check(self.lib.cm_tweb_tclass(self.h, _ptr(labels), labels.numel(), codes, len(cwt_used), _ptr(out)))
        return out

    # ------------------------------------------------------- synthetic mocks
    def zeldovich(self, NL, L, A=1e4, ns=-1.5, seed=1, out=None):
        """NL^3 Zel'dovich particles (float32, (NL^3, 3), device) from a seeded
        fixed-amplitude Gaussian field with P(k) = A k^ns."""
        torch = _torch()
        if out is None:
            out = torch.empty((NL ** 3, 3), dtype=torch.float32, device=f"cuda:{self.device}")
        o = _abi.ZeldovichOpts(int(NL), float(L), float(A), float(ns), int(seed))
        self._check(self.lib.cm_zeldovich(self.h, C.byref(o), _ptr(out)))
        return out

    def uniform(self, n, L, seed=1, out=None):
        """n uniform positions in [0, L)^3 (float32, (n, 3), device)."""
        torch = _torch()
        if out is None:
            out = torch.empty((n, 3), dtype=torch.float32, device=f"cuda:{self.device}")
        self._check(self.lib.cm_uniform(self.h, int(n), float(L), int(seed), _ptr(out)))
        return out

    # ---------------------------------------------------------- whole mock
    def measure(self, pos, N, L, scheme, nshells, kmin, dk_bin, weights=None, mas_correction=True,
                sn_correction=True, float_shell_rule=False, do_tweb=False, R_smooth=0.0,
                lambda_th=0.0, margin_tol=0.0, fused=True, interlace_shift=0.0, labels=None, interlace=False,
                log_binning=False, dlog=0.0):
        torch = _torch()
        n = pos.shape[0]
        o = _measure_opts(N, L, scheme, nshells, kmin, dk_bin, mas_correction=mas_correction,
                          sn_correction=sn_correction, float_shell_rule=float_shell_rule, do_tweb=do_tweb,
                          R_smooth=R_smooth, lambda_th=lambda_th, margin_tol=margin_tol, fused=fused,
                          interlace_shift=interlace_shift, interlace=interlace, log_binning=log_binning,
                          dlog=dlog)
        p0, p2, p4 = (np.zeros(nshells) for _ in range(3))
        nm = np.zeros(nshells, np.int64)
        if do_tweb and labels is None:
            labels = torch.empty((N, N, N), dtype=torch.int8, device=pos.device)
        info = _abi.MeasureInfo()
        self._check(self.lib.cm_measure(self.h, C.byref(o), _ptr(pos), _ptr(weights), n,
                                        p0.ctypes.data_as(C.c_void_p), p2.ctypes.data_as(C.c_void_p),
                                        p4.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p),
                                        _ptr(labels), C.byref(info)))
        return {"pk0": p0, "pk2": p2, "pk4": p4, "nmodes": nm, "labels": labels,
                "n_obj": info.n_obj, "shot_noise": info.shot_noise, "normal_power": info.normal_power,
                "tweb_counts": list(info.tweb.counts), "tweb_marginal": info.tweb.n_marginal,
                "tweb_tol": info.tweb.margin_tol}

    def measure_packed(self, mocks):
        """cm_measure_packed over a list of mocks, each a dict with the
        arguments of measure() (pos, N, L, scheme, nshells, kmin, dk_bin and the
        keywords).  Returns one result dict per mock, as measure() does; a
        failed mock raises after the batch with its index."""
        torch = _torch()
        arr = (_abi.PackedMock * len(mocks))()
        outs = []
        for i, m in enumerate(mocks):
            m = dict(m)
            pos, N, nshells = m.pop("pos"), m.pop("N"), m.pop("nshells")
            weights, labels = m.pop("weights", None), m.pop("labels", None)
            L, scheme, kmin, dk_bin = m.pop("L"), m.pop("scheme"), m.pop("kmin"), m.pop("dk_bin")
            o = _measure_opts(N, L, scheme, nshells, kmin, dk_bin, **m)
            if o.do_tweb and labels is None:
                labels = torch.empty((N, N, N), dtype=torch.int8, device=pos.device)
            out = {"pk0": np.zeros(nshells), "pk2": np.zeros(nshells), "pk4": np.zeros(nshells),
                   "nmodes": np.zeros(nshells, np.int64), "labels": labels, "_info": _abi.MeasureInfo(),
                   "_keep": (pos, weights)}
            arr[i] = _abi.PackedMock(o, _ptr(pos), _ptr(weights), pos.shape[0],
                                     out["pk0"].ctypes.data_as(C.c_void_p), out["pk2"].ctypes.data_as(C.c_void_p),
                                     out["pk4"].ctypes.data_as(C.c_void_p), out["nmodes"].ctypes.data_as(C.c_void_p),
                                     _ptr(labels), C.pointer(out["_info"]))
            outs.append(out)
        sts = np.zeros(len(mocks), np.int32)
        self.lib.cm_measure_packed(self.h, arr, len(mocks), sts.ctypes.data_as(C.c_void_p))
        res = []
        for i, (out, s_) in enumerate(zip(outs, sts)):
            if s_ != 0:
                raise CosmoError(int(s_), f"packed mock {i}: {self.lib.cm_last_error(self.h).decode()}")
            info = out.pop("_info")
            out.pop("_keep")
            out.update({"n_obj": info.n_obj, "shot_noise": info.shot_noise, "normal_power": info.normal_power,
                        "tweb_counts": list(info.tweb.counts), "tweb_marginal": info.tweb.n_marginal,
                        "tweb_tol": info.tweb.margin_tol})
            res.append(out)
        return res


def _tweb_stats(st):
    return {"counts": list(st.counts), "n_marginal": st.n_marginal, "margin_tol": st.margin_tol}


# margin_tol asking for the rounding tolerance K(N) u max|phi| / d^2
TOL_ROUNDING = -1.0


def _fkp_result(p0, p2, p4, nm, pw, info):
    return {"pk0": p0, "pk2": p2, "pk4": p4, "nmodes": nm, "pk_window": pw, "W_g": info.n_obj,
            "shot_noise": info.shot_noise, "normal_power": info.normal_power}


def _measure_opts(N, L, scheme, nshells, kmin, dk_bin, mas_correction=True, sn_correction=True,
                  float_shell_rule=False, do_tweb=False, R_smooth=0.0, lambda_th=0.0, margin_tol=0.0, fused=True,
                  interlace_shift=0.0, interlace=False, log_binning=False, dlog=0.0):
    md = _abi.MasDesc(N, float(L), (C.c_double * 3)(0.0, 0.0, 0.0), CM_AOS, CM_F32, _scheme(scheme),
                      float(interlace_shift), CM_MAS_AUTO)
    return _abi.MeasureOpts(md, nshells, float(kmin), float(dk_bin), int(mas_correction), int(sn_correction),
                            int(float_shell_rule), int(do_tweb), float(R_smooth), float(lambda_th),
                            float(margin_tol), int(fused), int(interlace), int(log_binning), float(dlog))


class Ticket:
    """A submitted request and its host/device result buffers."""

    def __init__(self, server, req, keep, out):
        self.server, self.req, self._keep, self.out = server, req, keep, out
        self.id = None

    def result(self, timeout=0.0):
        lat = C.c_double()
        err = C.create_string_buffer(512)
        st = self.server.lib.cm_server_wait(self.server.h, self.id, float(timeout), C.byref(lat), err, 512)
        if st != _abi.CM_ERR_TIMEOUT:  # the job is finished: its buffers are no longer written
            self.server._live.pop(self.id, None)
        if st != 0:
            raise CosmoError(st, err.value.decode() or self.server.lib.cm_status_string(st).decode())
        info = self.out["_info"]
        self.out.update({"n_obj": info.n_obj, "shot_noise": info.shot_noise, "normal_power": info.normal_power,
                         "tweb_counts": list(info.tweb.counts), "tweb_marginal": info.tweb.n_marginal,
                         "latency_ms": lat.value})
        return self.out


class Server:
    """Continuous-batching measurement server (cm_server_*): `slots` requests
    run concurrently, each on its own stream, admitted by HBM footprint."""

    def __init__(self, device=0, slots=2, hbm_budget=0):
        self.lib = _abi.load()
        self.device = device
        # submitted tickets whose results have not been collected: the server
        # writes into their host and device buffers until the job finishes, so
        # they live until result() or close(), whatever the caller keeps
        self._live = {}
        h = C.c_void_p()
        st = self.lib.cm_server_create(device, slots, hbm_budget, C.byref(h))
        if st != 0:
            raise CosmoError(st, f"cm_server_create failed: {self.lib.cm_status_string(st).decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.cm_server_destroy(self.h)  # cancels queued jobs, waits for running ones
            self.h = None
        getattr(self, "_live", {}).clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stats(self):
        """Admission counters: evictions, bytes evicted, peak held, budget, and
        the requests re-run alone after outgrowing their reservation."""
        e, eb, ph, b, rr = C.c_int64(), C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_int64()
        st = self.lib.cm_server_stats(self.h, C.byref(e), C.byref(eb), C.byref(ph), C.byref(b), C.byref(rr))
        if st != 0:
            raise CosmoError(st, "cm_server_stats failed")
        return {"evictions": e.value, "evicted_bytes": eb.value, "peak_held": ph.value, "budget": b.value,
                "reruns": rr.value}

    def request(self, N, L, scheme, nshells, kmin, dk_bin, pos=None, gen=None, labels=None, **kw):
        """Build a request: device particles `pos` ((n,3) float32) or
        gen=dict(NL=, A=, ns=, seed=) for a GPU-generated mock in the same box."""
        torch = _torch()
        o = _measure_opts(N, L, scheme, nshells, kmin, dk_bin, **kw)
        out = {"pk0": np.zeros(nshells), "pk2": np.zeros(nshells), "pk4": np.zeros(nshells),
               "nmodes": np.zeros(nshells, np.int64), "_info": _abi.MeasureInfo()}
        if o.do_tweb and labels is None:
            labels = torch.empty((N, N, N), dtype=torch.int8, device=f"cuda:{self.device}")
        out["labels"] = labels
        g = _abi.ZeldovichOpts(0, float(L), 0.0, 0.0, 0)
        if gen is not None:
            g = _abi.ZeldovichOpts(int(gen["NL"]), float(L), float(gen.get("A", 1e4)), float(gen.get("ns", -1.5)),
                                   int(gen.get("seed", 1)))
        req = _abi.Request(o, _ptr(pos), None, 0 if pos is None else int(pos.shape[0]), g,
                           out["pk0"].ctypes.data_as(C.c_void_p), out["pk2"].ctypes.data_as(C.c_void_p),
                           out["pk4"].ctypes.data_as(C.c_void_p), out["nmodes"].ctypes.data_as(C.c_void_p),
                           _ptr(labels), C.pointer(out["_info"]))
        return Ticket(self, req, (pos, labels), out)

    def footprint(self, ticket):
        return int(self.lib.cm_request_footprint(C.byref(ticket.req)))

    def submit(self, ticket):
        t = C.c_int64()
        st = self.lib.cm_server_submit(self.h, C.byref(ticket.req), C.byref(t))
        if st != 0:
            raise CosmoError(st, f"cm_server_submit: {self.lib.cm_status_string(st).decode()}")
        ticket.id = t.value
        self._live[ticket.id] = ticket
        return ticket

    def batch(self, tickets):
        """Run tickets with cm_measure_batch; returns their results."""
        n = len(tickets)
        arr = (_abi.Request * n)(*[t.req for t in tickets])
        sts = np.zeros(n, np.int32)
        lat = np.zeros(n)
        self.lib.cm_measure_batch(self.h, arr, n, sts.ctypes.data_as(C.c_void_p), lat.ctypes.data_as(C.c_void_p))
        res = []
        for t, s_, l_ in zip(tickets, sts, lat):
            if s_ != 0:
                msg = C.create_string_buffer(512)
                self.lib.cm_server_last_error(self.h, msg, 512)
                raise CosmoError(int(s_), f"batch request failed: {self.lib.cm_status_string(int(s_)).decode()}"
                                          f" ({msg.value.decode()})")
            info = t.out["_info"]
            t.out.update({"n_obj": info.n_obj, "shot_noise": info.shot_noise, "normal_power": info.normal_power,
                          "tweb_counts": list(info.tweb.counts), "tweb_marginal": info.tweb.n_marginal,
                          "latency_ms": float(l_)})
            res.append(t.out)
        return res


def binning(N, L, ndel=1, DeltaKmin=0.0):
    """Params::derived_pars binning (Source/Params.cpp:4091-4104): k_f=2pi/L,
    kmin = DeltaKmin*k_f, dk = ndel*k_f, N_shells = N/ndel/2."""
    kf = 2.0 * math.pi / L
    return {"kmin": DeltaKmin * kf, "dk": ndel * kf, "nshells": N // ndel // 2, "kf": kf}


def binning_log(N, L, DeltaKmin, N_log_bins):
    """Params::derived_pars with type_of_binning = log (Source/Params.cpp:4099-4106):
    kmin = DeltaKmin*k_f, kmax = N/2 k_f, Deltal = log10(kmax/kmin)/N_log_bins,
    N_log_bins shells; the k of shell s is kmin 10^((s - 1/2) Deltal)."""
    kf = 2.0 * math.pi / L
    kmin = DeltaKmin * kf
    dlog = math.log10((N / 2) * kf / kmin) / N_log_bins
    return {"kmin": kmin, "dk": 0.0, "nshells": N_log_bins, "kf": kf, "dlog": dlog, "log_binning": True,
            "k": [kmin * 10.0 ** ((s - 0.5) * dlog) for s in range(N_log_bins)]}
