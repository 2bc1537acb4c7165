"""cm_pk_shells_noalloc (shells with a fixed device work array, for callers that
cannot allocate) against cm_pk_shells through the C ABI: two calls in a row,
the second with a smaller input, for both dtypes and the linear, log and
window binnings; sizes beyond the fixed array are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _opts(N, L, ns, kmin, dk, scheme, sn, normal, log=False, dlog=0.0, window=False):
    from cosmicatlas_amd import _abi
    return _abi.PkOpts(N, float(L), ns, float(kmin), float(dk), scheme, 1, float(sn), float(normal), 0,
                       int(log), float(dlog), int(window))


def _call(ctx, fn, o, dk):
    from cosmicatlas_amd import CM_F32, CM_F64
    ns = o.nshells
    p0, p2, p4 = (np.zeros(ns) for _ in range(3))
    nm = np.zeros(ns, np.int64)
    code = CM_F32 if dk.dtype == torch.complex64 else CM_F64
    st = getattr(ctx.lib, fn)(ctx.h, C.byref(o), code, C.c_void_p(dk.data_ptr()), p0.ctypes.data_as(C.c_void_p),
                              p2.ctypes.data_as(C.c_void_p), p4.ctypes.data_as(C.c_void_p),
                              nm.ctypes.data_as(C.c_void_p))
    return st, (p0, p2, p4, nm)


def _spectrum(N, seed, dtype):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    return torch.from_numpy(a).to(dtype).cuda()


CASES = [
    # (N, ns, kmin/kf, dk/kf, scheme, log, window)
    (96, 48, 0.0, 1.0, 2, False, False),
    (64, 16, 1.0, 2.0, 1, False, False),
    (128, 8, 1.0, 0.0, 3, True, False),
    (48, 24, 0.0, 1.0, 0, False, True),
]


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_noalloc_equals_pk_shells(dtype):
    import cosmicatlas_amd as cm
    L = 300.0
    kf = 2 * np.pi / L
    with cm.Context(0) as ctx:
        # consecutive calls: larger then smaller inputs, over every case twice
        for rep in range(2):
            for (N, ns, dmin, ndel, scheme, log, window) in CASES:
                dlog = np.log10((N / 2) / dmin) / ns if log else 0.0
                o = _opts(N, L, ns, dmin * kf, ndel * kf, scheme, 0.0 if window else 1.7, 2.3, log, dlog, window)
                dk = _spectrum(N, 100 * rep + N, dtype)
                sa, a = _call(ctx, "cm_pk_shells_noalloc", o, dk)
                sb, b = _call(ctx, "cm_pk_shells", o, dk)
                assert sa == 0 and sb == 0, ctx.lib.cm_last_error(ctx.h)
                assert np.array_equal(a[3], b[3]) and a[3].sum() > 0
                scale = np.abs(b[0]).max()
                for x, y in zip(a[:3], b[:3]):
                    assert np.max(np.abs(x - y)) <= 1e-12 * scale


def test_noalloc_refuses_sizes_beyond_its_array():
    import cosmicatlas_amd as cm
    with cm.Context(0) as ctx:
        dk = _spectrum(16, 1, torch.complex64)
        st, _ = _call(ctx, "cm_pk_shells_noalloc", _opts(2048, 100.0, 8, 0.0, 0.1, 1, 0.0, 1.0), dk)
        assert st == 2 and b"beyond its fixed scratch" in ctx.lib.cm_last_error(ctx.h)
        st, _ = _call(ctx, "cm_pk_shells_noalloc", _opts(16, 100.0, 2000, 0.0, 0.01, 1, 0.0, 1.0), dk)
        assert st == 2
        # a valid call afterwards still works
        st, r = _call(ctx, "cm_pk_shells_noalloc", _opts(16, 100.0, 8, 0.0, 2 * np.pi / 100.0, 1, 0.0, 1.0), dk)
        assert st == 0 and r[3].sum() > 0
