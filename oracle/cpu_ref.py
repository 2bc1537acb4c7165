"""ctypes front end of libcpuref.so, the OpenMP CPU baseline (test/bench
infrastructure only; see cpu_ref.c)."""
import ctypes as C
import math
import os

import numpy as np

from . import HERE, build

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(HERE, "libcpuref.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.cr_time_pipeline.argtypes = [C.c_int, C.c_float, C.c_void_p, C.c_size_t, C.c_int, C.c_float,
                                       C.c_float, C.c_float, C.c_int, C.c_void_p]
        _LIB = L
    return _LIB


def time_pipeline(pos, N, L, scheme=2, ndel=1, R=2.0, lambda_th=0.0, threads=0, repeats=1, warmup=0,
                  progress=None):
    """Seconds per stage (mas, fft, pk, tweb, tweb_fft) of the TSC pipeline at N^3:
    the run with the median total of `repeats` timed runs after `warmup` untimed ones,
    plus "runs", the totals of the timed runs.  progress(i, seconds) is called
    after each run."""
    assert scheme == 2, "the CPU baseline restates the TSC path"
    pos = np.ascontiguousarray(pos, np.float32)
    kf = 2 * math.pi / L
    ns = N // ndel // 2
    runs = []
    for r in range(warmup + repeats):
        out = np.zeros(5)
        lib().cr_time_pipeline(N, L, pos.ctypes.data_as(C.c_void_p), len(pos), ns, ndel * kf, R, lambda_th,
                               threads, out.ctypes.data_as(C.c_void_p))
        if progress:
            progress(r, float(out[:4].sum()))
        if r >= warmup:
            runs.append(out)
    totals = [float(o[:4].sum()) for o in runs]
    runs.sort(key=lambda o: o[:4].sum())
    med = runs[len(runs) // 2]
    return {"mas": med[0], "fft": med[1], "pk": med[2], "tweb": med[3], "tweb_fft": med[4], "runs": totals}
