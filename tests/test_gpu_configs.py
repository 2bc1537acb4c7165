"""The large, batched and serving configurations at their stated sizes.

C3: 5.4e8 Zel'dovich particles (every other Lagrangian x plane of a 1024^3
    field, P = 1e4 k^-1.5, L = 2000 Mpc/h) plus 20% uniform randoms,
    interlaced TSC onto 1024^3, 256 shells: per-shell float32-vs-FP64 error of
    the raw power below 1e-4 and mode counts equal to the exact enumeration.
C4: 64 uneven mocks (N in {128, 192, 256, 384, 512}, NL^3 particles
    log-uniform in [1e5, 5e7] with NL rounded to the generator's grids, L in
    [250, 2000], NGP/CIC/TSC, 64 shells, T-web) through cm_measure_batch:
    each equals its alone run, P(k) within 1e-4 of the FP64 pipeline, mode
    counts exact.
C5: a Poisson stream of requests from 128^3 to 768^3 with 1e5 to 1.3e8
    particles (one above 1e8), optional T-web, admitted by HBM footprint:
    each result equals the same request alone."""
import math
import threading
import time

import numpy as np
import pytest
import torch

from gpu_util import exact_mode_counts

pytestmark = pytest.mark.gpu

GEN_NL = (48, 64, 96, 128, 192, 256, 384, 512)


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


def _fp64_raw(ctx, pos, N, L, scheme, b, interlace=False):
    """Raw (shot noise not subtracted) FP64 shells of the same particles:
    FP64 deposit(s), FFT, interlacing combination and shells on the GPU."""
    import cosmicatlas_amd as cm
    m = ctx.mesh_assign(pos, N, L, scheme, algo="plain", mesh_dtype=torch.float64)
    nobj = ctx.mesh_sum(m)
    d = ctx.fft_r2c(m, subtract=nobj / N ** 3)
    del m
    if interlace:
        m = ctx.mesh_assign(pos, N, L, scheme, algo="plain", mesh_dtype=torch.float64, interlace_shift=0.5)
        d2 = ctx.fft_r2c(m, subtract=nobj / N ** 3)
        del m
        j = torch.arange(N, device="cuda", dtype=torch.float64)
        sj = torch.where(j <= N // 2, j, j - N)
        kz = torch.arange(N // 2 + 1, device="cuda", dtype=torch.float64)
        for ax, v in ((0, sj), (1, sj), (2, kz)):
            shape = [1, 1, 1]
            shape[ax] = -1
            d2 *= torch.exp(1j * math.pi * v / N).view(shape)
        d += d2
        d *= 0.5
        del d2
    V = L ** 3
    p0, _, _, nm = ctx.pk_shells(d, N, L, b["nshells"], b["kmin"], b["dk"], scheme, 0.0, nobj ** 2 / V)
    del d
    torch.cuda.empty_cache()
    return p0, nm, nobj


def test_c3_large_interlaced_1024(ctx):
    import cosmicatlas_amd as cm
    N, L, NLg = 1024, 2000.0, 1024
    zel = ctx.zeldovich(NLg, L, A=1e4, ns=-1.5, seed=3)
    # every other Lagrangian x plane: 5.4e8 particles of the same field
    pos = zel.view(NLg, NLg * NLg, 3)[::2].reshape(-1, 3).contiguous()
    del zel
    torch.cuda.empty_cache()
    nz = pos.shape[0]
    assert nz >= 5e8
    ran = ctx.uniform(nz // 5, L, seed=4)
    pos = torch.cat([pos, ran])
    del ran
    b = cm.binning(N, L, ndel=2)
    assert b["nshells"] == 256
    out = ctx.measure(pos, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], interlace=True, sn_correction=False)
    assert out["n_obj"] == pytest.approx(pos.shape[0], rel=1e-6)
    p64, nm64, nobj = _fp64_raw(ctx, pos, N, L, "TSC", b, interlace=True)
    del pos
    torch.cuda.empty_cache()
    assert np.array_equal(out["nmodes"], nm64)
    assert np.array_equal(nm64, exact_mode_counts(N, 256, 0, 2))
    ok = nm64 > 0
    err = np.abs(out["pk0"][ok] - p64[ok]) / np.abs(p64[ok])
    assert err.max() < 1e-4, err.max()


def _c4_specs(n=64, seed=7):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        Nm = int(rng.choice([128, 192, 256, 384, 512]))
        nl = math.exp(rng.uniform(math.log(1e5), math.log(5e7))) ** (1.0 / 3.0)
        NLm = min(GEN_NL[:-1], key=lambda g: abs(math.log(g / nl)))
        Lm = float(rng.uniform(250.0, 2000.0))
        sch = ["NGP", "CIC", "TSC"][int(rng.integers(0, 3))]
        ndel = 1 if Nm == 192 else Nm // 128
        specs.append((Nm, NLm, Lm, sch, ndel, True, 5000 + i))
    return specs


def _request(srv, spec, labels=None):
    import cosmicatlas_amd as cm
    Nm, NLm, Lm, sch, ndel, tw, sd = spec
    bm = cm.binning(Nm, Lm, ndel=ndel)
    kw = dict(do_tweb=True, R_smooth=2.0 * Lm / Nm, margin_tol=cm.TOL_ROUNDING) if tw else {}
    return srv.request(Nm, Lm, sch, 64, bm["kmin"], bm["dk"], gen=dict(NL=NLm, seed=sd), labels=labels, **kw)


def _alone(ctx, spec):
    import cosmicatlas_amd as cm
    Nm, NLm, Lm, sch, ndel, tw, sd = spec
    bm = cm.binning(Nm, Lm, ndel=ndel)
    kw = dict(do_tweb=True, R_smooth=2.0 * Lm / Nm, margin_tol=cm.TOL_ROUNDING) if tw else {}
    pos = ctx.zeldovich(NLm, Lm, A=1e4, ns=-1.5, seed=sd)
    return ctx.measure(pos, Nm, Lm, sch, 64, bm["kmin"], bm["dk"], **kw), pos, bm


def _same(a, b, spec):
    """In the mix as alone: equal mode counts and object count, P(k) to the
    rounding of FP64 shell sums and of the mesh cells that crowded tiles (more
    than 4096 records, split over workgroups) add with float atomics in
    schedule order (measured up to 1e-7 of |P| + SN), labels equal outside
    the counted margin band."""
    Nm, Lm = spec[0], spec[2]
    assert np.array_equal(a["nmodes"], b["nmodes"])
    assert a["n_obj"] == pytest.approx(b["n_obj"], rel=1e-9)
    sn = Lm ** 3 / b["n_obj"]
    assert np.max(np.abs(a["pk0"] - b["pk0"]) / (np.abs(b["pk0"]) + sn)) < 1e-6
    if spec[5]:
        assert int((a["labels"] != b["labels"]).sum()) <= max(a["tweb_marginal"], b["tweb_marginal"])
        assert sum(a["tweb_counts"]) == Nm ** 3


def test_c4_batched_64_uneven_mocks(ctx):
    import cosmicatlas_amd as cm
    specs = _c4_specs()
    counts = {}
    with cm.Server(0, slots=3) as srv:
        res = srv.batch([_request(srv, s) for s in specs])
        adm = srv.stats()
    assert len(res) == 64
    # with the default budget no request outgrows its reservation (a re-run
    # alone would serialise the server)
    assert adm["reruns"] == 0 and adm["peak_held"] <= adm["budget"]
    for spec, r in zip(specs, res):
        a, pos, bm = _alone(ctx, spec)
        _same(r, a, spec)
        Nm, NLm, Lm, sch, ndel = spec[:5]
        key = (Nm, ndel)
        if key not in counts:
            counts[key] = exact_mode_counts(Nm, Nm // ndel // 2, 0, ndel)[:64]
        assert np.array_equal(r["nmodes"], counts[key])
        # float32 against FP64 on the raw power (no shot-noise subtraction)
        p64, nm64, _ = _fp64_raw(ctx, pos, Nm, Lm, sch, dict(nshells=64, kmin=bm["kmin"], dk=bm["dk"]))
        raw32 = ctx.measure(pos, Nm, Lm, sch, 64, bm["kmin"], bm["dk"], sn_correction=False)["pk0"]
        del pos
        assert np.array_equal(nm64, r["nmodes"])
        ok = nm64 > 0
        assert np.max(np.abs(raw32[ok] - p64[ok]) / np.abs(p64[ok])) < 1e-4


def test_c4_packed_64_uneven_mocks(ctx):
    """The same 64 mocks through cm_measure_packed (one deposit launch per
    scheme over the union (mock, tile) work list, one T-web launch for all):
    each equals its alone run."""
    import cosmicatlas_amd as cm
    specs = _c4_specs()
    args = []
    for (Nm, NLm, Lm, sch, ndel, tw, sd) in specs:
        bm = cm.binning(Nm, Lm, ndel=ndel)
        args.append(dict(pos=ctx.zeldovich(NLm, Lm, A=1e4, ns=-1.5, seed=sd), N=Nm, L=Lm, scheme=sch, nshells=64,
                         kmin=bm["kmin"], dk_bin=bm["dk"], do_tweb=True, R_smooth=2.0 * Lm / Nm,
                         margin_tol=cm.TOL_ROUNDING))
    res = ctx.measure_packed(args)
    del args
    torch.cuda.empty_cache()
    for spec, r in zip(specs, res):
        a, pos, _ = _alone(ctx, spec)
        del pos
        _same(r, a, spec)


def test_c5_poisson_stream_mixed(ctx):
    import cosmicatlas_amd as cm
    rng = np.random.default_rng(11)
    specs = []
    for i in range(24):
        Nm = int(rng.choice([128, 192, 256, 384, 512, 768]))
        nl = math.exp(rng.uniform(math.log(1e5), math.log(1e8))) ** (1.0 / 3.0)
        NLm = min(GEN_NL[:-1], key=lambda g: abs(math.log(g / nl)))
        Lm = float(rng.uniform(250.0, 2000.0))
        sch = ["NGP", "CIC", "TSC"][int(rng.integers(0, 3))]
        ndel = 1 if Nm in (128, 192) else Nm // 128
        specs.append((Nm, NLm, Lm, sch, ndel, bool(rng.integers(0, 2)), 7000 + i))
    specs[5] = (768, 512, 1500.0, "TSC", 6, True, 7777)  # 1.3e8 particles
    with cm.Server(0, slots=3) as srv:
        tickets = [None] * len(specs)

        def producer():
            for k, sp in enumerate(specs):
                tickets[k] = srv.submit(_request(srv, sp))
                time.sleep(float(rng.exponential(0.004)))

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        res = [t.result(timeout=120.0) for t in tickets]
        adm = srv.stats()
    assert adm["reruns"] == 0 and adm["peak_held"] <= adm["budget"]
    assert max(sp[1] ** 3 for sp in specs) >= 1e8
    for spec, r in zip(specs, res):
        a, pos, _ = _alone(ctx, spec)
        del pos
        _same(r, a, spec)
        assert r["latency_ms"] > 0
