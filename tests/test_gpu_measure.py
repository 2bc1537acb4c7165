"""Whole-mock measurement (particles -> P(k) + T-web) on the GPU against the
oracle pipeline, plain and fused paths."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


def _oracle_pk(orc, pos, N, L, scheme, b, mas_correction=True):
    mesh = orc.mas(pos, N, L, scheme)
    nobj = mesh.sum()
    V = L ** 3
    dk = orc.fft_r2c(mesh - nobj / N ** 3)
    o = orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, V / nobj, nobj ** 2 / V,
                    mas_correction=mas_correction)
    p0, p2, p4, nm = orc.pk_shells(dk, o)
    return p0, p2, p4, nm, mesh, nobj, V / nobj


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_measure_zeldovich_pk(ctx, orc, fused, scheme):
    import cosmicatlas_amd as cm
    from oracle import zeldovich as synth
    NL, N, L = 32, 64, 250.0
    pos = synth.zeldovich(NL, L, A=2e3, seed=11)
    b = cm.binning(N, L, ndel=1, DeltaKmin=0.0)
    out = ctx.measure(torch.from_numpy(pos).cuda(), N, L, scheme, b["nshells"], b["kmin"], b["dk"], fused=fused)
    p0, p2, p4, nm, _, nobj, SN = _oracle_pk(orc, pos, N, L, scheme, b)
    assert np.array_equal(out["nmodes"], nm)
    assert out["n_obj"] == pytest.approx(nobj, rel=1e-6)
    # error relative to the un-subtracted power (P + SN*Sq/W^2 >= P): 1e-4
    denom = np.abs(p0) + SN
    assert np.max(np.abs(out["pk0"] - p0) / denom) < 1e-4
    assert np.max(np.abs(out["pk2"] - p2) / denom) < 1e-4


def test_measure_tweb_fused_equals_plain_and_oracle(ctx, orc):
    import cosmicatlas_amd as cm
    from oracle import zeldovich as synth
    NL, N, L, R = 32, 32, 200.0, 12.0
    pos = synth.zeldovich(NL, L, A=2e3, seed=3)
    b = cm.binning(N, L)
    posd = torch.from_numpy(pos).cuda()
    kw = dict(do_tweb=True, R_smooth=R, lambda_th=0.0, margin_tol=1e-4)
    a = ctx.measure(posd, N, L, 2, b["nshells"], b["kmin"], b["dk"], fused=False, **kw)
    la = a["labels"].cpu().numpy()
    c = ctx.measure(posd, N, L, 2, b["nshells"], b["kmin"], b["dk"], fused=True, **kw)
    lc = c["labels"].cpu().numpy()
    mesh = orc.mas(pos, N, L, orc.TSC)
    mean = mesh.sum() / N ** 3
    lab_ref, lam_ref, _, _ = orc.tweb(mesh / mean - 1.0, L, R, 0.0)
    ok = np.abs(lam_ref).min(axis=0) > 1e-4 * np.abs(lam_ref).max()
    assert np.array_equal(la[ok], lab_ref[ok])
    assert np.array_equal(lc[ok], lab_ref[ok])
    assert sum(c["tweb_counts"]) == N ** 3


def test_c1_plain_pipeline_f64(ctx, orc):
    """Config 1: 200,000 float64 uniform positions, L=500, CIC on a 128^3 float64
    mesh, r2c, 64 shells with DeltaKmin=1; NGP counts bit-exact."""
    import cosmicatlas_amd as cm
    N, L, n = 128, 500.0, 200000
    pos = np.random.default_rng(1).uniform(0, L, (n, 3))
    posd = torch.from_numpy(pos).cuda()
    ngp = ctx.mesh_assign(posd, N, L, "NGP", mesh_dtype=torch.float64).cpu().numpy()
    assert np.array_equal(ngp, orc.mas(pos, N, L, orc.NGP))
    mesh = ctx.mesh_assign(posd, N, L, "CIC", mesh_dtype=torch.float64)
    nobj = ctx.mesh_sum(mesh)
    dk = ctx.fft_r2c(mesh, subtract=nobj / N ** 3)
    b = cm.binning(N, L, ndel=1, DeltaKmin=1.0)
    V = L ** 3
    p0, p2, p4, nm = ctx.pk_shells(dk, N, L, b["nshells"], b["kmin"], b["dk"], 1, V / nobj, nobj ** 2 / V)
    ref_mesh = orc.mas(pos, N, L, orc.CIC)
    rdk = orc.fft_r2c(ref_mesh - ref_mesh.sum() / N ** 3)
    o = orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], 1, V / ref_mesh.sum(), ref_mesh.sum() ** 2 / V)
    r0, _, _, rn = orc.pk_shells(rdk, o)
    assert np.array_equal(nm, rn)
    assert np.max(np.abs(p0 - r0) / (np.abs(r0) + V / nobj)) < 1e-10


def test_measure_fused_record_overflow_retry(ctx, orc):
    """Every particle on a tile corner (8 records each with TSC on 16^3 tiles)
    exceeds the fused path's record estimate: the measurement is repeated with
    the exact record count and must equal the oracle and the plain path."""
    import cosmicatlas_amd as cm
    N, L = 64, 256.0
    d = L / N
    rng = np.random.default_rng(5)
    # cells with every coordinate == 0 mod 16 (tile corners), jittered inside the cell
    corners = np.array([(i, j, k) for i in range(0, N, 16) for j in range(0, N, 16) for k in range(0, N, 16)])
    cells = corners[rng.integers(0, len(corners), 60000)]
    pos = ((cells + rng.uniform(0.05, 0.95, cells.shape)) * d).astype(np.float32)
    b = cm.binning(N, L, ndel=1, DeltaKmin=0.0)
    posd = torch.from_numpy(pos).cuda()
    fz = ctx.measure(posd, N, L, 2, b["nshells"], b["kmin"], b["dk"], fused=True)
    pl = ctx.measure(posd, N, L, 2, b["nshells"], b["kmin"], b["dk"], fused=False)
    p0, _, _, nm, _, nobj, SN = _oracle_pk(orc, pos, N, L, 2, b)
    assert np.array_equal(fz["nmodes"], nm)
    assert fz["n_obj"] == pytest.approx(nobj, rel=1e-6)
    denom = np.abs(p0) + SN
    assert np.max(np.abs(fz["pk0"] - p0) / denom) < 1e-4
    assert np.max(np.abs(fz["pk0"] - pl["pk0"]) / denom) < 1e-4


def _alias_tables(N, scheme):
    """Per-axis aliasing sums of the shot noise: C1 (the reference's closed
    forms, Miscelanious.cpp:2532-2554) and the alternating sum
    A = sum_n (-1)^n sinc^{2p}(x + n pi) by brute force (|n| <= 2000)."""
    i = np.arange(N // 2 + 1)
    x = np.pi * i / N
    s2 = np.sin(x) ** 2
    C = [np.ones_like(x), 1 - 2 / 3 * s2, 1 - s2 + 2 / 15 * s2 ** 2,
         1 - 4 / 3 * s2 + 2 / 5 * s2 ** 2 - 4 / 315 * s2 ** 3][scheme]
    n = np.arange(-2000, 2001)
    y = x[:, None] + np.pi * n[None, :]
    sinc = np.where(np.abs(y) > 0, np.sin(y) / np.where(np.abs(y) > 0, y, 1.0), 1.0)
    A = np.sum(np.where(n % 2 == 0, 1.0, -1.0)[None, :] * sinc ** (2 * (scheme + 1)), axis=1)
    return C, A


def _interlaced_ref(orc, pos, N, L, scheme, b):
    """Oracle spectrum of interlaced meshes: the average of the plain spectrum
    and the half-cell-shifted one times exp(i k.(d/2)(1,1,1)), minus the aliased
    shot noise (prod C1 + prod A)/2 per mode."""
    mA = orc.mas(pos, N, L, scheme)
    mB = orc.mas(pos, N, L, scheme, interlace_shift=0.5)
    nobj = mA.sum()
    V = L ** 3
    dA = orc.fft_r2c(mA - nobj / N ** 3)
    dB = orc.fft_r2c(mB - nobj / N ** 3)
    j = np.arange(N)
    sj = np.where(j <= N // 2, j, j - N)
    kz = np.arange(N // 2 + 1)
    ph = np.exp(1j * np.pi * (sj[:, None, None] + sj[None, :, None] + kz[None, None, :]) / N)
    d = 0.5 * (dA + ph * dB)
    normal = nobj ** 2 / V
    lk = dict(log_binning=True, dlog=b["dlog"]) if b.get("log_binning") else {}
    p0, p2, p4, nm = orc.pk_shells(d, orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, 0.0, normal, **lk))
    C, A = _alias_tables(N, scheme)
    f = np.where(j <= N // 2, j, N - j)
    Cx, Ax = C[f], A[f]
    X = 0.5 * (Cx[:, None, None] * Cx[None, :, None] * C[None, None, :]
               + Ax[:, None, None] * Ax[None, :, None] * A[None, None, :])
    sn_shell = orc.pk_shells(np.sqrt(X), orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, 0.0, 1.0,
                                                    **lk))[0]
    SN = V / nobj
    return p0 - SN * sn_shell, p2, nm, SN, d, nobj, C, A


@pytest.mark.parametrize("scheme,do_tweb", [(1, False), (2, True), (0, False), (3, False)])
def test_measure_interlaced_matches_oracle(ctx, orc, scheme, do_tweb):
    """Interlaced meshes (no reference counterpart), both deposited from one
    binning: the records carry one more ghost cell towards the tile above, and
    the second mesh moves each record's fraction by half a cell."""
    import cosmicatlas_amd as cm
    from oracle import zeldovich as zref
    N, L = 32, 200.0
    pos = zref.zeldovich(24, L, A=2e3, seed=21)
    b = cm.binning(N, L)
    out = ctx.measure(torch.from_numpy(pos).cuda(), N, L, scheme, b["nshells"], b["kmin"], b["dk"],
                      interlace=True, do_tweb=do_tweb, R_smooth=12.0, margin_tol=1e-4)
    ref0, p2, nm, SN, d, nobj, C, A = _interlaced_ref(orc, pos, N, L, scheme, b)
    assert np.array_equal(out["nmodes"], nm)
    denom = np.abs(ref0) + SN
    assert np.max(np.abs(out["pk0"] - ref0) / denom) < 1e-4
    assert np.max(np.abs(out["pk2"] - p2) / denom) < 1e-4
    # the interlacing removes most of the aliased power: the alternating sums
    # are well below the plain ones near the Nyquist frequency
    assert A[-1] < 0.5 * C[-1]
    if do_tweb:
        mean = nobj / N ** 3
        delta = np.fft.irfftn(d, s=(N, N, N)) / mean
        lab_ref, lam_ref, _, _ = orc.tweb(delta, L, 12.0, 0.0)
        ok = np.abs(lam_ref).min(axis=0) > 1e-4 * np.abs(lam_ref).max()
        assert np.array_equal(out["labels"].cpu().numpy()[ok], lab_ref[ok])


@pytest.mark.parametrize("scheme,hist", [(2, 32), (1, 48), (3, 40), (0, 16)])
def test_measure_interlaced_grouped_binning(orc, scheme, hist):
    """The shared binning of the interlaced meshes through x-slab groups (forced
    on a small mesh by a small histogram budget): copies across group edges
    include the extra ghost cell of the shifted mesh, with the periodic wrap."""
    import os
    import cosmicatlas_amd as cm
    old = os.environ.get("CM_TUNE")
    os.environ["CM_TUNE"] = f"mas_hist={hist}"
    try:
        c = cm.Context(0)
    finally:
        if old is None:
            del os.environ["CM_TUNE"]
        else:
            os.environ["CM_TUNE"] = old
    try:
        N, L, n = 64, 300.0, 120000
        rng = np.random.default_rng(scheme + hist)
        pos = rng.uniform(0.0, L, (n, 3)).astype(np.float32)
        b = cm.binning(N, L)
        out = c.measure(torch.from_numpy(pos).cuda(), N, L, scheme, b["nshells"], b["kmin"], b["dk"], interlace=True)
    finally:
        c.close()
    ref0, p2, nm, SN, _, _, _, _ = _interlaced_ref(orc, pos, N, L, scheme, b)
    assert np.array_equal(out["nmodes"], nm)
    denom = np.abs(ref0) + SN
    assert np.max(np.abs(out["pk0"] - ref0) / denom) < 1e-4
    assert np.max(np.abs(out["pk2"] - p2) / denom) < 1e-4


def test_measure_interlaced_ngp_dense_tiles(ctx, orc):
    """NGP with the shifted mesh copies a particle to the tile above when it sits
    in a tile's last cell plane, so a tile holds more records than particles.  A
    dense small mesh (N = 32: 8 tiles of 16^3, ~30k records each) splits every
    tile over several deposit workgroups; the records must be bounded by the
    ghost copies, and the estimate-sized record buffer must not overflow (so
    the third call replays a captured graph instead of the exact eager path)."""
    import cosmicatlas_amd as cm
    N, L, n = 32, 100.0, 200000
    rng = np.random.default_rng(77)
    pos = rng.uniform(0.0, L, (n, 3)).astype(np.float32)
    b = cm.binning(N, L)
    with cm.Context(0) as c:
        pd = torch.from_numpy(pos).cuda()
        outs = [c.measure(pd, N, L, 0, b["nshells"], b["kmin"], b["dk"], interlace=True) for _ in range(3)]
        captured, launched = c.graph_counts()
    # call 1 eager, call 2 captured and launched, call 3 launched
    assert (captured, launched) == (1, 2), (captured, launched)
    ref0, p2, nm, SN, _, nobj, _, _ = _interlaced_ref(orc, pos, N, L, 0, b)
    assert nobj == n
    for out in outs:
        assert out["n_obj"] == n
        assert np.array_equal(out["nmodes"], nm)
        denom = np.abs(ref0) + SN
        assert np.max(np.abs(out["pk0"] - ref0) / denom) < 1e-4
        assert np.max(np.abs(out["pk2"] - p2) / denom) < 1e-4


@pytest.mark.parametrize("scheme,interlace", [(2, False), (1, True), (0, False)])
def test_measure_log_binning_matches_oracle(ctx, orc, scheme, interlace):
    """The fused x pass walks the log shells with the same integer thresholds
    as the shell kernel: counts equal the oracle's log10 rule, P(k) within
    1e-4 (type_of_binning = log, Params.cpp:4104-4106)."""
    import cosmicatlas_amd as cm
    N, L, n = 64, 300.0, 150000
    rng = np.random.default_rng(31 + scheme)
    pos = rng.uniform(0.0, L, (n, 3)).astype(np.float32)
    b = cm.binning_log(N, L, DeltaKmin=1.0, N_log_bins=8)
    out = ctx.measure(torch.from_numpy(pos).cuda(), N, L, scheme, b["nshells"], b["kmin"], 0.0, interlace=interlace,
                      log_binning=True, dlog=b["dlog"])
    if interlace:
        ref0, _, nm, SN, _, _, _, _ = _interlaced_ref(orc, pos, N, L, scheme, b)
    else:
        mesh = orc.mas(pos, N, L, scheme)
        nobj = mesh.sum()
        V = L ** 3
        SN = V / nobj
        o = orc.pk_opts(N, L, b["nshells"], b["kmin"], 0.0, scheme, SN, nobj ** 2 / V, log_binning=True,
                        dlog=b["dlog"])
        ref0, _, _, nm = orc.pk_shells(orc.fft_r2c(mesh - nobj / N ** 3), o)
    assert np.array_equal(out["nmodes"], nm)
    assert nm.sum() > 0
    assert np.max(np.abs(out["pk0"] - ref0) / (np.abs(ref0) + SN)) < 1e-4
