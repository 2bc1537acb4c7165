"""FKP estimator with a random catalogue (cm_measure_fkp) against the oracle
pipeline, a closed form and the box estimator."""
import numpy as np
import pytest
import torch

from oracle import zeldovich as zref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


def _sq_table(N, scheme):
    i = np.arange(N // 2 + 1)
    s2 = np.sin(np.pi * i / N) ** 2
    return [np.ones_like(s2), 1 - 2 / 3 * s2, 1 - s2 + 2 / 15 * s2 ** 2,
            1 - 4 / 3 * s2 + 2 / 5 * s2 ** 2 - 4 / 315 * s2 ** 3][scheme]


def test_randoms_equal_to_galaxies_closed_form(ctx, orc):
    """Randoms at the galaxies' positions: alpha = 1 and delta = 0, so P0 is
    minus the shot noise 2/nbar times the shell mean of prod Sq / W^2."""
    import cosmicatlas_amd as cm
    N, L, n = 32, 100.0, 30000
    pos = zref.uniform(n, L, 3)
    posd = torch.from_numpy(pos).cuda()
    b = cm.binning(N, L)
    out = ctx.measure_fkp(posd, posd, N, L, "CIC", b["nshells"], b["kmin"], b["dk"])
    nbar = n / L ** 3
    assert out["shot_noise"] == pytest.approx(2.0 / nbar, rel=1e-12)
    assert out["normal_power"] == pytest.approx(n * nbar, rel=1e-12)
    C = _sq_table(N, 1)
    f = np.where(np.arange(N) <= N // 2, np.arange(N), N - np.arange(N))
    X = C[f][:, None, None] * C[f][None, :, None] * C[None, None, :]
    m = orc.pk_shells(np.sqrt(X), orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], 1, 0.0, 1.0))[0]
    ref = -2.0 / nbar * m
    assert np.allclose(out["pk0"], ref, rtol=1e-5, atol=0)
    assert np.max(np.abs(out["pk2"])) < 1e-6 * 2.0 / nbar


def test_fkp_matches_oracle(ctx, orc):
    import cosmicatlas_amd as cm
    N, L, scheme = 32, 200.0, 2
    gal = zref.zeldovich(24, L, A=2e3, seed=5)
    ran = zref.uniform(3 * len(gal), L, 6)
    rng = np.random.default_rng(1)
    wg = rng.uniform(0.5, 1.5, len(gal)).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, len(ran)).astype(np.float32)
    nbar, P_est = 1.5e-3, 2e3
    b = cm.binning(N, L)
    out = ctx.measure_fkp(torch.from_numpy(gal).cuda(), torch.from_numpy(ran).cuda(), N, L, scheme, b["nshells"],
                          b["kmin"], b["dk"], weights=torch.from_numpy(wg).cuda(),
                          rweights=torch.from_numpy(wr).cuda(), nbar=nbar, P_est=P_est, fkp_weight=True, window=True)
    wf = 1.0 / (1.0 + nbar * P_est)
    Wg = wf * wg.astype(np.float64).sum()
    Wr = wf * wr.astype(np.float64).sum()
    Sr = wf ** 2 * (wr.astype(np.float64) ** 2).sum()
    alpha = Wg / Wr
    normal = alpha * nbar * Sr
    SN = (alpha + 1) * Sr / (nbar * Sr)
    mesh = (orc.mas(gal, N, L, scheme, weights=(wf * wg).astype(np.float32))
            - alpha * orc.mas(ran, N, L, scheme, weights=(wf * wr).astype(np.float32)))
    dk = orc.fft_r2c(mesh)
    p0, p2, p4, nm = orc.pk_shells(dk, orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, SN, normal))
    assert np.array_equal(out["nmodes"], nm)
    assert out["shot_noise"] == pytest.approx(SN, rel=1e-9)
    denom = np.abs(p0) + SN
    assert np.max(np.abs(out["pk0"] - p0) / denom) < 1e-4
    assert np.max(np.abs(out["pk2"] - p2) / denom) < 1e-4
    assert np.max(np.abs(out["pk4"] - p4) / denom) < 1e-4
    # window: the reference's mode set of the randoms' raw spectrum (zero mode
    # in shell 0 at kmin = 0, first octant left out), over normal_power/alpha^2
    mr = orc.mas(ran, N, L, scheme, weights=(wf * wr).astype(np.float32))
    pw, _ = orc.window_shells(orc.fft_r2c(mr), orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, 0.0,
                                                             normal / alpha ** 2))
    assert pw[0] > 100 * pw[1:].max()  # |W(0)|^2 dominates the first shell
    assert np.allclose(out["pk_window"], pw, rtol=1e-4, atol=0)


@pytest.mark.parametrize("dmin,logb", [(1.0, False), (1.0, True)])
def test_fkp_window_binnings(ctx, orc, dmin, logb):
    """The window's zero mode goes to shell 0 under log binning (kv = 0 takes
    index 0, FftwFunctions.cpp:1110-1113) and to no shell when kmin > 0 under
    linear binning."""
    import cosmicatlas_amd as cm
    N, L, scheme = 32, 200.0, 1
    gal = zref.zeldovich(24, L, A=2e3, seed=15)
    ran = zref.uniform(4 * len(gal), L, 16)
    if logb:
        b = cm.binning_log(N, L, DeltaKmin=dmin, N_log_bins=6)
        kw = dict(log_binning=True, dlog=b["dlog"])
    else:
        b = cm.binning(N, L, DeltaKmin=dmin)
        kw = {}
    out = ctx.measure_fkp(torch.from_numpy(gal).cuda(), torch.from_numpy(ran).cuda(), N, L, scheme, b["nshells"],
                          b["kmin"], b["dk"], window=True, **kw)
    alpha = len(gal) / len(ran)
    normal = alpha * (len(gal) / L ** 3) * len(ran)
    mr = orc.mas(ran, N, L, scheme)
    pw, nm = orc.window_shells(orc.fft_r2c(mr), orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, 0.0,
                                                              normal / alpha ** 2, **kw))
    assert np.allclose(out["pk_window"], pw, rtol=1e-4, atol=0)
    if logb:
        assert pw[0] > 0 and nm[0] >= 1


def test_fkp_tends_to_box_estimator(ctx):
    """With many more randoms than galaxies the FKP estimate of a periodic box
    approaches the box estimator on shells where the signal dominates."""
    import cosmicatlas_amd as cm
    N, L = 64, 300.0
    gal = torch.from_numpy(zref.zeldovich(32, L, A=3e3, seed=8)).cuda()
    ran = ctx.uniform(40 * gal.shape[0], L, seed=9)
    b = cm.binning(N, L)
    fk = ctx.measure_fkp(gal, ran, N, L, "TSC", b["nshells"], b["kmin"], b["dk"])
    box = ctx.measure(gal, N, L, "TSC", b["nshells"], b["kmin"], b["dk"])
    SN = L ** 3 / gal.shape[0]
    sig = box["pk0"] > 5 * SN
    assert sig.sum() >= 3
    rel = np.abs(fk["pk0"][sig] - box["pk0"][sig]) / box["pk0"][sig]
    assert rel.max() < 0.15, rel


def _fkp_oracle(orc, gal, ran, wg, wr, nbg, nbr, P_est, N, L, scheme, b):
    """The estimator assembled from oracle deposits, FFT and shells with the
    per-object sums of Catalog::get_interpolated_density_field
    (Source/Catalog.cpp:14142-14171) and the random branch of
    get_parameters_estimator (Source/FftwFunctions.cpp:180-186)."""
    fg = wg.astype(np.float64) / (1.0 + P_est * nbg.astype(np.float64))
    fr = wr.astype(np.float64) / (1.0 + P_est * nbr.astype(np.float64))
    alpha = fg.sum() / fr.sum()
    normal_r = (nbr.astype(np.float64) * fr ** 2).sum()
    SN = (alpha + 1) * (fr ** 2).sum() / normal_r
    mesh = (orc.mas(gal, N, L, scheme, weights=fg.astype(np.float32))
            - alpha * orc.mas(ran, N, L, scheme, weights=fr.astype(np.float32)))
    p = orc.pk_shells(orc.fft_r2c(mesh), orc.pk_opts(N, L, b["nshells"], b["kmin"], b["dk"], scheme, SN,
                                                     alpha * normal_r))
    return p, SN, alpha * normal_r


@pytest.mark.parametrize("scheme", [0, 1, 3])
def test_per_object_mean_density(ctx, orc, scheme):
    """n(r) per object (a gradient along x, as a selection function gives):
    weights 1/(1 + P_est n_i) and normal = sum n_i w_i^2 against the oracle,
    for NGP, CIC and PCS deposits."""
    import cosmicatlas_amd as cm
    N, L, P_est = 32, 200.0, 1e4
    gal = zref.zeldovich(24, L, A=2e3, seed=11)
    ran = zref.uniform(3 * len(gal), L, 12)
    rng = np.random.default_rng(13)
    wg = rng.uniform(0.5, 1.5, len(gal)).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, len(ran)).astype(np.float32)
    nbg = (1e-4 * (1.0 + 4.0 * gal[:, 0] / L)).astype(np.float32)
    nbr = (1e-4 * (1.0 + 4.0 * ran[:, 0] / L)).astype(np.float32)
    b = cm.binning(N, L)
    cu = lambda a: torch.from_numpy(a).cuda()
    out = ctx.measure_fkp(cu(gal), cu(ran), N, L, scheme, b["nshells"], b["kmin"], b["dk"], weights=cu(wg),
                          rweights=cu(wr), P_est=P_est, fkp_weight=True, nbar_g=cu(nbg), nbar_r=cu(nbr))
    (p0, p2, p4, nm), SN, normal = _fkp_oracle(orc, gal, ran, wg, wr, nbg, nbr, P_est, N, L, scheme, b)
    assert np.array_equal(out["nmodes"], nm)
    assert out["shot_noise"] == pytest.approx(SN, rel=1e-9)
    assert out["normal_power"] == pytest.approx(normal, rel=1e-9)
    denom = np.abs(p0) + SN
    for got, ref in ((out["pk0"], p0), (out["pk2"], p2), (out["pk4"], p4)):
        assert np.max(np.abs(got - ref) / denom) < 1e-4
    # the weights vary by a factor (1 + 5e-4 P_est) / (1 + 1e-4 P_est) = 3 across the box:
    # the mean density taken constant changes the shot noise S_r / normal_r by ~30%
    const = ctx.measure_fkp(cu(gal), cu(ran), N, L, scheme, b["nshells"], b["kmin"], b["dk"], weights=cu(wg),
                            rweights=cu(wr), nbar=float(nbr.mean()), P_est=P_est, fkp_weight=True)
    assert abs(out["shot_noise"] / const["shot_noise"] - 1) > 0.15


def test_constant_per_object_density_equals_constant(ctx):
    """Arrays that hold the constant nbar reproduce the constant-nbar result, with
    FKP weights on and off (off: nbar enters the normalisation only)."""
    import cosmicatlas_amd as cm
    N, L, nbar = 32, 150.0, 2.5e-3
    gal = torch.from_numpy(zref.zeldovich(20, L, A=2e3, seed=14)).cuda()
    ran = ctx.uniform(4 * gal.shape[0], L, seed=15)
    b = cm.binning(N, L)
    for fkp in (True, False):
        kw = dict(nbar=nbar, P_est=3e3, fkp_weight=fkp)
        a = ctx.measure_fkp(gal, ran, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], **kw)
        c = ctx.measure_fkp(gal, ran, N, L, "TSC", b["nshells"], b["kmin"], b["dk"],
                            nbar_g=torch.full((gal.shape[0],), nbar, device="cuda"),
                            nbar_r=torch.full((ran.shape[0],), nbar, device="cuda"), **kw)
        assert c["normal_power"] == pytest.approx(a["normal_power"], rel=1e-6)
        assert c["shot_noise"] == pytest.approx(a["shot_noise"], rel=1e-6)
        assert np.array_equal(a["nmodes"], c["nmodes"])
        assert np.max(np.abs(a["pk0"] - c["pk0"]) / (np.abs(a["pk0"]) + a["shot_noise"])) < 1e-5


def test_fkp_errors(ctx):
    """Empty randoms, a zero mean density for every random, and per-object
    arrays of the wrong length are refused before any deposit."""
    import cosmicatlas_amd as cm
    N, L = 32, 100.0
    gal = ctx.uniform(1000, L, seed=16)
    ran = ctx.uniform(3000, L, seed=17)
    b = cm.binning(N, L)
    args = (N, L, "CIC", b["nshells"], b["kmin"], b["dk"])
    with pytest.raises(cm.CosmoError, match="empty catalogue"):
        ctx.measure_fkp(gal, ran[:0], *args)
    with pytest.raises(cm.CosmoError, match="nbar w\\^2"):
        ctx.measure_fkp(gal, ran, *args, nbar_r=torch.zeros(3000, device="cuda"))
    with pytest.raises(ValueError):
        ctx.measure_fkp(gal, ran, *args, nbar_g=torch.ones(999, device="cuda"))
    # the context still measures after the refusals
    out = ctx.measure_fkp(gal, ran, *args)
    assert np.isfinite(out["pk0"]).all()
