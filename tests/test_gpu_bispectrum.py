"""Bispectrum by shell inverse transforms (cm_bispectrum) against a brute-force
enumeration of closed triangles and the single-particle closed form."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _shell(q, Dmin, ndel):
    """exact integer shell of |n|^2 = q: (s ndel + Dmin)^2 <= q < ((s+1) ndel + Dmin)^2"""
    s = np.floor((np.sqrt(q) - Dmin) / ndel).astype(int)
    for _ in range(2):
        lo = s * ndel + Dmin
        hi = lo + ndel
        s = np.where(q < lo * lo, s - 1, np.where(q >= hi * hi, s + 1, s))
    return np.where(q < Dmin * Dmin, -1, s)


def _brute(mesh, N, ns, Dmin, ndel, bins):
    """sum of d1 d2 d3 over closed triangles k1 + k2 + k3 = 0 (full grid, no wrap)."""
    d = np.fft.fftn(mesh - mesh.mean())
    f = np.fft.fftfreq(N, 1.0 / N).astype(int)
    kx, ky, kz = np.meshgrid(f, f, f, indexing="ij")
    q = kx ** 2 + ky ** 2 + kz ** 2
    sh = np.where(q > 0, _shell(q, Dmin, ndel), -1)
    sel = (sh >= 0) & (sh < ns)
    K = np.stack([kx[sel], ky[sel], kz[sel]], 1)
    S = sh[sel]
    D = d[sel]
    lut = {tuple(k): i for i, k in enumerate(K)}
    out = {}
    for (i, j, k) in bins:
        tot, cnt = 0.0, 0
        A = np.where(S == i)[0]
        B = np.where(S == j)[0]
        for a in A:
            k3 = -(K[a][None, :] + K[B])
            for b_, kk in zip(B, k3):
                c = lut.get(tuple(kk))
                if c is not None and S[c] == k:
                    tot += (D[a] * D[b_] * D[c]).real
                    cnt += 1
        out[(i, j, k)] = (tot, cnt)
    return out


def test_bispectrum_matches_brute_force():
    import cosmicatlas_amd as cm
    N, L = 16, 100.0
    rng = np.random.default_rng(4)
    mesh = rng.poisson(3.0, (N, N, N)).astype(np.float32)
    ns, ndel = 5, 1
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=float(mesh.mean()))
        n = float(mesh.sum())
        out = ctx.bispectrum(dk, N, L, ns, 0.0, 2 * np.pi / L * ndel, "NGP", n, mas_correction=False,
                             sn_correction=False)
    bins = [tuple(b) for b in out["ijk"]]
    assert len(bins) > 10
    ref = _brute(mesh.astype(np.float64), N, ns, 0, ndel, bins)
    V = L ** 3
    normal_b = n ** 3 / V ** 2
    for b, bk, nt in zip(bins, out["bk"], out["ntri"]):
        tot, cnt = ref[b]
        assert nt == cnt, (b, nt, cnt)
        if cnt:
            expect = tot / cnt / normal_b
            assert abs(bk - expect) <= 1e-4 * (abs(expect) + 1e-3 * np.abs(out["bk"]).max()), (b, bk, expect)


def test_single_particle_closed_form():
    """One particle at a cell centre with NGP: delta_k = exp(-i k.x) for k != 0,
    so every closed triangle gives 1 and B = V^2 / n^3 with n = 1."""
    import cosmicatlas_amd as cm
    N, L = 32, 64.0
    mesh = np.zeros((N, N, N), np.float32)
    mesh[5, 9, 20] = 1.0
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=1.0 / N ** 3)
        out = ctx.bispectrum(dk, N, L, 8, 0.0, 2 * np.pi / L, "NGP", 1.0, mas_correction=False,
                             sn_correction=False)
    V = L ** 3
    ok = out["ntri"] > 0
    assert ok.sum() > 20
    assert np.allclose(out["bk"][ok], V ** 2, rtol=1e-4)


def test_bin_subset_independent_of_shell_count():
    """Bins (i, j, k) with i < 16 are the same triangles whether 16 or 24 shells
    are asked for; 24 shells take 256-cell chunks in LDS, 16 shells 512-cell
    chunks, so the two sums run through different tile and chunk layouts."""
    import cosmicatlas_amd as cm
    N, L = 96, 300.0
    rng = np.random.default_rng(11)
    mesh = rng.poisson(2.0, (N, N, N)).astype(np.float32)
    kf = 2 * np.pi / L
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=float(mesh.mean()))
        n = float(mesh.sum())
        a = ctx.bispectrum(dk, N, L, 16, 2 * kf, kf, "CIC", n)
        b = ctx.bispectrum(dk, N, L, 24, 2 * kf, kf, "CIC", n)
    sub = b["ijk"][:, 0] < 16
    assert len(a["bk"]) == 507 and len(b["bk"]) == 1521 and sub.sum() == 507
    np.testing.assert_array_equal(a["ijk"], b["ijk"][sub])
    assert np.all(np.abs(a["ntri"] - b["ntri"][sub]) <= 1 + 1e-6 * a["ntri"])
    # float32 sums in two orders: a bin's rounding scales with its terms, not with
    # its (often cancelling) total, hence the floor of 1e-7 of the largest bin
    scale = np.abs(a["bk"]).max()
    assert np.all(np.abs(a["bk"] - b["bk"][sub]) <= 1e-5 * (np.abs(a["bk"]) + 1e-2 * scale))
    assert np.all(np.abs(a["sn"] - b["sn"][sub]) <= 1e-5 * np.abs(a["sn"]).max())


def test_shells_beyond_third_of_grid_rejected():
    import cosmicatlas_amd as cm
    N, L = 32, 100.0
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.zeros((N, N, N), device="cuda"))
        with pytest.raises(cm.CosmoError):
            ctx.bispectrum(dk, N, L, 12, 0.0, 2 * np.pi / L, "NGP", 1.0)


# ------------------------------------------------- direct (brute-force) estimator
def _direct_case(N, L, seed, scheme, mas):
    import cosmicatlas_amd as cm
    rng = np.random.default_rng(seed)
    mesh = rng.poisson(2.0, (N, N, N)).astype(np.float32)
    n = float(mesh.astype(np.float64).sum())
    return cm, mesh, n


@pytest.mark.parametrize("N,nb,Dmin,ndel,scheme,mas", [
    (16, 13, 1.0, 1.0, "NGP", False),   # every mode, both Nyquist copies
    (16, 6, 0.5, 1.7, "CIC", True),     # fractional edges (FP64 bin rule)
    (24, 9, 2.0, 1.0, "TSC", True),     # modes below the first edge left out
])
def test_direct_matches_oracle(orc, N, nb, Dmin, ndel, scheme, mas):
    """cm_bispectrum_direct against or_bk_direct on the same complex64 half
    spectrum: counts bit-exact, sums to FP64 summation-order rounding."""
    L = 200.0
    kf = 2 * np.pi / L
    cm, mesh, n = _direct_case(N, L, N + nb, scheme, mas)
    sid = {"NGP": 0, "CIC": 1, "TSC": 2, "PCS": 3}[scheme]
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=n / N ** 3)
        out = ctx.bispectrum_direct(dk, N, L, nb, Dmin * kf, ndel * kf, scheme, n, mas_correction=mas)
        dkh = dk.cpu().numpy().astype(np.complex128)
    b, sn, cnt, ps, pc = orc.bk_direct(dkh, Dmin, ndel, nb, sid, mas)
    assert np.array_equal(out["ntri"], cnt)
    assert np.array_equal(out["nk"], pc)
    assert cnt.sum() > 1000
    # |Re d1 d2 d3| <= ((|d1|^2+|d2|^2+|d3|^2)/3)^1.5 per term bounds the scale of each sum
    scale = np.where(cnt > 0, cnt * (sn / np.maximum(cnt, 1) / 3) ** 1.5, 0)
    assert np.all(np.abs(out["bsum"] - b) <= 1e-11 * scale + 1e-300)
    np.testing.assert_allclose(out["snsum"], sn, rtol=1e-12, atol=0)
    V = L ** 3
    ok = pc > 0
    np.testing.assert_allclose(out["pk"][ok], ps[ok] / pc[ok] / (n * n / V) - V / n, rtol=1e-11)
    if N == 16 and nb == 13:
        h = N // 2
        assert cnt.sum() == ((N + 1) ** 2 - h * (h + 1)) ** 3 - 3 * (N + 1) ** 3 + 2


def test_direct_single_particle_is_pure_shot_noise():
    """One object (NGP, one cell): d_k = exp(-i k.x) for k != 0, so P = V/n^2 - V/n = 0
    and B = V^2/n^3 minus the Poisson shot noise 3 V^2/n^3 - 3V^2/n^2 + V^2/n^2 = 0."""
    import cosmicatlas_amd as cm
    N, L = 16, 50.0
    mesh = np.zeros((N, N, N), np.float32)
    mesh[3, 11, 6] = 1.0
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=1.0 / N ** 3)
        out = ctx.bispectrum_direct(dk, N, L, 12, 2 * np.pi / L, 2 * np.pi / L, "NGP", 1.0, mas_correction=False)
    V = L ** 3
    ok = out["ntri"] > 0
    assert ok.sum() > 100
    assert np.max(np.abs(out["bk"][ok])) < 1e-5 * V ** 2
    assert np.max(np.abs(out["pk"])) < 1e-5 * V


def test_direct_equals_shell_transform_estimator():
    """Below N/3 (no Nyquist component, no wrapped triangle) the direct sums are
    the shell-transform estimator's over every ordering of the bin triple: same
    B for every bin, and ntri = (number of distinct orderings) x closed triangles."""
    import itertools
    import cosmicatlas_amd as cm
    N, L = 32, 120.0
    kf = 2 * np.pi / L
    rng = np.random.default_rng(9)
    mesh = rng.poisson(3.0, (N, N, N)).astype(np.float32)
    n = float(mesh.astype(np.float64).sum())
    ns = 8  # shells [s, s+1) k_f + 1 k_f: last edge 9 < 32/3
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.from_numpy(mesh).cuda(), subtract=n / N ** 3)
        fast = ctx.bispectrum(dk, N, L, ns, kf, kf, "CIC", n, mas_correction=True, sn_correction=False)
        direct = ctx.bispectrum_direct(dk, N, L, ns, kf, kf, "CIC", n, mas_correction=True, sn_correction=False)
    checked = 0
    for (i, j, k), bk, nt in zip(fast["ijk"], fast["bk"], fast["ntri"]):
        a, b, c = sorted((i, j, k))
        perms = len(set(itertools.permutations((a, b, c))))
        assert direct["ntri"][a, b, c] == perms * nt, (i, j, k)
        if nt:
            assert abs(direct["bk"][a, b, c] - bk) <= 2e-4 * (abs(bk) + 1e-3 * np.abs(fast["bk"]).max())
            checked += 1
    assert checked > 50
