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


def test_shells_beyond_third_of_grid_rejected():
    import cosmicatlas_amd as cm
    N, L = 32, 100.0
    with cm.Context(0) as ctx:
        dk = ctx.fft_r2c(torch.zeros((N, N, N), device="cuda"))
        with pytest.raises(cm.CosmoError):
            ctx.bispectrum(dk, N, L, 12, 0.0, 2 * np.pi / L, "NGP", 1.0)
