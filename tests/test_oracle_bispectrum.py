"""Pins of the oracle's direct bispectrum (or_bk_direct, the restatement of
FftwFunctions::bispectrum_fkp, Source/FftwFunctions.cpp:3105-3297) against
things that do not share its code (CPU only):

* a closed form for the number of ordered pairs on the (N+1)^3 grid,
* the phase closure d1 d2 d3 = 1 of a single point field (checks the
  half-spectrum lookup and its conjugation on every octant),
* an enumeration of closed triangles on numpy's full complex FFT grid for bins
  below N/2 (no Nyquist component), with and without the MAS division.
"""
import itertools

import numpy as np
import pytest


def _total_pairs(N):
    """Ordered (q1, q2) in [-h, h]^3 with q1 + q2 in the box, none of q1, q2,
    q3 zero: per axis (N+1)^2 - h(h+1) pairs, minus 3 (N+1)^3 - 2 (inclusion-
    exclusion of q1 = 0, q2 = 0, q1 + q2 = 0, which meet only at (0, 0))."""
    h = N // 2
    per_axis = (N + 1) ** 2 - h * (h + 1)
    return per_axis ** 3 - 3 * (N + 1) ** 3 + 2


@pytest.mark.parametrize("N", [2, 4, 6, 8, 10])
def test_total_pairs_closed_form(orc, N):
    rng = np.random.default_rng(N)
    m = rng.standard_normal((N, N, N))
    dk = orc.fft_r2c(m)
    nb = int(np.ceil(np.sqrt(3) * N / 2)) + 1  # every nonzero mode in a bin
    _, _, cnt, _, pcnt = orc.bk_direct(dk, 1, 1, nb)
    assert cnt.sum() == _total_pairs(N)
    assert pcnt.sum() == (N + 1) ** 3 - 1
    # only the sorted corner a <= b <= c is written
    a, b, c = np.nonzero(cnt)
    assert np.all((a <= b) & (b <= c))


@pytest.mark.parametrize("x0", [(0, 0, 0), (1, 2, 3), (5, 0, 7), (7, 7, 1)])
def test_point_field_phase_closure(orc, x0):
    """delta(x) = [x == x0]: d(q) = exp(-2 pi i q.x0/N) and q1 + q2 + q3 = 0, so
    every term Re(d1 d2 d3) is exactly 1 and every |d|^2 is 1."""
    N = 8
    m = np.zeros((N, N, N))
    m[x0] = 1.0
    b, sn, cnt, ps, pc = orc.bk_direct(orc.fft_r2c(m), 1, 1, 8)
    np.testing.assert_allclose(b, cnt, rtol=0, atol=1e-9)
    np.testing.assert_allclose(sn, 3 * cnt, rtol=0, atol=1e-9)
    np.testing.assert_allclose(ps, pc, rtol=0, atol=1e-9)


def _numpy_triangles(mesh, nb, Dmin, ndel, scheme=None):
    """Closed triangles k1 + k2 + k3 = 0 on numpy's full complex FFT grid (q in
    [-N/2, N/2-1]) with all |q| < Dmin + nb ndel <= N/2, summed per ordered bin
    triple and then over the distinct permutations of each sorted triple."""
    N = mesh.shape[0]
    d = np.fft.fftn(mesh)
    f = np.fft.fftfreq(N, 1.0 / N).astype(int)
    qx, qy, qz = np.meshgrid(f, f, f, indexing="ij")
    q2 = qx ** 2 + qy ** 2 + qz ** 2
    if scheme is not None:
        d = d / (np.sinc(qx / N) * np.sinc(qy / N) * np.sinc(qz / N)) ** (scheme + 1)
    r = np.sqrt(q2)
    s = np.floor((r - Dmin) / ndel + 1e-12).astype(int)
    keep = (q2 > 0) & (r >= Dmin) & (s < nb)
    K = np.stack([qx[keep], qy[keep], qz[keep]], 1)
    S, D = s[keep], d[keep]
    h = N // 2
    cube = -np.ones((2 * N + 1,) * 3, int)  # q offset by N, holds the index into K
    cube[K[:, 0] + N, K[:, 1] + N, K[:, 2] + N] = np.arange(len(K))
    T = np.zeros((nb,) * 3)
    B = np.zeros((nb,) * 3)
    for i1 in range(len(K)):
        k3 = -(K[i1][None, :] + K)
        idx = cube[k3[:, 0] + N, k3[:, 1] + N, k3[:, 2] + N]
        ok = idx >= 0
        i2 = np.nonzero(ok)[0]
        i3 = idx[ok]
        np.add.at(T, (S[i1], S[i2], S[i3]), 1)
        np.add.at(B, (S[i1], S[i2], S[i3]), (D[i1] * D[i2] * D[i3]).real)
    Ts, Bs = np.zeros_like(T), np.zeros_like(B)
    for a, b, c in itertools.product(range(nb), repeat=3):
        if a <= b <= c:
            for p in set(itertools.permutations((a, b, c))):
                Ts[a, b, c] += T[p]
                Bs[a, b, c] += B[p]
    assert Dmin + nb * ndel <= h
    return Ts, Bs


@pytest.mark.parametrize("scheme", [None, 0, 2])
def test_matches_numpy_triangles(orc, scheme):
    N = 12
    rng = np.random.default_rng(7)
    mesh = rng.standard_normal((N, N, N))
    Ts, Bs = _numpy_triangles(mesh, 5, 1, 1, scheme)
    b, _, cnt, _, _ = orc.bk_direct(orc.fft_r2c(mesh), 1, 1, 5, scheme or 0, scheme is not None)
    assert np.array_equal(cnt, Ts.astype(np.int64))
    np.testing.assert_allclose(b, Bs, rtol=1e-10, atol=1e-9 * np.abs(Bs).max())


def test_fractional_bins_match_numpy(orc):
    """Bin edges that are not integers (FP64 rule)."""
    N = 12
    mesh = np.random.default_rng(8).standard_normal((N, N, N))
    Ts, Bs = _numpy_triangles(mesh, 3, 0.5, 1.7)
    b, _, cnt, _, _ = orc.bk_direct(orc.fft_r2c(mesh), 0.5, 1.7, 3)
    assert np.array_equal(cnt, Ts.astype(np.int64))
    np.testing.assert_allclose(b, Bs, rtol=1e-10, atol=1e-9 * np.abs(Bs).max())
