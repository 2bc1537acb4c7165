"""Pins of the FP64 oracle against independent mathematics (CPU only).

Each test checks the oracle against something that does not share its code:
brute-force DFTs and numpy.fft, closed-form MAS stencils, an independent
mask-based enumeration of the estimator's modes with integer shell edges,
closed-form single-mode and shot-noise spectra, numpy.linalg.eigvalsh and a
spectral (Fourier-multiplier) restatement of the finite-difference Hessian.
"""
import numpy as np
import pytest

TWO_PI = 2.0 * np.pi


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 24, 25, 27, 49, 96])
def test_fft_1d_bruteforce(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n)
    np.testing.assert_allclose(orc.fft_1d(x, -1), F @ x, atol=1e-12 * max(1, n))
    np.testing.assert_allclose(orc.fft_1d(x, +1), F.conj() @ x, atol=1e-12 * max(1, n))


@pytest.mark.parametrize("N", [4, 6, 8, 12, 16, 24])
def test_fft_r2c_matches_numpy(orc, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N))
    out = orc.fft_r2c(x)
    ref = np.fft.rfftn(x)
    np.testing.assert_allclose(out, ref, atol=1e-11 * N ** 1.5)
    # c2r divides by N^3 and inverts
    np.testing.assert_allclose(orc.fft_c2r(out, N), x, atol=1e-12 * N)


def test_fft_closed_forms(orc):
    N = 8
    d = np.zeros((N, N, N)); d[0, 0, 0] = 1.0          # delta -> constant 1
    np.testing.assert_allclose(orc.fft_r2c(d), np.ones((N, N, N // 2 + 1)), atol=1e-14)
    a, b, c = 1, 3, 2                                   # plane wave -> one mode
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x = np.cos(TWO_PI * (a * i + b * j + c * k) / N)
    h = orc.fft_r2c(x)
    exp = np.zeros_like(h); exp[a, b, c] = N ** 3 / 2
    np.testing.assert_allclose(h, exp, atol=1e-10)
    # Parseval on the half spectrum (kz=0 and kz=N/2 planes once, others twice)
    rng = np.random.default_rng(1)
    y = rng.standard_normal((N, N, N))
    H = orc.fft_r2c(y)
    w = np.full(N // 2 + 1, 2.0); w[0] = w[-1] = 1.0
    assert np.isclose((np.abs(H) ** 2 * w).sum() / N ** 3, (y ** 2).sum(), rtol=1e-12)


# --------------------------------------------------------------------------- MAS
def test_ngp_counts_match_histogram(orc):
    rng = np.random.default_rng(3)
    N, L = 32, 500.0
    d = L / N
    pos = rng.uniform(0, L, (20000, 3))
    # keep particles clear of cell edges so an edge-based histogram is unambiguous
    frac = pos / d - np.floor(pos / d)
    pos = pos[np.all((frac > 1e-6) & (frac < 1 - 1e-6), axis=1)]
    mesh = orc.mas(pos, N, L, orc.NGP)
    edges = [np.linspace(0, L, N + 1)] * 3
    H, _ = np.histogramdd(pos, bins=edges)
    assert np.array_equal(mesh, H)
    assert mesh.sum() == len(pos)


@pytest.mark.parametrize("s", [-0.49, -0.2, 0.0, 0.13, 0.4999])
def test_tsc_cic_single_particle_stencil(orc, s):
    N, L = 16, 16.0          # d = 1
    c = np.array([5, 9, 12])
    p = (c + 0.5 + s).reshape(1, 3)
    tsc = orc.mas(p, N, L, orc.TSC)
    w1 = np.array([0.5 * (0.5 - s) ** 2, 0.75 - s * s, 0.5 * (0.5 + s) ** 2])
    exp = np.zeros((N, N, N))
    exp[c[0] - 1:c[0] + 2, c[1] - 1:c[1] + 2, c[2] - 1:c[2] + 2] = np.einsum("i,j,k->ijk", w1, w1, w1)
    np.testing.assert_allclose(tsc, exp, atol=1e-15)
    cic = orc.mas(p, N, L, orc.CIC)
    if s >= 0:
        w = {0: 1 - s, 1: s}
    else:
        w = {-1: -s, 0: 1 + s}
    exp = np.zeros((N, N, N))
    for a, wa in w.items():
        for b, wb in w.items():
            for e, we in w.items():
                exp[c[0] + a, c[1] + b, c[2] + e] += wa * wb * we
    np.testing.assert_allclose(cic, exp, atol=1e-15)


def test_pcs_at_node_center(orc):
    N, L = 16, 16.0
    p = np.array([[4.5, 7.5, 10.5]])
    m = orc.mas(p, N, L, orc.PCS)
    w1 = np.array([1 / 6, 4 / 6, 1 / 6])
    exp = np.zeros((N, N, N))
    exp[3:6, 6:9, 9:12] = np.einsum("i,j,k->ijk", w1, w1, w1)
    np.testing.assert_allclose(m, exp, atol=1e-15)


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_mas_conservation_translation_wrap(orc, scheme):
    rng = np.random.default_rng(10 + scheme)
    N, L = 16, 100.0
    d = L / N
    pos = rng.uniform(0, L, (3000, 3))
    w = rng.uniform(0.5, 2.0, 3000)
    m = orc.mas(pos, N, L, scheme, weights=w)
    assert np.isclose(m.sum(), w.sum(), rtol=1e-13)
    shifted = pos.copy(); shifted[:, 1] = (shifted[:, 1] + d) % L
    m2 = orc.mas(shifted, N, L, scheme, weights=w)
    np.testing.assert_allclose(m2, np.roll(m, 1, axis=1), atol=1e-11)
    # periodic images land on the same mesh
    m3 = orc.mas(pos + np.array([L, -L, 2 * L]), N, L, scheme, weights=w)
    np.testing.assert_allclose(m3, m, atol=1e-11)


def test_mas_soa_equals_aos_and_f32_wrap(orc):
    rng = np.random.default_rng(5)
    N, L = 16, 1000.0
    pos = rng.uniform(0, L, (500, 3)).astype(np.float32)
    a = orc.mas(pos, N, L, orc.TSC)
    b = orc.mas(pos.T.copy(), N, L, orc.TSC, soa=True)
    np.testing.assert_array_equal(a, b)
    # float32 wrap: x = -1e-6 wraps to exactly 0 in float32 (L + x rounds to L),
    # to the last cell in float64
    t32 = orc.mas(np.array([[-1e-6, 1.0, 1.0]], np.float32), N, L, orc.NGP)
    t64 = orc.mas(np.array([[-1e-6, 1.0, 1.0]], np.float64), N, L, orc.NGP)
    assert t32[0, 0, 0] == 1 and t64[N - 1, 0, 0] == 1


def test_interlaced_single_particle(orc):
    """Interlaced mesh = particle shifted by +d/2 on every axis (closed form)."""
    N, L = 16, 16.0
    p = np.array([[5.2, 9.7, 1.1]])
    a = orc.mas(p, N, L, orc.CIC, interlace_shift=0.5)
    b = orc.mas(p + 0.5, N, L, orc.CIC)
    np.testing.assert_allclose(a, b, atol=1e-15)


# --------------------------------------------------------------------------- P(k)
def _stored_mode_mask(N):
    """Which stored r2c modes power_spectrum_fkp visits, as a mask over the
    N x N x (N/2+1) half grid (independent of the octant-loop formulation)."""
    ix = np.arange(N)[:, None, None]
    iy = np.arange(N)[None, :, None]
    iz = np.arange(N // 2 + 1)[None, None, :]
    h = N // 2
    xl, xh = ix < h, ix > h
    yl, yh = iy < h, iy > h
    m = (xl & yl & ((ix + iy + iz) > 0)) | (xl & yh & (iz > 0)) \
        | (xh & yl & ((iy > 0) | (iz > 0))) | (xh & yh & (iz > 0))
    return m


def _folded_q(N):
    ix = np.arange(N); fx = np.minimum(ix, N - ix)
    iz = np.arange(N // 2 + 1)
    return fx[:, None, None] ** 2 + fx[None, :, None] ** 2 + iz[None, None, :] ** 2


@pytest.mark.parametrize("N,ndel,dmin", [(16, 1, 0), (32, 1, 1), (32, 2, 0), (24, 3, 0), (48, 2, 1)])
def test_mode_counts_exact(orc, N, ndel, dmin):
    L = 500.0
    kf = TWO_PI / L
    ns = N // ndel // 2
    o = orc.pk_opts(N, L, ns, dmin * kf, ndel * kf, 1, 0.0, 1.0)
    _, _, _, nm = orc.pk_shells(np.ones((N, N, N // 2 + 1), complex), o)
    mask = _stored_mode_mask(N)
    q = _folded_q(N)
    exp = np.zeros(ns, np.int64)
    for s in range(ns):
        lo, hi = (s * ndel + dmin) ** 2, ((s + 1) * ndel + dmin) ** 2
        exp[s] = np.count_nonzero(mask & (q >= lo) & (q < hi))
    np.testing.assert_array_equal(nm, exp)


@pytest.mark.parametrize("N,dmin,nlog", [(128, 1.0, 6), (64, 1.5, 10), (48, 2.0, 7), (32, 0.5, 5)])
def test_log_shell_counts_exact(orc, N, dmin, nlog):
    """type_of_binning = log: shell s = floor(log10(|k|/kmin)/Deltal) + 1
    (FftwFunctions.cpp:1103-1108), Deltal = log10(kmax/kmin)/N_log_bins
    (Params.cpp:4106).  Independent count: the edges
    q_s = (DeltaKmin 10^((s-1) Deltal))^2 in 60-digit decimal arithmetic
    against the integer q = |k/k_f|^2 of every visited mode; an edge within
    1e-30 of an integer is that integer (the exact rule puts a mode on an edge
    in the shell above).  N = 128, DeltaKmin = 1, 6 bins puts every edge on a
    power of 4 (10^Deltal = 2)."""
    import decimal
    D = decimal.Decimal
    decimal.getcontext().prec = 60
    L = 400.0
    kf = TWO_PI / L
    dlog = np.log10((N / 2) / dmin) / nlog
    o = orc.pk_opts(N, L, nlog, dmin * kf, 0.0, 1, 0.0, 1.0, log_binning=True, dlog=dlog)
    _, _, _, nm = orc.pk_shells(np.ones((N, N, N // 2 + 1), complex), o)
    dl = (D(N // 2) / D(str(dmin))).log10() / D(nlog)
    edges = [(D(str(dmin)) * (D(10) ** ((D(s) - 1) * dl))) ** 2 for s in range(nlog + 1)]
    edges = [e.to_integral_value() if abs(e - e.to_integral_value()) < D("1e-30") * max(e, D(1)) else e
             for e in edges]
    q = _folded_q(N)[_stored_mode_mask(N)]
    vals, cnt = np.unique(q, return_counts=True)
    exp = np.zeros(nlog, np.int64)
    for v, c in zip(vals.tolist(), cnt.tolist()):
        for s in range(nlog):
            if edges[s] <= v < edges[s + 1]:
                exp[s] += c
    np.testing.assert_array_equal(nm, exp)
    if (N, dmin, nlog) == (128, 1.0, 6):
        # edges 4^(s-1): shell s >= 1 holds 4^(s-1) <= q < 4^s
        assert [int(e) for e in edges[1:]] == [4 ** (s - 1) for s in range(1, nlog + 1)]


@pytest.mark.parametrize("dmin,logb", [(0, False), (1, False), (1, True)])
def test_window_mode_set(orc, dmin, logb):
    """The window shells of the random branch visit the zero mode and the
    stored modes outside the first octant (ix, iy < N/2), not the first
    octant's other modes.  Independent mask over the stored half grid; and a
    constant field, whose only mode is k = 0 with d = c N^3, gives
    (c N^3)^2 / normal in shell 0 when that shell holds k = 0."""
    N, L, ndel = 16, 200.0, 1
    kf = TWO_PI / L
    ns = N // 2
    lk = dict(log_binning=True, dlog=np.log10(N / 2 / dmin) / ns) if logb else {}
    o = orc.pk_opts(N, L, ns, dmin * kf, ndel * kf, 1, 0.0, 2.0, mas_correction=True, **lk)
    rng = np.random.default_rng(4)
    dk = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    pw, nm = orc.window_shells(dk, o)
    ix = np.arange(N)[:, None, None]
    iy = np.arange(N)[None, :, None]
    mask = _stored_mode_mask(N) & ~((ix < N // 2) & (iy < N // 2))
    mask[0, 0, 0] = True
    q = _folded_q(N)
    fx = np.minimum(np.arange(N), N - np.arange(N))
    kz = np.arange(N // 2 + 1)
    corr = (np.sinc(fx / N)[:, None, None] * np.sinc(fx / N)[None, :, None] * np.sinc(kz / N)[None, None, :]) ** 2
    P = np.abs(dk) ** 2 / corr ** 2
    if logb:
        with np.errstate(divide="ignore"):
            t = np.log10(np.sqrt(q) / dmin) / lk["dlog"]
        sh = np.where(q == 0, 0, np.floor(t + 1e-12).astype(np.int64) + 1)
    else:
        with np.errstate(invalid="ignore"):
            sh = np.floor((np.sqrt(q) - dmin) / ndel + 1e-12).astype(np.int64)
            sh[np.sqrt(q) < dmin] = -1
    exp_n = np.array([np.count_nonzero(mask & (sh == s)) for s in range(ns)])
    exp_p = np.array([P[mask & (sh == s)].sum() for s in range(ns)]) / np.maximum(exp_n, 1) / 2.0
    np.testing.assert_array_equal(nm, exp_n)
    np.testing.assert_allclose(pw, exp_p, rtol=1e-12)
    c = 0.7
    const = np.zeros_like(dk)
    const[0, 0, 0] = c * N ** 3
    pw1, nm1 = orc.window_shells(const, o)
    if dmin == 0 or logb:
        assert np.isclose(pw1[0] * nm1[0] * 2.0, (c * N ** 3) ** 2, rtol=1e-14)
    else:
        assert np.all(pw1 == 0)


def test_float_shell_rule_survey_counts(orc):
    """The reference's float32 shell index changes per-shell mode counts by a
    summed 2,612 (of 548,955 modes) in 47 of 64 shells at N=128, L=500, ndel=1,
    DeltaKmin=0, relative to the exact rule -- an independent numpy emulation
    of FftwFunctions.cpp:1061-1099 gives the same figures."""
    N, L = 128, 500.0
    kf = TWO_PI / L
    ex = orc.pk_opts(N, L, 64, 0.0, kf, 1, 0.0, 1.0)
    fl = orc.pk_opts(N, L, 64, 0.0, kf, 1, 0.0, 1.0, float_shell_rule=True)
    z = np.ones((N, N, N // 2 + 1), complex)
    n_ex = orc.pk_shells(z, ex)[3]
    n_fl = orc.pk_shells(z, fl)[3]
    assert n_ex.sum() == 548955
    # per-shell count differences between the two rules (the survey's metric:
    # sum over shells of |count_float - count_exact|, and shells that differ)
    assert int(np.abs(n_ex - n_fl).sum()) == 2612
    assert np.count_nonzero(n_ex != n_fl) == 47


def test_single_mode_power_closed_form(orc):
    N, L = 16, 100.0
    kf = TWO_PI / L
    A = 0.7
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    for (a, b, c) in [(1, 2, 3), (3, 1, 0), (0, 0, 5), (2, 5, 8)]:
        x = A * np.cos(TWO_PI * (a * i + b * j + c * k) / N)
        # 14 shells so that the z-Nyquist plane (|k| >= N/2 k_f) is inside the range
        NS = 14
        o = orc.pk_opts(N, L, NS, 0.0, kf, 0, 0.0, 1.0, mas_correction=False)
        p0, p2, p4, nm = orc.pk_shells(orc.fft_r2c(x), o)
        s = orc.shell_index(o, a, b, c)
        # on the z-Nyquist plane the mode and its conjugate are both stored and
        # both counted (FftwFunctions.cpp:1084 runs k up to N/2 inclusive)
        mult = 2 if c == N // 2 else 1
        assert 0 <= s < NS
        exp = np.zeros(NS); exp[s] = mult * (A * N ** 3 / 2) ** 2 / nm[s]
        np.testing.assert_allclose(p0, exp, rtol=1e-12, atol=1e-6)
        mu = c / np.sqrt(a * a + b * b + c * c)
        np.testing.assert_allclose(p2[s], 5 * 0.5 * (3 * mu * mu - 1) * exp[s], rtol=1e-12, atol=1e-6)
        # MAS deconvolution divides by (sinc_x sinc_y sinc_z)^(2(p+1)), TSC p=2
        o2 = orc.pk_opts(N, L, NS, 0.0, kf, 2, 0.0, 1.0, mas_correction=True)
        p0c = orc.pk_shells(orc.fft_r2c(x), o2)[0]
        sinc = np.prod([np.sinc(v / N) for v in (a, b, c)])
        np.testing.assert_allclose(p0c[s], exp[s] / sinc ** 6, rtol=1e-12)


def test_shot_noise_closed_form(orc):
    """A mode-free field leaves only the aliased shot-noise term: NGP gives -SN in
    every shell; CIC over the |n|=1 modes gives -SN(1-2/3 sin^2(pi/N))."""
    N, L, SN, npow = 16, 100.0, 3.5, 2.0
    kf = TWO_PI / L
    z = np.zeros((N, N, N // 2 + 1), complex)
    o = orc.pk_opts(N, L, N // 2, 0.0, kf, 0, SN, npow, mas_correction=False)
    p0, _, _, nm = orc.pk_shells(z, o)
    np.testing.assert_allclose(p0[nm > 0], -SN, rtol=1e-14)
    # shell 0 = [k_f, 1.2 k_f) holds only the |n|=1 modes
    o1 = orc.pk_opts(N, L, 4, kf, 0.2 * kf, 1, SN, npow, mas_correction=False)
    p0, _, _, nm = orc.pk_shells(z, o1)
    s1 = np.sin(np.pi / N)
    assert nm[0] == 3
    assert np.isclose(p0[0], -SN * (1 - 2 / 3 * s1 ** 2), rtol=1e-14)


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_aliased_shot_noise_all_schemes(orc, scheme):
    """The aliased shot-noise factor per axis, C1(x) = sum_n W^2(k + 2 pi n/d),
    W = sinc^(p) (p = scheme + 1, the B-spline window of NGP/CIC/TSC/PCS), at
    x = pi i/N, against its brute-force sum over 400001 images plus the
    integral of the 1/n^2 tail for NGP (Jing 2005; the closed forms the
    estimator uses, Miscelanious.cpp:2532-2554).  Shells [i, i + 0.05) k_f with
    i = 2^j hold only the six axis vectors (r3(4^j) = 6, and the next integer
    |k|^2 = i^2 + 1 lies beyond (i + 0.05)^2 for i < 10), so the shell value of a
    mode-free field is -SN C1(pi i/N) C1(0)^2 exactly."""
    N, L, SN = 32, 100.0, 2.5
    kf = TWO_PI / L
    z = np.zeros((N, N, N // 2 + 1), complex)
    M = 200000
    n = np.arange(-M, M + 1)
    for i in (1, 2, 4, 8):
        o = orc.pk_opts(N, L, 1, i * kf, 0.05 * kf, scheme, SN, 1.0, mas_correction=False)
        p0, _, _, nm = orc.pk_shells(z, o)
        assert nm[0] == 3  # the +x, +y, +z axis modes of the half spectrum
        x = np.pi * i / N
        y = x + np.pi * n
        brute = np.sum((np.sin(y) / y) ** (2 * (scheme + 1)))
        if scheme == 0:  # sum over |n| > M of sin^2 x / (pi n)^2
            brute += np.sin(x) ** 2 * 2.0 / (np.pi ** 2 * (M + 0.5))
        np.testing.assert_allclose(p0[0], -SN * brute, rtol=1e-9)


# --------------------------------------------------------------------------- T-web
def test_eigen3_matches_eigvalsh(orc):
    rng = np.random.default_rng(7)
    for _ in range(300):
        A = rng.standard_normal((3, 3))
        m9 = A.T.reshape(9)            # m[i+3j] = A(i,j): column-major storage
        ev = orc.eigen3(m9)
        low = np.tril(A) + np.tril(A, -1).T   # GSL reads the lower triangle
        ref = np.sort(np.linalg.eigvalsh(low))[::-1]
        np.testing.assert_allclose(ev, ref, atol=1e-12 * np.abs(ref).max())
    # degenerate / diagonal cases
    np.testing.assert_allclose(orc.eigen3(np.diag([1.0, 1.0, 1.0]).reshape(9)), [1, 1, 1])
    np.testing.assert_allclose(orc.eigen3(np.diag([-2.0, 5.0, 0.0]).reshape(9)), [5, 0, -2])


def _fd_multiplier(N, L):
    d = L / N
    k = TWO_PI / L * np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N)
    return (8 * np.sin(k * d) - np.sin(2 * k * d)) / (6 * d), k


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_gradfindif_fourier_multiplier(orc, dim):
    N, L = 16, 50.0
    m = 3
    D, k = _fd_multiplier(N, L)
    c = np.arange(N)
    f1 = np.sin(TWO_PI * m * c / N)
    shape = [1, 1, 1]; shape[dim - 1] = N
    f = np.broadcast_to(f1.reshape(shape), (N, N, N)).copy()
    g = orc.gradfindif(f, L, dim)
    exp1 = D[m] * np.cos(TWO_PI * m * c / N)
    np.testing.assert_allclose(g, np.broadcast_to(exp1.reshape(shape), (N, N, N)), atol=1e-12)


def _spectral_tweb(delta, L, R, lth):
    """Independent restatement: numpy FFT, Gaussian*(-1/k^2), FD Hessian as the
    Fourier multiplier -D(k_a)D(k_b), eigvalsh."""
    N = delta.shape[0]
    D, k = _fd_multiplier(N, L)
    kz = k[: N // 2 + 1].copy(); kz[N // 2] = abs(kz[N // 2])
    Dz = D[: N // 2 + 1]
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    with np.errstate(divide="ignore"):
        f = np.where(k2 > 0, -np.exp(-0.5 * k2 * R * R) / k2, 0.0)
    phik = np.fft.rfftn(delta) * f
    Ds = np.meshgrid(D, D, Dz, indexing="ij")
    H = np.empty((N, N, N, 3, 3))
    for a in range(3):
        for b in range(3):
            H[..., a, b] = np.fft.irfftn(-Ds[a] * Ds[b] * phik, s=(N, N, N))
    ev = np.linalg.eigvalsh(H)[..., ::-1]
    labels = 4 - (ev > lth).sum(-1)
    return ev, labels


def test_tweb_matches_spectral_restatement(orc):
    rng = np.random.default_rng(11)
    N, L, R, lth = 16, 64.0, 6.0, 0.05
    delta = rng.standard_normal((N, N, N))
    labels, lam, _, counts = orc.tweb(delta, L, R, lth)
    ev, lab_ref = _spectral_tweb(delta, L, R, lth)
    lam_c = np.moveaxis(lam, 0, -1)
    np.testing.assert_allclose(lam_c, ev, atol=1e-10 * np.abs(ev).max())
    margin = np.abs(ev - lth).min(-1)
    ok = margin > 1e-8
    assert np.array_equal(labels[ok], lab_ref[ok])
    assert counts.sum() == N ** 3
    assert np.array_equal(np.bincount(labels.ravel(), minlength=5)[1:], counts)


def test_tweb_plane_wave(orc):
    """delta = A cos(k x) along x: one non-zero eigenvalue lambda = A W D(k)^2/k^2
    cos(kx), the other two zero -- a closed form for eigenvalues and labels."""
    N, L, R, A, m = 16, 64.0, 3.0, 0.4, 2
    D, k = _fd_multiplier(N, L)
    x = np.arange(N)
    c = np.cos(TWO_PI * m * x / N)
    delta = np.broadcast_to((A * c)[:, None, None], (N, N, N)).copy()
    labels, lam, _, _ = orc.tweb(delta, L, R, 0.0)
    kk = k[m]
    lam_exp = A * np.exp(-0.5 * kk * kk * R * R) * D[m] ** 2 / kk ** 2 * c
    l1 = np.maximum(lam_exp, 0.0)
    l3 = np.minimum(lam_exp, 0.0)
    np.testing.assert_allclose(lam[0], np.broadcast_to(l1[:, None, None], (N, N, N)), atol=1e-12)
    np.testing.assert_allclose(lam[2], np.broadcast_to(l3[:, None, None], (N, N, N)), atol=1e-12)
    np.testing.assert_allclose(lam[1], 0.0, atol=1e-12)
    # the two zero eigenvalues sit on lambda_th=0 up to round-off, so classify
    # just above and just below it: +eps -> sheets/voids, -eps -> knots/filaments
    sel = np.abs(c) > 1e-6
    lab_hi = orc.tweb(delta, L, R, 1e-9, want_lambdas=False)[0]
    lab_lo = orc.tweb(delta, L, R, -1e-9, want_lambdas=False)[0]
    assert np.array_equal(lab_hi[sel, 0, 0], np.where(c > 0, 3, 4)[sel])
    assert np.array_equal(lab_lo[sel, 0, 0], np.where(c > 0, 1, 2)[sel])
    assert np.all(lab_hi == lab_hi[:, :1, :1]) and np.all(lab_lo == lab_lo[:, :1, :1])


@pytest.mark.parametrize("q,R", [((2, 1, 3), 0.0), ((1, -2, 1), 0.0), ((3, 0, -1), 5.0)])
def test_tweb_spectral_plane_wave(orc, q, R):
    """Spectral tidal tensor of delta = A cos(2 pi q.i / N): T_ab = A W(k)
    k_a k_b / k^2 cos(.) exactly (k = q k_f), so one eigenvalue is A W cos(.)
    along q and two are zero; a negative component of q pins the sign of the
    off-diagonal products.  With lambda_th = -0.3 A the zero eigenvalues are
    above the threshold: label = 2 - [A W cos > -0.3 A]."""
    N, L, A = 16, 80.0, 0.7
    i = np.arange(N)
    ph = TWO_PI * (q[0] * i[:, None, None] + q[1] * i[None, :, None] + q[2] * i[None, None, :]) / N
    c = np.cos(ph)
    delta = A * c
    kf = TWO_PI / L
    k2 = kf * kf * (q[0] ** 2 + q[1] ** 2 + q[2] ** 2)
    lam1 = A * np.exp(-0.5 * k2 * R * R) * c
    lth = -0.3 * A
    labels, lam, counts = orc.tweb_spectral(delta, L, R, lth)
    np.testing.assert_allclose(lam[0], np.maximum(lam1, 0.0), atol=1e-12)
    np.testing.assert_allclose(lam[1], 0.0, atol=1e-12)
    np.testing.assert_allclose(lam[2], np.minimum(lam1, 0.0), atol=1e-12)
    sel = np.abs(lam1 - lth) > 1e-9
    assert np.array_equal(labels[sel], (2 - (lam1 > lth).astype(int))[sel])
    assert counts.sum() == N ** 3


def test_tweb_spectral_trace_is_smoothed_delta(orc):
    """The trace of T_ab = k_a k_b / k^2 W delta_k is W delta_k without its k = 0
    mode: the three eigenvalues sum to the smoothed, mean-free field (numpy's
    FFT), on a random field; and the spectral tensor of the oracle agrees with
    a numpy restatement built from numpy's FFT (all nine components)."""
    rng = np.random.default_rng(5)
    N, L, R = 16, 60.0, 4.0
    delta = rng.standard_normal((N, N, N))
    labels, lam, counts = orc.tweb_spectral(delta, L, R, 0.0)
    kf = TWO_PI / L
    kx = np.fft.fftfreq(N, 1.0 / N) * kf
    kz = np.fft.rfftfreq(N, 1.0 / N) * kf
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    dk = np.fft.rfftn(delta) * np.exp(-0.5 * k2 * R * R)
    dk[0, 0, 0] = 0.0
    np.testing.assert_allclose(lam.sum(axis=0), np.fft.irfftn(dk, s=(N, N, N)), atol=1e-12)
    # the Nyquist index of numpy's fftfreq is -N/2; the reference's coords take
    # +N/2; the mixed products differ only there, and a field with no power at
    # the Nyquist planes (zeroed here) makes both agree
    K = [KX, KY, KZ]
    K[0] = np.where(np.abs(KX) >= N // 2 * kf, 0.0, KX)
    K[1] = np.where(np.abs(KY) >= N // 2 * kf, 0.0, KY)
    nyq = (np.abs(KX) >= N // 2 * kf) | (np.abs(KY) >= N // 2 * kf) | (KZ >= N // 2 * kf)
    dk2 = np.where(nyq, 0.0, dk)
    delta2 = np.fft.irfftn(np.fft.rfftn(delta) * np.where(nyq, 0.0, 1.0), s=(N, N, N))
    _, lam2, _ = orc.tweb_spectral(delta2, L, R, 0.0)
    H = np.empty((N, N, N, 3, 3))
    with np.errstate(invalid="ignore", divide="ignore"):
        for a in range(3):
            for b in range(3):
                m = np.where(k2 > 0, K[a] * K[b] / k2, 0.0)
                H[..., a, b] = np.fft.irfftn(m * dk2, s=(N, N, N))
    ev = np.linalg.eigvalsh(H)[..., ::-1]
    np.testing.assert_allclose(np.moveaxis(lam2, 0, -1), ev, atol=1e-11)
