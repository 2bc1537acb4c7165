"""GPU P(k) shells against the oracle estimator on the same spectrum: mode
counts bit-exact (exact and float32 rules), sums to FP64 round-off."""
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


def _spectrum(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))


@pytest.mark.parametrize("N,ndel,dmin,scheme", [(16, 1, 0, 0), (32, 1, 1, 1), (64, 2, 0, 2), (48, 3, 0, 3),
                                                (128, 1, 0, 1), (96, 1, 1, 2)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_shells_match_oracle(ctx, orc, N, ndel, dmin, scheme, dtype):
    L = 500.0
    kf = 2 * np.pi / L
    ns = N // ndel // 2
    dk = _spectrum(N, N + ndel)
    SN, normal = 2.5, 3.0
    o = orc.pk_opts(N, L, ns, dmin * kf, ndel * kf, scheme, SN, normal)
    r0, r2, r4, rn = orc.pk_shells(dk, o)
    dkt = torch.from_numpy(dk).to(dtype).cuda()
    p0, p2, p4, nm = ctx.pk_shells(dkt, N, L, ns, dmin * kf, ndel * kf, scheme, SN, normal)
    assert np.array_equal(nm, rn)
    tol = 1e-11 if dtype == torch.complex128 else 2e-6
    scale = np.abs(r0).max()
    assert np.max(np.abs(p0 - r0)) < tol * scale
    assert np.max(np.abs(p2 - r2)) < tol * scale
    assert np.max(np.abs(p4 - r4)) < tol * scale


@pytest.mark.parametrize("N,L,ndel", [(128, 500.0, 1), (64, 1000.0, 2)])
def test_float_shell_rule_counts(ctx, orc, N, L, ndel):
    kf = 2 * np.pi / L
    ns = N // ndel // 2
    ones = torch.ones((N, N, N // 2 + 1), dtype=torch.complex64, device="cuda")
    for rule in (False, True):
        o = orc.pk_opts(N, L, ns, 0.0, ndel * kf, 1, 0.0, 1.0, float_shell_rule=rule)
        rn = orc.pk_shells(np.ones((N, N, N // 2 + 1), complex), o)[3]
        nm = ctx.pk_shells(ones, N, L, ns, 0.0, ndel * kf, 1, 0.0, 1.0, float_shell_rule=rule)[3]
        assert np.array_equal(nm, rn), rule


def test_single_mode_closed_form_on_gpu(ctx):
    """A cos wave through the GPU FFT and shells lands in one shell with
    P = (A N^3/2)^2 / n_s (normal_power 1, no SN, no MAS correction)."""
    N, L, A = 32, 100.0, 0.3
    kf = 2 * np.pi / L
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x = A * np.cos(2 * np.pi * (3 * i + 2 * j + 5 * k) / N)
    h = ctx.fft_r2c(torch.from_numpy(x).cuda())
    p0, _, _, nm = ctx.pk_shells(h, N, L, N // 2, 0.0, kf, 0, 0.0, 1.0, mas_correction=False)
    s = int(np.floor(np.sqrt(9 + 4 + 25)))
    exp = np.zeros(N // 2)
    exp[s] = (A * N ** 3 / 2) ** 2 / nm[s]
    assert np.max(np.abs(p0 - exp)) < 1e-9 * exp.max()


@pytest.mark.parametrize("N,dmin,nlog", [(128, 1.0, 6), (64, 1.5, 10), (48, 2.0, 7), (96, 0.5, 12)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_log_shells_match_oracle(ctx, orc, N, dmin, nlog, dtype):
    """type_of_binning = log: integer |k|^2 thresholds on the GPU against the
    oracle's FP64 log10 rule (pinned by a 60-digit enumeration): mode counts
    bit-exact, including the modes that sit on an edge (N = 128, 6 bins)."""
    L = 400.0
    kf = 2 * np.pi / L
    dlog = np.log10((N / 2) / dmin) / nlog
    dk = _spectrum(N, 3 * N + nlog)
    SN, normal = 1.5, 2.0
    o = orc.pk_opts(N, L, nlog, dmin * kf, 0.0, 2, SN, normal, log_binning=True, dlog=dlog)
    r0, r2, r4, rn = orc.pk_shells(dk, o)
    p0, p2, p4, nm = ctx.pk_shells(torch.from_numpy(dk).to(dtype).cuda(), N, L, nlog, dmin * kf, 0.0, 2, SN, normal,
                                   log_binning=True, dlog=dlog)
    assert np.array_equal(nm, rn)
    tol = 1e-11 if dtype == torch.complex128 else 2e-6
    scale = np.abs(r0).max()
    for a, r in ((p0, r0), (p2, r2), (p4, r4)):
        assert np.max(np.abs(a - r)) < tol * scale


def test_log_binning_argument_checks(ctx):
    import cosmicatlas_amd as cm
    dk = torch.zeros((16, 16, 9), dtype=torch.complex64, device="cuda")
    with pytest.raises(cm.CosmoError):  # kmin = 0 has no logarithm
        ctx.pk_shells(dk, 16, 100.0, 4, 0.0, 0.0, 1, 0.0, 1.0, log_binning=True, dlog=0.3)
    with pytest.raises(cm.CosmoError):  # the float32 rule is the linear one's
        ctx.pk_shells(dk, 16, 100.0, 4, 0.1, 0.0, 1, 0.0, 1.0, log_binning=True, dlog=0.3, float_shell_rule=True)


@pytest.mark.parametrize("N,dmin,logb", [(32, 0, False), (48, 1, False), (64, 1, True)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_window_set_matches_oracle(ctx, orc, N, dmin, logb, dtype):
    """window_set: the zero mode and the stored modes outside the first octant
    (FftwFunctions.cpp:1248-1262, 1354-1568) against the oracle's window loop."""
    L = 300.0
    kf = 2 * np.pi / L
    ns = N // 2
    lk = dict(log_binning=True, dlog=float(np.log10((N / 2) / dmin) / ns)) if logb else {}
    dk = _spectrum(N, 5 * N + dmin)
    dk[0, 0, 0] = 40.0 + 0j  # a dominant zero mode, as a random catalogue's
    o = orc.pk_opts(N, L, ns, dmin * kf, kf, 1, 0.0, 3.0, **lk)
    rw, rn = orc.window_shells(dk, o)
    p0, _, _, nm = ctx.pk_shells(torch.from_numpy(dk).to(dtype).cuda(), N, L, ns, dmin * kf, kf, 1, 0.0, 3.0,
                                 window_set=True, **lk)
    assert np.array_equal(nm, rn)
    tol = 1e-11 if dtype == torch.complex128 else 2e-6
    assert np.max(np.abs(p0 - rw) / np.maximum(np.abs(rw), 1e-300)) < tol
