"""GPU T-web against the oracle: eigenvalues to float32 round-off, labels
identical outside the eigenvalue margin (those cells are counted)."""
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


def _smooth_field(N, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((N, N, N))
    k = np.fft.fftfreq(N)
    kz = np.fft.rfftfreq(N)
    w = np.exp(-20 * (k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2))
    return np.fft.irfftn(np.fft.rfftn(x) * w, s=(N, N, N)) * 30


@pytest.mark.parametrize("N,L,R,lth", [(32, 64.0, 3.0, 0.0), (48, 100.0, 4.0, 0.02), (64, 128.0, 0.0, -0.01)])
def test_tweb_matches_oracle(ctx, orc, N, L, R, lth):
    delta = _smooth_field(N, N).astype(np.float32)
    lab_ref, lam_ref, _, cnt_ref = orc.tweb(delta.astype(np.float64), L, R, lth)
    tol = 1e-4 * np.abs(lam_ref).max()
    labels, lam, st = ctx.tweb(torch.from_numpy(delta).cuda(), L, R, lth, margin_tol=tol, want_lambdas=True)
    lam = lam.cpu().numpy()
    labels = labels.cpu().numpy()
    assert np.max(np.abs(lam - lam_ref)) < 2e-5 * np.abs(lam_ref).max()
    margin = np.abs(lam_ref - lth).min(axis=0)
    ok = margin >= tol
    assert np.array_equal(labels[ok], lab_ref[ok])
    assert sum(st["counts"]) == N ** 3
    assert np.array_equal(np.bincount(labels.ravel(), minlength=5)[1:], st["counts"])
    # cells whose labels differ are all among the marginal cells the kernel counts
    assert int(np.count_nonzero(labels != lab_ref)) <= st["n_marginal"] + int((~ok).sum())


def test_tweb_from_phi_matches_oracle(ctx, orc):
    N, L = 32, 50.0
    phi = _smooth_field(N, 3)
    lab_ref, lam_ref, _ = orc.tweb_from_phi(phi, L, 0.0)
    labels, lam, _ = ctx.tweb_from_phi(torch.from_numpy(phi.astype(np.float32)).cuda(), L, 0.0, want_lambdas=True)
    assert np.max(np.abs(lam.cpu().numpy() - lam_ref)) < 2e-5 * np.abs(lam_ref).max()
    ok = np.abs(lam_ref).min(axis=0) > 1e-4 * np.abs(lam_ref).max()
    assert np.array_equal(labels.cpu().numpy()[ok], lab_ref[ok])


def test_poisson_matches_oracle(ctx, orc):
    N, L, R = 48, 100.0, 5.0
    delta = _smooth_field(N, 5)
    dk = orc.fft_r2c(delta)
    ref = orc.fft_c2r(orc.smooth_poisson(dk, N, L, R), N)
    phi = ctx.poisson(torch.from_numpy(delta.astype(np.float32)).cuda(), L, R).cpu().numpy()
    assert np.max(np.abs(phi - ref)) < 1e-5 * np.abs(ref).max()


def _tclass_ref(labels, codes):
    """get_Tclassification (Source/Cwclass.cpp:1983-2030), restated per cell."""
    combos = {12: (1, 2), 23: (2, 3), 34: (3, 4), 123: (1, 2, 3), 234: (2, 3, 4)}
    if len(codes) == 1:
        return np.full(labels.shape, codes[0], np.int32)
    out = np.zeros(labels.shape, np.int32)
    for ic, c in enumerate(codes):
        hit = (labels == c) if c < 5 else np.isin(labels, combos.get(c, ()))
        out[hit] = ic
    return out


@pytest.mark.parametrize("codes", [[234], [0], [1, 234], [12, 34], [4, 3, 2, 1], [123, 4], [0, 2], [13, 2]])
def test_tweb_tclass(codes):
    import cosmicatlas_amd as cm
    rng = np.random.default_rng(len(codes))
    lab = rng.integers(1, 5, 5000).astype(np.int8)
    with cm.Context(0) as ctx:
        out = ctx.tweb_tclass(torch.from_numpy(lab).cuda(), codes).cpu().numpy()
    assert np.array_equal(out, _tclass_ref(lab, codes))
    # hand-checked cells: knot (1), filament (2), sheet (3), void (4)
    if codes == [1, 234]:
        assert [int(_tclass_ref(np.array([v], np.int8), codes)[0]) for v in (1, 2, 3, 4)] == [0, 1, 1, 1]
    if codes == [12, 34]:
        assert [int(_tclass_ref(np.array([v], np.int8), codes)[0]) for v in (1, 2, 3, 4)] == [0, 0, 1, 1]


def test_tweb_lambda_variant_equals_count_variant(ctx):
    """Without the lambda output the marching kernel decides uncertain signs and
    margin bands in an inline FP64 branch; with it, every cell takes the
    out-of-line FP64 branch.  Both must give the same labels, class counts and
    marginal count, also when a large tol makes a good part of the mesh
    marginal."""
    N, L = 48, 100.0
    phi = torch.from_numpy(_smooth_field(N, 11).astype(np.float32)).cuda()
    for lth, frac in ((0.0, 1e-3), (0.01, 0.3)):
        _, lam, _ = ctx.tweb_from_phi(phi, L, lth, want_lambdas=True)
        tol = frac * float(lam.abs().max())
        la, _, sa = ctx.tweb_from_phi(phi, L, lth, margin_tol=tol, want_lambdas=True)
        lb, _, sb = ctx.tweb_from_phi(phi, L, lth, margin_tol=tol, want_lambdas=False)
        assert torch.equal(la, lb)
        assert sa == sb
        assert sum(sb["counts"]) == N ** 3
        if frac > 0.1:
            assert sb["n_marginal"] > N ** 3 // 20  # a good part of the mesh is marginal


def test_tweb_uncertain_signs_settled_in_fp64(ctx):
    """phi varying along x only: T_yy = T_zz = T_xy = ... = 0 exactly, so two
    eigenvalues of T - 0 I are exactly zero and every float sign test is
    uncertain; the FP64 count gives label 3 where T_xx > 0 and 4 elsewhere."""
    N, L = 32, 64.0
    x = np.arange(N)
    phi1 = np.cos(2 * np.pi * 3 * x / N) + 0.3 * np.sin(2 * np.pi * 5 * x / N)
    phi = np.broadcast_to(phi1[:, None, None], (N, N, N)).astype(np.float32)
    d = L / N
    c = np.array([1.0, -8.0, 0.0, 8.0, -1.0]) / (12 * d)
    p = phi1.astype(np.float64)
    g = sum(c[i] * np.roll(p, 2 - i) for i in range(5))
    txx = sum(c[i] * np.roll(g, 2 - i) for i in range(5))
    expect = np.where(txx > 0, 3, 4)
    clear = np.abs(txx) > 1e-4 * np.abs(txx).max()
    for want in (False, True):
        lab, _, st = ctx.tweb_from_phi(torch.from_numpy(phi.copy()).cuda(), L, 0.0, want_lambdas=want)
        lab = lab.cpu().numpy()
        assert np.array_equal(lab[clear][:, 0, 0], expect[clear])
        assert np.all(lab == lab[:, :1, :1])  # constant over y, z
        assert sum(st["counts"]) == N ** 3


@pytest.mark.parametrize("frac", [1e-4, 1e-2, 0.1])
def test_tweb_marginal_count_matches_eigenvalues(ctx, frac):
    """The marginal count (cells with an eigenvalue within tol of lambda_th)
    comes from the sign counts of T - (lambda_th -+ tol) I, certified in float
    or decided in FP64.  Reference: the same count taken from the returned
    eigenvalues, up to cells within float rounding of the tol boundary."""
    N, L, lth = 48, 100.0, 0.005
    phi = torch.from_numpy(_smooth_field(N, 21).astype(np.float32)).cuda()
    _, lam, _ = ctx.tweb_from_phi(phi, L, lth, want_lambdas=True)
    tol = frac * float(lam.abs().max())
    _, lam, st_l = ctx.tweb_from_phi(phi, L, lth, margin_tol=tol, want_lambdas=True)
    _, _, st = ctx.tweb_from_phi(phi, L, lth, margin_tol=tol, want_lambdas=False)
    m = (lam.double() - lth).abs().min(dim=0).values
    exact = int((m < tol).sum())
    edge = int(((m - tol).abs() < 1e-5 * tol + 1e-6 * float(lam.abs().max())).sum())
    assert exact > 0
    assert abs(st["n_marginal"] - exact) <= edge
    assert st["n_marginal"] == st_l["n_marginal"]


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("N,L,R,lth", [(32, 64.0, 3.0, 0.0), (48, 100.0, 4.0, 0.02)])
def test_spectral_hessian_matches_oracle(ctx, orc, dtype, N, L, R, lth):
    """hessian="spectral": T_ab = c2r(W k_a k_b / k^2 delta_k) (the reference's
    _USE_GFFT_EIGENV_ branch) against the oracle's restatement: eigenvalues to
    float32 (FP64) round-off, labels equal outside the margin, counts
    consistent.  The fields carry no power at the Nyquist planes (smoothing),
    where the mixed multipliers with +N/2 k_f are not Hermitian."""
    delta = _smooth_field(N, N + 1)
    lab_ref, lam_ref, cnt_ref = orc.tweb_spectral(delta, L, R, lth)
    scale = np.abs(lam_ref).max()
    if dtype == "f32":
        tol = 1e-4 * scale
        labels, lam, st = ctx.tweb(torch.from_numpy(delta.astype(np.float32)).cuda(), L, R, lth, margin_tol=tol,
                                   want_lambdas=True, hessian="spectral")
        err = 2e-5
    else:
        # the closed-form (trigonometric) eigenvalues carry errors up to about
        # sqrt(u) |lambda| where two eigenvalues nearly coincide (acos near
        # +-1): measured 1e-8 of the largest here, against the oracle's
        # iterative solve
        tol = 1e-6 * scale
        labels, lam, _, st = ctx.tweb_f64(torch.from_numpy(delta).cuda(), L, R, lth, margin_tol=tol,
                                          want_lambdas=True, hessian="spectral")
        err = 1e-7
    lam = lam.cpu().numpy()
    labels = labels.cpu().numpy()
    assert np.max(np.abs(lam - lam_ref)) < err * scale
    ok = np.abs(lam_ref - lth).min(axis=0) >= tol
    assert np.array_equal(labels[ok], lab_ref[ok])
    assert sum(st["counts"]) == N ** 3
    assert np.array_equal(np.bincount(labels.ravel(), minlength=5)[1:], st["counts"])
    assert int(np.count_nonzero(labels != lab_ref)) <= st["n_marginal"] + int((~ok).sum())


def test_spectral_hessian_plane_wave_and_rounding_tol(ctx):
    """Plane wave delta = A cos(2 pi q.i / N) (FP64): eigenvalues {A W cos, 0,
    0} in closed form, labels from them; the marginal count and the rounding
    tolerance of the spectral tensor."""
    import cosmicatlas_amd as cm
    N, L, A, q, R = 32, 100.0, 0.5, (3, -1, 2), 2.0
    i = np.arange(N)
    c = np.cos(2 * np.pi * (q[0] * i[:, None, None] + q[1] * i[None, :, None] + q[2] * i[None, None, :]) / N)
    kf = 2 * np.pi / L
    lam1 = A * np.exp(-0.5 * kf * kf * sum(x * x for x in q) * R * R) * c
    labels, lam, _, st = ctx.tweb_f64(torch.from_numpy(A * c).cuda(), L, R, -0.3 * A, want_lambdas=True,
                                      hessian="spectral")
    lam = lam.cpu().numpy()
    # a double zero eigenvalue is the worst case of the closed form: its roots
    # come out within about sqrt(u) |lambda_1| (measured 4e-9 here)
    eps = 1e-7 * A
    np.testing.assert_allclose(lam[0], np.maximum(lam1, 0.0), atol=eps)
    np.testing.assert_allclose(lam[1], 0.0, atol=eps)
    np.testing.assert_allclose(lam[2], np.minimum(lam1, 0.0), atol=eps)
    sel = np.abs(lam1 + 0.3 * A) > eps
    assert np.array_equal(labels.cpu().numpy()[sel], (2 - (lam1 > -0.3 * A).astype(int))[sel])
    # lambda_th = 0: two eigenvalues vanish analytically, so every cell is
    # within 1e-6 of the threshold (the closed form puts them within 4e-9)
    _, _, _, st1 = ctx.tweb_f64(torch.from_numpy(A * c).cuda(), L, R, 0.0, margin_tol=1e-6, hessian="spectral")
    assert st1["n_marginal"] == N ** 3
    _, _, _, st0 = ctx.tweb_f64(torch.from_numpy(A * c).cuda(), L, R, 0.0, margin_tol=cm.TOL_ROUNDING,
                                hessian="spectral")
    assert 0.0 < st0["margin_tol"] < 1e-12


def test_hessian_selector_errors(ctx):
    import ctypes as C
    import cosmicatlas_amd as cm
    from cosmicatlas_amd import _abi
    N, L = 16, 50.0
    delta = torch.zeros((N, N, N), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        ctx.tweb(delta, L, 0.0, 0.0, hessian="fourier")
    labels = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
    o = _abi.TwebOpts(N, L, 0.0, 0.0, 0.0, 1.0, 5)
    assert ctx.lib.cm_tweb(ctx.h, C.byref(o), C.c_void_p(delta.data_ptr()), C.c_void_p(labels.data_ptr()), None,
                           None) == _abi.CM_ERR_INVALID
    o = _abi.TwebOpts(N, L, 0.0, 0.0, 0.0, 1.0, 1)
    assert ctx.lib.cm_tweb_from_phi(ctx.h, C.byref(o), C.c_void_p(delta.data_ptr()), C.c_void_p(labels.data_ptr()),
                                    None, None) == _abi.CM_ERR_UNSUPPORTED
    with pytest.raises(cm.CosmoError):
        ctx.tweb_f64(delta.double(), L, 0.0, 0.0, want_phi=True, hessian="spectral")
