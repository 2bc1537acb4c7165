"""The single-mock workload at its full size (256^3 Zel'dovich particles, TSC onto
512^3, 128 shells, T-web) and the grouped binning at 768^3, checked through
properties that need no CPU pass at these sizes: float32 against a float64 mesh,
FFT and shells on the GPU (the P(k) tolerance 1e-4), equal mode counts, the
particle count, invariance under a translation by whole cells, and agreement of
the fused and plain paths; the T-web labels against an FP64 T-web of the same
particles, every differing label inside the counted margin band."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, L, NL = 512, 1000.0, 256


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


def _measure(ctx, pos, N_, L_, ndel, fused=True, tweb=True):
    import cosmicatlas_amd as cm
    b = cm.binning(N_, L_, ndel=ndel)
    kw = dict(do_tweb=True, R_smooth=2.0, lambda_th=0.0, margin_tol=cm.TOL_ROUNDING) if tweb else {}
    return ctx.measure(pos, N_, L_, "TSC", b["nshells"], b["kmin"], b["dk"], fused=fused, **kw), b


def _fp64_pk(ctx, pos, N_, L_, b):
    mesh64 = ctx.mesh_assign(pos, N_, L_, "TSC", mesh_dtype=torch.float64)
    nobj = ctx.mesh_sum(mesh64)
    dk64 = ctx.fft_r2c(mesh64, subtract=nobj / N_ ** 3)
    del mesh64
    V = L_ ** 3
    p0, _, _, nm = ctx.pk_shells(dk64, N_, L_, b["nshells"], b["kmin"], b["dk"], "TSC", V / nobj, nobj ** 2 / V)
    raw = ctx.pk_shells(dk64, N_, L_, b["nshells"], b["kmin"], b["dk"], "TSC", 0.0, nobj ** 2 / V)[0]
    del dk64
    torch.cuda.empty_cache()
    return p0, nm, raw, nobj


@pytest.fixture(scope="module")
def mock(ctx):
    pos = ctx.zeldovich(NL, L, A=1e4, ns=-1.5, seed=1)
    out, b = _measure(ctx, pos, N, L, 2)
    return pos, out, b


def test_fullsize_pk_against_fp64(ctx, mock):
    pos, out, b = mock
    p0, nm, raw, nobj = _fp64_pk(ctx, pos, N, L, b)
    assert np.array_equal(out["nmodes"], nm)
    assert out["n_obj"] == pytest.approx(NL ** 3, rel=1e-6) and nobj == pytest.approx(NL ** 3, rel=1e-12)
    ok = nm > 0
    # relative to the power before shot-noise subtraction (what the FFT computes)
    assert np.max(np.abs(out["pk0"][ok] - p0[ok]) / np.abs(raw[ok])) < 1e-4


def test_fullsize_tweb_labels_against_fp64(ctx, mock):
    """Labels equal the FP64 T-web (FP64 deposit, FFTs, FD Hessian and
    eigenvalues) except for cells whose eigenvalue margin lies below the
    rounding tolerance K(N) u ||phi||_rms / d^2, and those cells are counted.
    The float32 T-web of cm_tweb_from_phi returns its eigenvalues, so there the
    claim is checked cell by cell: every differing label has a float32 margin
    below tol, the counted cells are those with such a margin, and tol bounds
    the eigenvalue error everywhere.  The fused measurement keeps no
    eigenvalues: every label it gets wrong must have an FP64 margin below tol
    (an eigenvalue on the other side of lambda_th), and there are no more such
    labels than counted cells."""
    import cosmicatlas_amd as cm
    pos, out, b = mock
    n = pos.shape[0]
    m64 = ctx.mesh_assign(pos, N, L, "TSC", algo="plain", mesh_dtype=torch.float64)
    ctx.mesh_affine(m64, N ** 3 / n, -1.0)
    lab64, lam64, _, st64 = ctx.tweb_f64(m64, L, 2.0, 0.0, margin_tol=cm.TOL_ROUNDING, want_lambdas=True)
    del m64
    assert sum(st64["counts"]) == N ** 3
    marg64 = lam64.abs().min(dim=0).values
    m32 = ctx.mesh_assign(pos, N, L, "TSC")
    ctx.mesh_affine(m32, N ** 3 / n, -1.0)
    phi32 = ctx.poisson(m32, L, R=2.0)
    del m32
    lab32, lam32, st32 = ctx.tweb_from_phi(phi32, L, 0.0, margin_tol=cm.TOL_ROUNDING, want_lambdas=True)
    del phi32
    tol = st32["margin_tol"]
    # the tolerance is the one the fused path applied (same ||phi||_rms to float rounding)
    assert out["tweb_tol"] == pytest.approx(tol, rel=1e-4)
    # the eigenvalue error itself lies below tol (the claim the tolerance makes;
    # measured 0.38 tol at this size)
    err = max(float((lam32[j].double() - lam64[j]).abs().max()) for j in range(3))
    del lam64
    assert err < tol
    # about 3% of the cells are counted at this size; far fewer labels differ
    assert 0 < st32["n_marginal"] < N ** 3 // 20
    marg32 = lam32.abs().min(dim=0).values
    del lam32
    in_band = marg32 < tol
    edge = int(((marg32 - tol).abs() < 1e-5 * tol).sum())
    assert abs(int(in_band.sum()) - st32["n_marginal"]) <= edge
    differ = lab32 != lab64
    assert int(differ.sum()) > 0
    assert bool(in_band[differ].all()), "a float32 label differs from FP64 outside the counted band"
    # fused path
    differ_f = out["labels"] != lab64
    nf = int(differ_f.sum())
    assert nf <= out["tweb_marginal"]
    assert bool((marg64[differ_f] < tol).all()), "a fused label differs from FP64 with an FP64 margin above tol"
    assert abs(out["tweb_marginal"] - st32["n_marginal"]) < 0.01 * st32["n_marginal"]


def test_fullsize_fused_equals_plain(ctx, mock):
    pos, out, b = mock
    plain, _ = _measure(ctx, pos, N, L, 2, fused=False)
    sn = out["shot_noise"]
    assert np.array_equal(out["nmodes"], plain["nmodes"])
    assert np.max(np.abs(out["pk0"] - plain["pk0"]) / (np.abs(plain["pk0"]) + sn)) < 1e-5
    # the plain path's eigenvalues are float32; the fused labels are exact up to the margin
    agree = float((out["labels"] == plain["labels"]).float().mean())
    assert agree > 0.999
    assert sum(out["tweb_counts"]) == N ** 3


def test_fullsize_translation_by_whole_cells(ctx, mock):
    """Moving every particle by (3, 5, 7) cells (d = L/N = 125/64 is exact in
    float32) rolls the mesh: the power spectrum is unchanged and the labels roll
    with it."""
    pos, out, b = mock
    d = L / N
    shift = torch.tensor([3 * d, 5 * d, 7 * d], dtype=torch.float32, device=pos.device)
    moved = torch.remainder(pos + shift, L)
    out2, _ = _measure(ctx, moved, N, L, 2)
    assert np.array_equal(out["nmodes"], out2["nmodes"])
    assert np.max(np.abs(out2["pk0"] - out["pk0"]) / (np.abs(out["pk0"]) + out["shot_noise"])) < 1e-5
    # phi is float32: the FFTs' rounding (~1e-7 of its norm, carried by the largest
    # scales) differs between the two meshes, and the second differences of the
    # Hessian amplify it by up to (k_Nyquist / k_f)^2; labels of cells whose
    # eigenvalue lies that close to lambda_th may differ (measured: 6e-4 of cells)
    rolled = torch.roll(out["labels"], shifts=(3, 5, 7), dims=(0, 1, 2))
    assert float((rolled == out2["labels"]).float().mean()) > 0.999


def test_grouped_binning_768_against_fp64(ctx):
    """768^3 (tiles beyond one workgroup's LDS histogram: x-slab groups) with the
    same particles plus 20% uniform randoms appended, against FP64."""
    N2, L2 = 768, 1500.0
    zel = ctx.zeldovich(NL, L2, A=1e4, ns=-1.5, seed=2)
    ran = ctx.uniform(NL ** 3 // 5, L2, seed=3)
    pos = torch.cat([zel, ran])
    out, b = _measure(ctx, pos, N2, L2, 6, tweb=False)
    p0, nm, raw, nobj = _fp64_pk(ctx, pos, N2, L2, b)
    assert np.array_equal(out["nmodes"], nm)
    assert out["n_obj"] == pytest.approx(pos.shape[0], rel=1e-6)
    ok = nm > 0
    assert np.max(np.abs(out["pk0"][ok] - p0[ok]) / np.abs(raw[ok])) < 1e-4
