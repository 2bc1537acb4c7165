"""Packed measurement of uneven mocks (cm_measure_packed): one deposit launch
per MAS scheme over the union (mock, tile) work list and one T-web launch over
all mocks' items.  Every mock must get what cm_measure gives it alone."""
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


# (N, NL, L, scheme, ndel, T-web, interlace, weighted, seed); the last two
# mocks are not packable (interlaced; T-web at N = 96, not a multiple of 64)
# and run through cm_measure inside the same call
SPECS = [
    (64, 32, 200.0, "NGP", 1, False, False, False, 11),
    (128, 64, 500.0, "CIC", 1, True, False, False, 12),
    (192, 48, 700.0, "TSC", 1, True, False, False, 13),
    (256, 96, 1000.0, "TSC", 2, True, False, False, 14),
    (128, 48, 300.0, "PCS", 1, True, False, False, 15),
    (64, 32, 250.0, "TSC", 1, True, False, True, 16),
    (256, 64, 800.0, "NGP", 2, True, False, False, 17),
    (192, 96, 900.0, "CIC", 1, False, False, True, 18),
    (128, 64, 400.0, "TSC", 1, True, True, False, 19),
    (96, 48, 300.0, "CIC", 1, True, False, False, 20),
]


def _args(ctx, spec):
    import cosmicatlas_amd as cm
    N, NL, L, sch, ndel, tw, il, wt, sd = spec
    b = cm.binning(N, L, ndel=ndel)
    pos = ctx.zeldovich(NL, L, A=1e4, ns=-1.5, seed=sd)
    w = None
    if wt:
        g = torch.Generator(device="cuda").manual_seed(sd)
        w = torch.rand(pos.shape[0], device="cuda", generator=g, dtype=torch.float32) + 0.5
    kw = dict(do_tweb=tw, interlace=il)
    if tw:
        kw.update(R_smooth=2.0 * L / N, margin_tol=cm.TOL_ROUNDING)
    return dict(pos=pos, N=N, L=L, scheme=sch, nshells=b["nshells"], kmin=b["kmin"], dk_bin=b["dk"], weights=w,
                **kw)


def _check_same(p, a, spec):
    """Packed equals alone: the tile deposits, FFTs and T-web are the same
    kernels on the same data.  What may differ is the order of additions: the
    FP64 shell sums, and the float atomics that join the chunks of a tile
    holding more than 4096 records (clustered regions), which move P(k) by up
    to ~1e-8 of |P| + SN and can flip labels only inside the counted margin
    band."""
    N, L = spec[0], spec[2]
    assert np.array_equal(p["nmodes"], a["nmodes"])
    assert p["n_obj"] == pytest.approx(a["n_obj"], rel=1e-9)
    sn = L ** 3 / a["n_obj"]
    for l in ("pk0", "pk2", "pk4"):
        assert np.max(np.abs(p[l] - a[l]) / (np.abs(a["pk0"]) + sn)) < 1e-7, l
    if spec[5]:
        assert sum(p["tweb_counts"]) == N ** 3
        assert int((p["labels"] != a["labels"]).sum()) <= max(p["tweb_marginal"], a["tweb_marginal"])
        assert p["tweb_marginal"] == pytest.approx(a["tweb_marginal"], rel=1e-3, abs=2)
        assert p["tweb_tol"] == pytest.approx(a["tweb_tol"], rel=1e-5)


def test_packed_mixed_batch_equals_alone(ctx):
    args = [_args(ctx, s) for s in SPECS]
    packed = ctx.measure_packed(args)
    assert len(packed) == len(SPECS)
    for spec, a, p in zip(SPECS, args, packed):
        kw = {k: v for k, v in a.items() if k not in ("pos", "N", "L", "scheme", "nshells", "kmin", "dk_bin")}
        alone = ctx.measure(a["pos"], a["N"], a["L"], a["scheme"], a["nshells"], a["kmin"], a["dk_bin"], **kw)
        _check_same(p, alone, spec)


def test_packed_repeat_is_reproducible(ctx):
    """A second call reuses the per-mock buffers (and the groups' arenas) and
    gives the same results."""
    args = [_args(ctx, s) for s in SPECS[:6]]
    first = ctx.measure_packed(args)
    second = ctx.measure_packed(args)
    for spec, p, q in zip(SPECS, first, second):
        _check_same(q, p, spec)


def test_packed_empty_mock(ctx):
    """A mock without particles beside others: the same result as alone."""
    import cosmicatlas_amd as cm
    a = _args(ctx, SPECS[1])
    e = dict(_args(ctx, SPECS[2]))
    e["pos"] = e["pos"][:0]
    res = ctx.measure_packed([a, e])
    kw = {k: v for k, v in e.items() if k not in ("pos", "N", "L", "scheme", "nshells", "kmin", "dk_bin")}
    alone = ctx.measure(e["pos"], e["N"], e["L"], e["scheme"], e["nshells"], e["kmin"], e["dk_bin"], **kw)
    assert res[1]["n_obj"] == alone["n_obj"] == 0.0
    assert np.array_equal(res[1]["nmodes"], alone["nmodes"])
    np.testing.assert_array_equal(res[1]["pk0"], alone["pk0"])
    assert torch.equal(res[1]["labels"], alone["labels"])


def test_packed_errors_are_per_mock(ctx):
    """A mock that fails its checks fails alone; the others are measured."""
    import cosmicatlas_amd as cm
    good = _args(ctx, SPECS[1])
    bad = dict(_args(ctx, SPECS[0]), N=100)  # no FFT of 100 points
    with pytest.raises(cm.CosmoError) as e:
        ctx.measure_packed([good, bad])
    assert "packed mock 1" in str(e.value)
    assert ctx.measure_packed([]) == []
