"""hipGraph replay of the fused measurement (cm_set_graphs, cm_graph_counts)
and the launch-shape autotuner (cm_autotune, cm_set_tune, cm_get_tune).

Graph replays are checked one by one against an eager context on the same
inputs, with eager work of the same context between replays, after the input
buffer is refilled in place, and after a work buffer of the context has been
reallocated (the graph must then be captured again)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, L, NL = 64, 300.0, 48
A = 100.0  # weak clustering: no tile exceeds one deposit workgroup, so the
           # deposit (integer LDS sums) and every later kernel are deterministic


@pytest.fixture(scope="module")
def cm():
    import cosmicatlas_amd as cm
    return cm


def _measure(ctx, pos, labels, N=N, L=L, scheme="TSC", do_tweb=True):
    import cosmicatlas_amd as cm
    b = cm.binning(N, L, ndel=1)
    return ctx.measure(pos, N, L, scheme, b["nshells"], b["kmin"], b["dk"], do_tweb=do_tweb,
                       R_smooth=2.0 * L / N if do_tweb else 0.0, margin_tol=1e-6, labels=labels)


def _same(a, b, la=None, lb=None):
    assert np.array_equal(a["nmodes"], b["nmodes"])
    sn = L ** 3 / a["n_obj"]
    assert a["n_obj"] == b["n_obj"]
    for k in ("pk0", "pk2", "pk4"):
        assert np.max(np.abs(a[k] - b[k]) / (np.abs(a[k]) + sn)) < 1e-9, k
    if la is not None:
        assert torch.equal(la, lb)
        assert a["tweb_counts"] == b["tweb_counts"]


def test_graph_replays_match_eager_with_work_between(cm):
    ce, cg = cm.Context(0), cm.Context(0)
    ce.set_graphs(False)
    try:
        pos = cg.zeldovich(NL, L, A=A, ns=-1.5, seed=3)
        le = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
        lg = torch.empty_like(le)
        ref = _measure(ce, pos, le)
        other = cg.zeldovich(32, L, A=A, seed=9)
        for it in range(5):
            out = _measure(cg, pos, lg)
            _same(ref, out, le, lg)
            # eager work of the same context between replays: another deposit and
            # a smaller measurement (no buffer grows)
            cg.mesh_assign(other, 32, L, "CIC")
            _measure(cg, other, None, N=32, do_tweb=False)
            torch.cuda.synchronize()
        cap, launched = cg.graph_counts()
        assert cap >= 1 and launched >= 3
        assert ce.graph_counts() == (0, 0)
    finally:
        ce.close()
        cg.close()


def test_graph_reads_inputs_refilled_in_place(cm):
    ce, cg = cm.Context(0), cm.Context(0)
    ce.set_graphs(False)
    try:
        pos = cg.zeldovich(NL, L, A=A, seed=1)
        lg = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
        le = torch.empty_like(lg)
        for seed in (1, 1, 2, 5, 2):  # eager, capture, then replays on new contents
            cg.zeldovich(NL, L, A=A, seed=seed, out=pos)
            out = _measure(cg, pos, lg)
            ref = _measure(ce, pos, le)
            _same(ref, out, le, lg)
        assert cg.graph_counts()[1] >= 3
    finally:
        ce.close()
        cg.close()


def test_graph_recaptured_after_buffer_reallocation(cm):
    ce, cg = cm.Context(0), cm.Context(0)
    ce.set_graphs(False)
    try:
        pos = cg.zeldovich(NL, L, A=A, seed=4)
        lg = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
        le = torch.empty_like(lg)
        ref = _measure(ce, pos, le)
        for _ in range(3):
            _same(ref, _measure(cg, pos, lg), le, lg)
        cap0 = cg.graph_counts()[0]
        # a larger mesh grows the context's work buffers (the old addresses are freed)
        big = cg.zeldovich(64, 600.0, A=A, seed=8)
        _measure(cg, big, torch.empty((128,) * 3, dtype=torch.int8, device="cuda"), N=128, L=600.0)
        for _ in range(2):
            _same(ref, _measure(cg, pos, lg), le, lg)
        assert cg.graph_counts()[0] == cap0 + 1
    finally:
        ce.close()
        cg.close()


def test_autotune_shapes_and_results(cm):
    ctx = cm.Context(0)
    try:
        pos = ctx.zeldovich(NL, L, seed=6)
        b = cm.binning(N, L, ndel=1)
        lab0 = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
        ctx.set_tune(N, None)
        ref = _measure(ctx, pos, lab0)
        res = ctx.autotune(pos, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], do_tweb=True,
                           R_smooth=2.0 * L / N, margin_tol=1e-6)
        t = res["tune"]
        assert t["xnb"] in (0, 8, 16) and t["xe"] in (0, 8, 16) and t["tz"] in (0, 16, 32, 64)
        assert t["tseg"] in (0, 1, 2, 4, 8) and t["dep_wg"] in (0, 4, 16)
        assert 0 < res["ms_after"] <= res["ms_before"]
        assert ctx.get_tune(N) == t
        lab1 = torch.empty_like(lab0)
        out = _measure(ctx, pos, lab1)
        # other x-pass shapes change the float32 FFT's rounding, not the estimator
        assert np.array_equal(out["nmodes"], ref["nmodes"])
        sn = L ** 3 / ref["n_obj"]
        assert np.max(np.abs(out["pk0"] - ref["pk0"]) / (np.abs(ref["pk0"]) + sn)) < 1e-5
        assert (lab1 != lab0).float().mean().item() < 1e-3
        # a fixed shape is honoured and can be cleared
        ctx.set_tune(N, {"xnb": 8, "tz": 32, "tseg": 2, "dep_wg": 4})
        assert ctx.get_tune(N) == {"xnb": 8, "xe": 0, "tz": 32, "tseg": 2, "dep_wg": 4}
        out2 = _measure(ctx, pos, lab1)
        assert np.max(np.abs(out2["pk0"] - ref["pk0"]) / (np.abs(ref["pk0"]) + sn)) < 1e-5
        ctx.set_tune(N, None)
        assert ctx.get_tune(N) == {"xnb": 0, "xe": 0, "tz": 0, "tseg": 0, "dep_wg": 0}
    finally:
        ctx.set_tune(N, None)
        ctx.close()


def test_sequences_with_host_sync_are_not_captured(cm):
    """The x-slab group binning of meshes with more tiles than one histogram
    holds synchronises with the host inside the sequence; such a call must run
    eagerly every time (no capture attempt) and give the eager results."""
    import os
    old = os.environ.get("CM_TUNE")
    os.environ["CM_TUNE"] = "mas_hist=32"  # N = 64: 64 tiles, two x-slab groups
    try:
        cg = cm.Context(0)
    finally:
        if old is None:
            del os.environ["CM_TUNE"]
        else:
            os.environ["CM_TUNE"] = old
    ce = cm.Context(0)
    ce.set_graphs(False)
    try:
        pos = cg.zeldovich(NL, L, A=A, seed=12)
        lg = torch.empty((N, N, N), dtype=torch.int8, device="cuda")
        le = torch.empty_like(lg)
        ref = _measure(ce, pos, le)
        for _ in range(4):
            _same(ref, _measure(cg, pos, lg), le, lg)
        assert cg.graph_counts() == (0, 0)
    finally:
        cg.close()
        ce.close()
