"""GPU mass assignment against the oracle: NGP counts bit-exact, CIC/TSC/PCS
to float32 round-off, plain and tiled deposits, wrap, weights, interlacing."""
import numpy as np
import pytest
import torch

from gpu_util import clustered_pos, rel_err, uniform_pos

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("algo", ["plain", "tiled"])
@pytest.mark.parametrize("N,n", [(32, 50000), (128, 400000)])
def test_ngp_counts_bit_exact(ctx, orc, algo, N, n):
    L = 500.0
    pos = uniform_pos(n, L, N + n)
    mesh = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, "NGP", algo=algo).cpu().numpy()
    ref = orc.mas(pos, N, L, orc.NGP)
    assert np.array_equal(mesh.astype(np.float64), ref)
    assert mesh.sum() == n


def test_ngp_f64_positions_and_mesh(ctx, orc):
    N, L, n = 128, 500.0, 200000
    pos = uniform_pos(n, L, 1, np.float64)
    mesh = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, "NGP", mesh_dtype=torch.float64).cpu().numpy()
    assert np.array_equal(mesh, orc.mas(pos, N, L, orc.NGP))


@pytest.mark.parametrize("scheme", ["CIC", "TSC", "PCS"])
@pytest.mark.parametrize("algo", ["plain", "tiled"])
def test_cic_tsc_pcs_match_oracle(ctx, orc, scheme, algo):
    N, L, n = 64, 300.0, 300000
    pos = uniform_pos(n, L, 5)
    mesh = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, scheme, algo=algo).cpu().numpy()
    ref = orc.mas(pos, N, L, {"CIC": orc.CIC, "TSC": orc.TSC, "PCS": orc.PCS}[scheme])
    # tiled, unit weights: 32-bit fixed point, quantum 2^-20 for the 4096-record
    # chunks of this dense mesh (half a quantum per add, ~30 adds per cell)
    assert rel_err(mesh, ref) < (2e-6 if algo == "plain" else 6e-6)
    assert abs(mesh.astype(np.float64).sum() - n) < 1e-5 * n


def test_c1_cic_f64_mesh_matches_oracle(ctx, orc):
    """Plain C1 deposit: 200,000 float64 positions in L=500, CIC on 128^3 float64."""
    N, L, n = 128, 500.0, 200000
    pos = np.random.default_rng(1).uniform(0, L, (n, 3))
    mesh = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, "CIC", mesh_dtype=torch.float64).cpu().numpy()
    ref = orc.mas(pos, N, L, orc.CIC)
    assert np.max(np.abs(mesh - ref)) < 1e-12


@pytest.mark.parametrize("scheme", ["CIC", "TSC"])
def test_lattice_order_blocked_walk(ctx, orc, scheme):
    """128^3 particles in lattice order (rows of 128 along z, displaced by up to
    3 cells): the record scatter walks each workgroup's chunk in blocks of 32
    rows x 32 columns; the mesh must equal the oracle's as for any order."""
    NL, N, L = 128, 128, 500.0
    rng = np.random.default_rng(11)
    g = (np.arange(NL) + 0.5) * (L / NL)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = ((lat + rng.uniform(-3, 3, lat.shape) * (L / N)) % L).astype(np.float32)
    sch = {"CIC": orc.CIC, "TSC": orc.TSC}[scheme]
    mesh = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, scheme, algo="tiled").cpu().numpy()
    assert rel_err(mesh, orc.mas(pos, N, L, sch)) < 6e-6
    # the same particles in random order (the walk then has no lattice rows):
    # tiles of more than 4096 records are deposited in chunks, each rounded to
    # its own fixed-point quantum, and which records share a chunk depends on
    # the order, so the two meshes agree to the quantisation bound above
    perm = rng.permutation(len(pos))
    mp = ctx.mesh_assign(torch.from_numpy(pos[perm]).cuda(), N, L, scheme, algo="tiled").cpu().numpy()
    assert rel_err(mp, mesh) < 6e-6


def test_tiled_equals_plain_on_clustered_input(ctx):
    """Clumps put >4096 particles in a tile: chunked tiles with global atomics."""
    N, L, n = 64, 100.0, 600000
    pos = torch.from_numpy(clustered_pos(n, L, 9)).cuda()
    a = ctx.mesh_assign(pos, N, L, "TSC", algo="plain").cpu().numpy()
    b = ctx.mesh_assign(pos, N, L, "TSC", algo="tiled").cpu().numpy()
    assert rel_err(b, a) < 1e-5
    assert a.max() > 4096 * 0.05


def test_weights_soa_interlace_and_accumulate(ctx, orc):
    N, L, n = 32, 64.0, 20000
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, L, (n, 3)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    posd = torch.from_numpy(pos).cuda()
    wd = torch.from_numpy(w).cuda()
    for algo in ("plain", "tiled"):
        m = ctx.mesh_assign(posd, N, L, "CIC", weights=wd, algo=algo).cpu().numpy()
        assert rel_err(m, orc.mas(pos, N, L, orc.CIC, weights=w)) < 2e-6
        soa = ctx.mesh_assign(torch.from_numpy(pos.T.copy()).cuda(), N, L, "CIC", weights=wd, soa=True,
                              algo=algo).cpu().numpy()
        assert rel_err(soa, m) < 1e-6
        # unit weights on the tiled path: 32-bit fixed point, quantum 2^-20 at
        # the ~3500-record tiles of this 32^3 mesh
        tol = 2e-6 if algo == "plain" else 6e-6
        il = ctx.mesh_assign(posd, N, L, "TSC", interlace_shift=0.5, algo=algo).cpu().numpy()
        assert rel_err(il, orc.mas(pos, N, L, orc.TSC, interlace_shift=0.5)) < tol
        acc = ctx.mesh_assign(posd, N, L, "TSC", algo=algo)
        ctx.mesh_assign(posd, N, L, "TSC", algo=algo, mesh=acc, accumulate=True)
        assert rel_err(acc.cpu().numpy(), 2 * orc.mas(pos, N, L, orc.TSC)) < tol


def test_wrap_images_and_edges(ctx, orc):
    """Periodic images, coordinates exactly 0 and L, tiny negatives (float32 wrap)."""
    N, L = 16, 100.0
    pos = np.array([[0.0, 0.0, 0.0], [100.0, 50.0, 25.0], [-1e-6, 3.0, 99.99999], [250.0, -75.0, 1000.0],
                    [99.9999, 99.9999, 99.9999], [6.25, 12.5, 18.75]], np.float32)
    for scheme in ("NGP", "CIC", "TSC", "PCS"):
        for algo in ("plain", "tiled"):
            m = ctx.mesh_assign(torch.from_numpy(pos).cuda(), N, L, scheme, algo=algo).cpu().numpy()
            ref = orc.mas(pos, N, L, {"NGP": 0, "CIC": 1, "TSC": 2, "PCS": 3}[scheme])
            assert np.max(np.abs(m - ref)) < 1e-6, (scheme, algo)


def test_empty_input_and_nonfinite_dropped(ctx):
    N, L = 16, 10.0
    empty = torch.zeros((0, 3), device="cuda")
    for algo in ("plain", "tiled"):
        m = ctx.mesh_assign(empty, N, L, "TSC", algo=algo)
        assert float(m.abs().sum()) == 0.0
    bad = torch.tensor([[1.0, float("nan"), 2.0], [float("inf"), 1.0, 1.0], [1.0, 1.0, 1.0]], device="cuda")
    for algo in ("plain", "tiled"):
        m = ctx.mesh_assign(bad, N, L, "NGP", algo=algo)
        assert float(m.sum()) == 1.0


def test_mesh_sum_fp64(ctx):
    x = torch.full((64, 64, 64), 0.1, dtype=torch.float32, device="cuda")
    s = ctx.mesh_sum(x)
    assert abs(s - 0.1 * 64 ** 3 * (np.float32(0.1) / 0.1)) < 1e-6


@pytest.mark.parametrize("N,hist,scheme", [(64, 32, "TSC"), (64, 48, "CIC"), (128, 128, "TSC"), (128, 200, "NGP"),
                                           (96, 72, "PCS")])
def test_xslab_group_binning_matches_oracle(orc, N, hist, scheme):
    """Meshes whose tiles do not fit one workgroup's histogram are binned in
    x-slab groups first.  A small histogram budget (CM_TUNE mas_hist) forces
    that path on small meshes: several groups, copies across group edges and
    the periodic wrap, checked against the FP64 oracle, with weights, SoA and
    the interlace shift, and NGP counts bit-exact."""
    import os
    import cosmicatlas_amd as cm
    old = os.environ.get("CM_TUNE")
    os.environ["CM_TUNE"] = f"mas_hist={hist}"
    try:
        c = cm.Context(0)
    finally:
        if old is None:
            del os.environ["CM_TUNE"]
        else:
            os.environ["CM_TUNE"] = old
    try:
        L, n = 300.0, 150000
        rng = np.random.default_rng(N + hist)
        pos = rng.uniform(-0.2 * L, 1.2 * L, (n, 3)).astype(np.float32)  # images beyond the box too
        w = rng.uniform(0.5, 1.5, n).astype(np.float32)
        sch = {"NGP": orc.NGP, "CIC": orc.CIC, "TSC": orc.TSC, "PCS": orc.PCS}[scheme]
        posd, wd = torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda()
        m = c.mesh_assign(posd, N, L, scheme, algo="tiled").cpu().numpy()
        ref = orc.mas(pos, N, L, sch)
        if scheme == "NGP":
            assert np.array_equal(m.astype(np.float64), ref)
        else:
            assert rel_err(m, ref) < 6e-6
        mw = c.mesh_assign(posd, N, L, scheme, weights=wd, algo="tiled").cpu().numpy()
        assert rel_err(mw, orc.mas(pos, N, L, sch, weights=w)) < 2e-6
        soa = c.mesh_assign(torch.from_numpy(pos.T.copy()).cuda(), N, L, scheme, soa=True, algo="tiled").cpu().numpy()
        assert np.array_equal(soa, m)
        il = c.mesh_assign(posd, N, L, scheme, interlace_shift=0.5, algo="tiled").cpu().numpy()
        assert rel_err(il, orc.mas(pos, N, L, sch, interlace_shift=0.5)) < 6e-6
    finally:
        c.close()
