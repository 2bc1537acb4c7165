"""GPU generator (cm_zeldovich, cm_uniform) against the FP64 restatement."""
import numpy as np
import pytest
import torch

from oracle import zeldovich as z

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("NL,L,seed", [(16, 100.0, 1), (48, 300.0, 7), (64, 1000.0, 123456789012)])
def test_zeldovich_matches_restatement(ctx, NL, L, seed):
    g = ctx.zeldovich(NL, L, A=1e4, ns=-1.5, seed=seed).cpu().numpy()
    ref = z.zeldovich(NL, L, A=1e4, ns=-1.5, seed=seed, dtype=np.float64)
    diff = np.abs((g.astype(np.float64) - ref + L / 2) % L - L / 2)
    # float32 FFT of the displacement: errors far below a cell (L/NL)
    assert diff.max() < 1e-4 * (L / NL), diff.max()
    assert g.min() >= 0 and g.max() < L


def test_uniform_bit_exact(ctx):
    n, L = 100003, 777.0
    g = ctx.uniform(n, L, seed=42).cpu().numpy()
    assert np.array_equal(g, z.uniform(n, L, 42))
    assert ctx.uniform(0, L, seed=1).shape == (0, 3)
