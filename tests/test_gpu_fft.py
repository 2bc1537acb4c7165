"""GPU r2c/c2r against the FP64 oracle FFT (same inputs)."""
import numpy as np
import pytest
import torch

from gpu_util import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import cosmicatlas_amd as cm
    c = cm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("N", [8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384])
def test_r2c_f32_matches_oracle(ctx, orc, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N)).astype(np.float32)
    out = ctx.fft_r2c(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = orc.fft_r2c(x.astype(np.float64))
    # float32 FFT: error relative to the largest mode, grows ~ log N
    assert rel_err(out, ref) < 2e-6 * np.log2(N), rel_err(out, ref)


@pytest.mark.parametrize("N", [8, 24, 64, 96])
def test_r2c_f64_matches_oracle(ctx, orc, N):
    rng = np.random.default_rng(100 + N)
    x = rng.standard_normal((N, N, N))
    out = ctx.fft_r2c(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = orc.fft_r2c(x)
    assert rel_err(out, ref) < 1e-13


@pytest.mark.parametrize("N,dt", [(16, torch.float32), (48, torch.float32), (128, torch.float32),
                                  (256, torch.float32), (384, torch.float32),
                                  (32, torch.float64), (96, torch.float64)])
def test_c2r_roundtrip_and_oracle(ctx, orc, N, dt):
    rng = np.random.default_rng(7 + N)
    x = rng.standard_normal((N, N, N))
    xt = torch.from_numpy(x).to(dt).cuda()
    h = ctx.fft_r2c(xt)
    href = orc.fft_r2c(xt.cpu().double().numpy())
    back = ctx.fft_c2r(h, N).cpu().numpy()
    tol = 1e-5 if dt == torch.float32 else 1e-12
    assert rel_err(back, xt.cpu().numpy()) < tol
    # c2r of the oracle spectrum equals the oracle c2r
    g = ctx.fft_c2r(torch.from_numpy(href).to(torch.complex64 if dt == torch.float32 else torch.complex128).cuda(), N)
    assert rel_err(g.cpu().numpy(), orc.fft_c2r(href, N)) < tol


def test_subtract_on_load_and_input_preserved(ctx, orc):
    N = 32
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((N, N, N)) + 5.0).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    h = ctx.fft_r2c(xt, subtract=5.0).cpu().numpy()
    ref = orc.fft_r2c(x.astype(np.float64) - 5.0)
    assert rel_err(h, ref) < 1e-5
    assert torch.equal(xt.cpu(), torch.from_numpy(x))


def test_unsupported_size_raises(ctx):
    import cosmicatlas_amd as cm
    with pytest.raises(cm.CosmoError):
        ctx.fft_r2c(torch.zeros((20, 20, 20), device="cuda"))
