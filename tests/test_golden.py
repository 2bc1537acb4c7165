"""The oracle and the GPU path against the committed golden fixtures of
tests/golden/ (written by tests/golden/make_golden.py from numpy and closed
forms): FFT values, eigenvalues and the MAS stencils of one particle."""
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEMES = {"NGP": 0, "CIC": 1, "TSC": 2, "PCS": 3}


def _stencils():
    with open(os.path.join(HERE, "stencils.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("n", [8, 12])
def test_oracle_fft_golden(orc, n):
    g = np.load(os.path.join(HERE, f"fft_{n}.npz"))
    out = orc.fft_r2c(g["x"])
    assert np.max(np.abs(out - g["k"])) < 1e-12 * np.max(np.abs(g["k"]))


def test_oracle_eigen_golden(orc):
    g = np.load(os.path.join(HERE, "eig.npz"))
    for m, ev in zip(g["m"], g["ev"]):
        got = np.sort(orc.eigen3(m.reshape(9)))
        assert np.max(np.abs(got - ev)) < 1e-12 * max(1.0, np.max(np.abs(ev)))


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_oracle_stencil_golden(orc, scheme):
    st = _stencils()
    for s in st["offsets"]:
        w = st["weights"][scheme][str(s)]
        for j, wj in zip(st["nodes"], w):
            assert orc.mas_kernel(SCHEMES[scheme], j - s) == pytest.approx(wj, abs=1e-15)


def _one_particle_mesh(N, L, scheme, c, s):
    """Expected mesh of one particle at (c_a + 1/2 + s_a) cells: the outer
    product of the golden 1D weights at its nodes (periodic)."""
    st = _stencils()
    m = np.zeros((N, N, N))
    ws = [st["weights"][scheme][str(s_a)] for s_a in s]
    for i, wi in zip(st["nodes"], ws[0]):
        for j, wj in zip(st["nodes"], ws[1]):
            for k, wk in zip(st["nodes"], ws[2]):
                m[(c[0] + i) % N, (c[1] + j) % N, (c[2] + k) % N] += wi * wj * wk
    return m


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_oracle_single_particle_golden(orc, scheme):
    N, L = 16, 32.0
    d = L / N
    c, s = (3, 15, 0), (0.125, -0.25, 0.3)
    pos = np.array([[(c[a] + 0.5 + s[a]) * d for a in range(3)]])
    m = orc.mas(pos, N, L, SCHEMES[scheme])
    np.testing.assert_allclose(m, _one_particle_mesh(N, L, scheme, c, s), rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_fft_golden():
    import torch
    import cosmicatlas_amd as cm
    g = np.load(os.path.join(HERE, "fft_8.npz"))
    with cm.Context(0) as ctx:
        x64 = torch.from_numpy(g["x"]).cuda()
        k64 = ctx.fft_r2c(x64).cpu().numpy()
        k32 = ctx.fft_r2c(x64.float()).cpu().numpy()
    scale = np.max(np.abs(g["k"]))
    assert np.max(np.abs(k64 - g["k"])) < 1e-12 * scale
    assert np.max(np.abs(k32 - g["k"])) < 2e-6 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("algo", ["plain", "tiled"])
def test_gpu_single_particle_golden(scheme, algo):
    import torch
    import cosmicatlas_amd as cm
    N, L = 16, 32.0
    d = L / N
    c, s = (3, 15, 0), (0.125, -0.25, 0.3)
    pos = torch.tensor([[(c[a] + 0.5 + s[a]) * d for a in range(3)]], dtype=torch.float32).cuda()
    with cm.Context(0) as ctx:
        m = ctx.mesh_assign(pos, N, L, scheme, algo=algo).cpu().numpy()
    np.testing.assert_allclose(m, _one_particle_mesh(N, L, scheme, c, s), rtol=0, atol=2e-7)
