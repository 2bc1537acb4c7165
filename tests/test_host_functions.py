"""The C++ host layer's free functions (reference names and argument
meanings) through tests/host_harness.cpp, against the oracle:
getDensity_{NGP,CIC,TSC} (Source/Miscelanious.cpp:65-147, 454-557, 693-815),
do_fftw_r2c / do_fftw_c2r (Source/NumericalMethods.cpp:1718-1753, 1818-1856),
PoissonSolver (Source/Miscelanious.cpp:1189-1274), EigenValuesTweb FD branch
(Source/Miscelanious.cpp:1668-1769) and get_nobjects (:1421-1426)."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "cosmicatlas_amd", "bin", "host_harness")


def _run(*args):
    r = subprocess.run([HARNESS, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


@pytest.mark.parametrize("scheme,code", [("NGP", 0), ("CIC", 1), ("TSC", 2)])
@pytest.mark.parametrize("weightmass", [False, True])
def test_getdensity_matches_oracle(tmp_path, orc, scheme, code, weightmass):
    N, L, n = 32, 120.0, 20000
    rng = np.random.default_rng(code + 10 * weightmass)
    pos = _f32(rng.uniform(0, L, (n, 3)))
    m = _f32(rng.uniform(0.5, 2.0, n))
    _f32(np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], m])).tofile(tmp_path / "p.bin")
    _run("getdensity", scheme, N, L, int(weightmass), tmp_path / "p.bin", tmp_path / "o.bin")
    got = np.fromfile(tmp_path / "o.bin", np.float32).reshape(N, N, N)
    ref = orc.mas(pos, N, L, code, weights=m if weightmass else None)
    if code == 0 and not weightmass:
        assert np.array_equal(got, ref)  # counts are exact
    else:
        assert rel_err(got, ref) < 6e-6


def test_getdensity_adds_onto_the_given_mesh(tmp_path, orc):
    """The reference adds every stencil onto delta (Miscelanious.cpp:454-557):
    a mesh of N^3 values passed in is accumulated onto."""
    N, L, n = 16, 50.0, 3000
    rng = np.random.default_rng(8)
    pos = _f32(rng.uniform(0, L, (n, 3)))
    _f32(np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], np.ones(n)])).tofile(tmp_path / "p.bin")
    init = _f32(rng.uniform(0, 1, (N, N, N)))
    init.tofile(tmp_path / "i.bin")
    _run("getdensity", "TSC", N, L, 0, tmp_path / "p.bin", tmp_path / "o.bin", tmp_path / "i.bin")
    got = np.fromfile(tmp_path / "o.bin", np.float32).reshape(N, N, N)
    assert rel_err(got, init + orc.mas(pos, N, L, orc.TSC)) < 6e-6


def test_fft_roundtrip_matches_oracle(tmp_path, orc):
    N = 24
    x = _f32(np.random.default_rng(3).standard_normal((N, N, N)))
    x.tofile(tmp_path / "x.bin")
    _run("r2c", N, tmp_path / "x.bin", tmp_path / "h.bin")
    h = np.fromfile(tmp_path / "h.bin", np.float32).view(np.complex64).reshape(N, N, N // 2 + 1)
    ref = orc.fft_r2c(x.astype(np.float64))
    assert rel_err(h, ref) < 2e-6 * np.log2(N)
    # c2r of the oracle spectrum equals the oracle c2r (divided by N^3, FFTW convention)
    ref.astype(np.complex64).tofile(tmp_path / "r.bin")
    _run("c2r", N, tmp_path / "r.bin", tmp_path / "y.bin")
    y = np.fromfile(tmp_path / "y.bin", np.float32).reshape(N, N, N)
    assert rel_err(y, orc.fft_c2r(ref, N)) < 1e-5
    assert rel_err(y, x) < 1e-5


def test_poisson_and_eigenvalues_match_oracle(tmp_path, orc):
    N, L = 32, 100.0
    rng = np.random.default_rng(9)
    k = np.fft.fftfreq(N)[:, None, None] ** 2 + np.fft.fftfreq(N)[None, :, None] ** 2 \
        + np.fft.rfftfreq(N)[None, None, :] ** 2
    delta = np.fft.irfftn(np.fft.rfftn(rng.standard_normal((N, N, N))) * np.exp(-40 * k), s=(N, N, N),
                          axes=(0, 1, 2))
    delta = _f32(delta / delta.std())
    delta.tofile(tmp_path / "d.bin")
    _run("poisson", N, L, tmp_path / "d.bin", tmp_path / "phi.bin")
    phi = np.fromfile(tmp_path / "phi.bin", np.float32).reshape(N, N, N)
    dk = orc.smooth_poisson(orc.fft_r2c(delta.astype(np.float64)), N, L, 0.0)
    assert rel_err(phi, orc.fft_c2r(dk, N)) < 1e-5
    _run("eigen", N, L, tmp_path / "phi.bin", tmp_path / "lam.bin")
    lam = np.fromfile(tmp_path / "lam.bin", np.float32).reshape(3, N, N, N)
    _, lam_ref, _ = orc.tweb_from_phi(phi.astype(np.float64), L, 0.0)
    assert np.max(np.abs(lam - lam_ref)) < 2e-5 * np.abs(lam_ref).max()


def test_get_nobjects(tmp_path):
    x = _f32(np.random.default_rng(2).uniform(0, 3, 1 << 20))
    x.tofile(tmp_path / "x.bin")
    _run("nobjects", tmp_path / "x.bin", tmp_path / "s.bin")
    s = np.fromfile(tmp_path / "s.bin", np.float32)[0]
    assert s == np.float32(x.astype(np.float64).sum())  # FP64 sum rounded once


def test_harness_reports_errors(tmp_path):
    r = subprocess.run([HARNESS, "r2c", "20", tmp_path / "none.bin", tmp_path / "o.bin"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot read" in r.stderr
