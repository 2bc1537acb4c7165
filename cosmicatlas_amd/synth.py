"""Seeded synthetic inputs of the benchmark configurations (host side, numpy).

* ``uniform``: i.i.d. uniform positions in [0, L)^3.
* ``zeldovich``: a Gaussian random field with P(k) = A k^ns on an NL^3
  Lagrangian grid (fixed amplitude, seeded uniform phases: the
  create_GARFIELD_FIXED_AMP semantics of Source/NumericalMethods.cpp:2147-2189,
  without its racy RNG), displaced by Psi(k) = i k delta(k)/k^2 (so that
  div Psi = -delta), positions q + Psi(q) wrapped into [0, L), float32.

The amplitude follows the estimator normalisation P = V |delta_k|^2 / N^6
(Source/FftwFunctions.cpp:159-174), so <P_hat(k)> = P(k).
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, L: float, seed: int, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, L, (n, 3)).astype(dtype)


def _kgrid(NL: int, L: float):
    kf = 2.0 * np.pi / L
    i = np.arange(NL)
    k1 = kf * np.where(i <= NL // 2, i, i - NL).astype(np.float64)
    kz = kf * np.arange(NL // 2 + 1, dtype=np.float64)
    return k1, kz


def gaussian_field_k(NL: int, L: float, A: float, ns: float, seed: int) -> np.ndarray:
    """Half spectrum delta_k (unnormalised forward convention)."""
    rng = np.random.default_rng(seed)
    k1, kz = _kgrid(NL, L)
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    kmod = np.sqrt(k2)
    with np.errstate(divide="ignore"):
        P = np.where(kmod > 0, A * kmod ** ns, 0.0)
    V = L ** 3
    amp = np.sqrt(P * NL ** 6 / V)
    theta = rng.uniform(0.0, 2.0 * np.pi, amp.shape)
    dk = amp * np.exp(1j * theta)
    dk[0, 0, 0] = 0.0
    # self-conjugate modes (all indices 0 or N/2) are kept real
    h = NL // 2
    for a in (0, h):
        for b in (0, h):
            for c in (0, h):
                dk[a, b, c] = dk[a, b, c].real
    return dk


def zeldovich(NL: int, L: float, A: float = 1e4, ns: float = -1.5, seed: int = 1,
              dtype=np.float32) -> np.ndarray:
    """(NL^3, 3) particle positions in [0, L)."""
    dk = gaussian_field_k(NL, L, A, ns, seed)
    k1, kz = _kgrid(NL, L)
    h = NL // 2
    # derivative wavenumbers with the Nyquist component set to zero
    kd = k1.copy(); kd[h] = 0.0
    kdz = kz.copy(); kdz[-1] = 0.0
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    inv = np.zeros_like(k2)
    inv[k2 > 0] = 1.0 / k2[k2 > 0]
    d = L / NL
    q = (np.arange(NL, dtype=np.float64) + 0.5) * d
    pos = np.empty((NL, NL, NL, 3), np.float64)
    comps = (kd[:, None, None], kd[None, :, None], kdz[None, None, :])
    grids = np.meshgrid(q, q, q, indexing="ij")
    for a in range(3):
        psi = np.fft.irfftn(1j * comps[a] * dk * inv, s=(NL, NL, NL))
        pos[..., a] = grids[a] + psi
    pos = np.mod(pos, L).reshape(-1, 3)
    out = pos.astype(dtype)
    out[out >= L] = 0.0  # float rounding of values just below L
    return out
