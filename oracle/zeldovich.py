"""FP64 numpy restatement of the synthetic-mock generator (Gaussian random
field with fixed amplitude and seeded phases, Zel'dovich displacement).

TEST INFRASTRUCTURE ONLY: the checker of the GPU generator (cm_zeldovich).

Reference semantics: create_GARFIELD_FIXED_AMP (Source/NumericalMethods.cpp:
2147-2189) draws |delta_k| = sqrt(P(k) N^6 / V) with a uniform random phase on
the r2c half spectrum; the reference fills the phases from a shared GSL
generator inside an OpenMP loop (a race), so its streams are not reproducible.
Here the phase of every mode comes from a counter-based generator,
Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC11), keyed by the seed and counted by the mode index, so the CPU
and the GPU draw the same numbers.  The kz = 0 and kz = N/2 planes are made
Hermitian (the mode (-i, -j) takes the conjugate of (i, j)); self-conjugate
modes keep amp*cos(theta).  The displacement follows the Zel'dovich
approximation with FFTW signs (LPT::get_displacement, Source/LPT.cpp:285):
Psi(k) = i k delta(k) / k^2 with the Nyquist component of k set to zero, so
div Psi = -delta; particles start at Lagrangian cell centres (i + 1/2) L/NL.
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32 with 10 rounds on uint32 arrays; returns the four output words."""
    c0 = np.asarray(c0, np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, np.uint32).astype(np.uint64)
    k0 = np.uint64(np.uint32(k0))
    k1 = np.uint64(np.uint32(k1))
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(W0)) & MASK32
            k1 = (k1 + np.uint64(W1)) & MASK32
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK32, lo1, (hi0 ^ c3 ^ k1) & MASK32, lo0
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def mode_phase(seed: int, index):
    """Uniform phase in (0, 2 pi) of half-spectrum mode `index` (linear index
    (i*NL + j)*(NL/2+1) + k): first Philox word of counter (index lo, index hi,
    0, 0) under key (seed lo, seed hi)."""
    index = np.asarray(index, np.uint64)
    lo = (index & MASK32).astype(np.uint32)
    hi = (index >> np.uint64(32)).astype(np.uint32)
    w = philox4x32_10(lo, hi, 0, 0, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)[0]
    return (w.astype(np.float64) + 0.5) * (2.0 * np.pi / 4294967296.0)


def gaussian_field_k(NL: int, L: float, A: float, ns: float, seed: int) -> np.ndarray:
    """Half spectrum delta_k, NL x NL x (NL/2+1), unnormalised forward convention."""
    kf = 2.0 * np.pi / L
    M1_ = NL // 2 + 1
    i = np.arange(NL)
    k1 = kf * np.where(i <= NL // 2, i, i - NL).astype(np.float64)
    kz = kf * np.arange(M1_, dtype=np.float64)
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    kmod = np.sqrt(k2)
    with np.errstate(divide="ignore"):
        P = np.where(kmod > 0, A * np.power(np.where(kmod > 0, kmod, 1.0), ns), 0.0)
    amp = np.sqrt(P * float(NL) ** 6 / L ** 3)
    ii, jj, kk = np.meshgrid(np.arange(NL), np.arange(NL), np.arange(M1_), indexing="ij")
    # Hermitian planes: the partner (-i, -j) of a kz = 0 or NL/2 mode; the mode
    # with the smaller linear index draws the phase, its partner negates it
    pi_, pj = (NL - ii) % NL, (NL - jj) % NL
    plane = (kk == 0) | (kk == NL // 2)
    own = (ii * NL + jj) * M1_ + kk
    partner = (pi_ * NL + pj) * M1_ + kk
    canon = np.where(plane, np.minimum(own, partner), own)
    theta = mode_phase(seed, canon.astype(np.uint64))
    sign = np.where(plane & (partner < own), -1.0, 1.0)
    dk = amp * np.exp(1j * sign * theta)
    selfc = plane & (partner == own)
    dk[selfc] = (amp * np.cos(theta))[selfc]
    dk[0, 0, 0] = 0.0
    return dk


def zeldovich(NL: int, L: float, A: float = 1e4, ns: float = -1.5, seed: int = 1,
              dtype=np.float32) -> np.ndarray:
    """(NL^3, 3) particle positions in [0, L), particle (i, j, k) at row
    (i*NL + j)*NL + k."""
    dk = gaussian_field_k(NL, L, A, ns, seed)
    kf = 2.0 * np.pi / L
    i = np.arange(NL)
    k1 = kf * np.where(i <= NL // 2, i, i - NL).astype(np.float64)
    kz = kf * np.arange(NL // 2 + 1, dtype=np.float64)
    h = NL // 2
    kd = k1.copy()
    kd[h] = 0.0
    kdz = kz.copy()
    kdz[-1] = 0.0
    k2 = k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2
    inv = np.zeros_like(k2)
    inv[k2 > 0] = 1.0 / k2[k2 > 0]
    d = L / NL
    q = (np.arange(NL, dtype=np.float64) + 0.5) * d
    pos = np.empty((NL, NL, NL, 3), np.float64)
    comps = (kd[:, None, None], kd[None, :, None], kdz[None, None, :])
    grids = np.meshgrid(q, q, q, indexing="ij")
    for a in range(3):
        psi = np.fft.irfftn(1j * comps[a] * dk * inv, s=(NL, NL, NL), axes=(0, 1, 2))
        pos[..., a] = grids[a] + psi
    pos = np.mod(pos, L).reshape(-1, 3)
    out = pos.astype(dtype)
    out[out >= L] = 0.0  # float rounding of values just below L
    return out


def uniform(n: int, L: float, seed: int) -> np.ndarray:
    """(n, 3) float32 uniform positions: Philox4x32-10 words of counter
    (i lo, i hi, 1, 0) under key (seed lo, seed hi), x = (w + 1/2) 2^-32 L."""
    i = np.arange(n, dtype=np.uint64)
    lo = (i & MASK32).astype(np.uint32)
    hi = (i >> np.uint64(32)).astype(np.uint32)
    w = philox4x32_10(lo, hi, 1, 0, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    out = np.stack([((x.astype(np.float64) + 0.5) * (1.0 / 4294967296.0) * L) for x in w[:3]], axis=1)
    out = out.astype(np.float32)
    out[out >= np.float32(L)] = 0.0
    return out
