"""Shared helpers of the GPU tests (imported by tests marked gpu only)."""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def uniform_pos(n, L, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (n, 3)).astype(dtype)


def clustered_pos(n, L, seed, dtype=np.float32, nclump=20, frac=0.7):
    """Uniform background plus tight clumps (tests tiles split into chunks)."""
    rng = np.random.default_rng(seed)
    nc = int(n * frac)
    centers = rng.uniform(0, L, (nclump, 3))
    idx = rng.integers(0, nclump, nc)
    clump = centers[idx] + rng.normal(0, 0.01 * L, (nc, 3))
    bg = rng.uniform(0, L, (n - nc, 3))
    return np.mod(np.concatenate([clump, bg]), L).astype(dtype)


def exact_mode_counts(N, ns, Dmin, ndel):
    """Modes per linear shell by integer enumeration of the stored half
    spectrum the estimator visits (x/y Nyquist planes left out, the kz = 0
    plane's half by the octant rule): shell s holds q = |k/k_f|^2 with
    (s ndel + Dmin)^2 <= q < ((s+1) ndel + Dmin)^2."""
    h = N // 2
    ix = np.arange(N)
    f = np.minimum(ix, N - ix)
    lo = ix < h
    hi = ix > h
    qxy = (f[:, None] ** 2 + f[None, :] ** 2).astype(np.int64)
    edges = np.array([(s * ndel + Dmin) ** 2 for s in range(ns + 1)], np.int64)
    counts = np.zeros(ns, np.int64)
    for kz in range(h + 1):
        if kz == 0:
            m = (lo[:, None] & lo[None, :] & ((ix[:, None] + ix[None, :]) > 0)) | (hi[:, None] & lo[None, :]
                                                                                  & (ix[None, :] > 0))
        else:
            m = (lo[:, None] | hi[:, None]) & (lo[None, :] | hi[None, :])
        q = qxy[m] + kz * kz
        s = np.searchsorted(edges, q, side="right") - 1
        s = s[(s >= 0) & (s < ns)]
        counts += np.bincount(s, minlength=ns)
    return counts
