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
