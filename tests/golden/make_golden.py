"""Writes the small golden fixtures of tests/golden/ (run from the repository
root: python tests/golden/make_golden.py).  The values come from library
routines and closed forms, never from the code under test:

* fft_8.npz: a seeded 8^3 float64 field and numpy.fft.rfftn of it (the
  unnormalised forward transform with FFTW's sign and half-spectrum layout);
* fft_12.npz: the same for a 12^3 field (radix 3 and 4);
* eig.npz: 32 seeded symmetric 3x3 matrices and numpy.linalg.eigvalsh of them;
* stencils.json: the closed-form NGP/CIC/TSC/PCS weights at a few in-cell
  offsets (the B-spline kernels of order 1..4 at the nodes -2..2).
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def bspline(order, x):
    """Centred cardinal B-spline of the given order (1 NGP .. 4 PCS) at x."""
    x = abs(x)
    if order == 1:
        return 1.0 if x < 0.5 else 0.0
    if order == 2:
        return max(0.0, 1.0 - x)
    if order == 3:
        if x < 0.5:
            return 0.75 - x * x
        if x < 1.5:
            return 0.5 * (1.5 - x) ** 2
        return 0.0
    if x < 1.0:
        return (4.0 - 6.0 * x * x + 3.0 * x ** 3) / 6.0
    if x < 2.0:
        return (2.0 - x) ** 3 / 6.0
    return 0.0


def main():
    rng = np.random.default_rng(20261203)
    for n in (8, 12):
        x = rng.standard_normal((n, n, n))
        np.savez(os.path.join(HERE, f"fft_{n}.npz"), x=x, k=np.fft.rfftn(x))
    m = rng.standard_normal((32, 3, 3))
    m = 0.5 * (m + m.transpose(0, 2, 1))
    np.savez(os.path.join(HERE, "eig.npz"), m=m, ev=np.linalg.eigvalsh(m))
    offsets = [-0.49, -0.25, 0.0, 0.125, 0.3, 0.49]
    st = {}
    for name, order in (("NGP", 1), ("CIC", 2), ("TSC", 3), ("PCS", 4)):
        # weight of node j (j = -2..2 relative to the nearest node) for a particle at offset s
        st[name] = {str(s): [bspline(order, j - s) for j in range(-2, 3)] for s in offsets}
    with open(os.path.join(HERE, "stencils.json"), "w") as f:
        json.dump({"offsets": offsets, "nodes": [-2, -1, 0, 1, 2], "weights": st}, f, indent=1)


if __name__ == "__main__":
    main()
