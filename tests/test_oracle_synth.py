"""Checks of the generator restatement (oracle/zeldovich.py): Philox4x32-10
known-answer vectors, the fixed amplitude and Hermitian planes of the field,
and the Zel'dovich relation div Psi = -delta."""
import numpy as np

from oracle import zeldovich as z


def test_philox_known_answers():
    # the published Philox4x32-10 test vectors (Random123 kat_vectors)
    cases = [
        ((0, 0, 0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 6, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for args, want in cases:
        got = tuple(int(x) for x in z.philox4x32_10(*args))
        assert got == want, [hex(g) for g in got]


def test_field_amplitude_and_hermitian_planes():
    NL, L, A, ns = 16, 200.0, 3e3, -1.5
    dk = z.gaussian_field_k(NL, L, A, ns, seed=9)
    kf = 2 * np.pi / L
    i = np.arange(NL)
    k1 = kf * np.where(i <= NL // 2, i, i - NL)
    kz = kf * np.arange(NL // 2 + 1)
    k = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp2 = np.where(k > 0, A * np.where(k > 0, k, 1.0) ** ns * NL ** 6 / L ** 3, 0.0)
    # |delta_k|^2 = P N^6 / V on every mode that is not self-conjugate
    h = NL // 2
    selfc = np.zeros(dk.shape, bool)
    for a in (0, h):
        for b in (0, h):
            for c in (0, h):
                selfc[a, b, c] = True
    assert np.allclose(np.abs(dk[~selfc]) ** 2, amp2[~selfc], rtol=1e-12)
    assert np.all(np.abs(dk[selfc]) ** 2 <= amp2[selfc] * (1 + 1e-12))
    # Hermitian planes: the real field reproduces the half spectrum exactly
    x = np.fft.irfftn(dk, s=(NL, NL, NL))
    assert np.allclose(np.fft.rfftn(x), dk, atol=1e-9 * np.abs(dk).max())
    # phases are uniform-looking: the mean of e^{i theta} over many modes is small
    ph = dk[1:, 1:, 1:h] / np.abs(dk[1:, 1:, 1:h])
    assert abs(ph.mean()) < 5.0 / np.sqrt(ph.size)


def test_zeldovich_divergence_is_minus_delta():
    NL, L = 16, 64.0
    pos = z.zeldovich(NL, L, A=20.0, ns=-1.5, seed=4, dtype=np.float64)
    d = L / NL
    q = (np.arange(NL) + 0.5) * d
    grid = np.stack(np.meshgrid(q, q, q, indexing="ij"), axis=-1).reshape(-1, 3)
    psi = (pos - grid + L / 2) % L - L / 2
    assert np.abs(psi).max() < L / 4
    psi = psi.reshape(NL, NL, NL, 3)
    kf = 2 * np.pi / L
    i = np.arange(NL)
    kd = kf * np.where(i <= NL // 2, i, i - NL).astype(float)
    kd[NL // 2] = 0.0
    kz = kf * np.arange(NL // 2 + 1).astype(float)
    kz[-1] = 0.0
    div = (1j * kd[:, None, None] * np.fft.rfftn(psi[..., 0]) + 1j * kd[None, :, None] * np.fft.rfftn(psi[..., 1])
           + 1j * kz[None, None, :] * np.fft.rfftn(psi[..., 2]))
    dk = z.gaussian_field_k(NL, L, 20.0, -1.5, 4)
    # modes with no Nyquist component: i k . Psi = i k . (i k delta / k^2) = -delta
    h = NL // 2
    m = np.ones(dk.shape, bool)
    m[h, :, :] = m[:, h, :] = False
    m[:, :, -1] = False
    m[0, 0, 0] = False
    assert np.allclose(div[m], -dk[m], atol=1e-9 * np.abs(dk).max())


def test_uniform_range_and_moments():
    u = z.uniform(200000, 50.0, seed=2)
    assert u.dtype == np.float32 and u.min() >= 0 and u.max() < 50.0
    assert np.allclose(u.mean(axis=0), 25.0, atol=0.2)
    assert not np.array_equal(u, z.uniform(200000, 50.0, seed=3))
