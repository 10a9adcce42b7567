"""Shared seeded inputs for the parity tests."""
import numpy as np

from amdkiit_amd.pwgrid import Cell, PWSetup


def small_setup(L=10.0, ecut=12.0):
    return PWSetup(Cell.cubic(L), ecut)


def random_bands(pw, nb, seed=1):
    """c_i(G) = (u1 - 1/2 + i (u2 - 1/2)) / (1 + |G|^2), Im c(0) = 0."""
    rng = np.random.default_rng(seed)
    g2 = pw.g2_pw
    c = (rng.random((nb, pw.ngpw)) - 0.5 + 1j * (rng.random((nb, pw.ngpw)) - 0.5)) / (1.0 + g2)
    g0 = np.where(np.all(pw.mill_pw == 0, axis=1))[0]
    c[:, g0] = c[:, g0].real
    return c


def smooth_potential(pw, seed=2, amp=1.0):
    """A smooth real V(r): a few low Fourier modes, dense (n3, n2, n1)."""
    rng = np.random.default_rng(seed)
    n1, n2, n3 = pw.n
    z, y, x = np.meshgrid(np.arange(n3) / n3, np.arange(n2) / n2, np.arange(n1) / n1, indexing="ij")
    v = np.full((n3, n2, n1), -0.3)
    for _ in range(6):
        k = rng.integers(-3, 4, size=3)
        ph = rng.random() * 2 * np.pi
        v += amp * 0.2 * rng.random() * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
    return v
