"""CPU checks that pin the FP64 oracle (no GPU)."""
import numpy as np
import pytest

from amdkiit_amd.pwgrid import ANG2BOHR, Cell, PWSetup
from oracle import pw_oracle as O
from tests.common import random_bands, small_setup, smooth_potential


def test_grid_and_sphere_sizes_match_reference_run():
    # tests/1_water/wf_opt/amdkiit.out: 108^3 grid, 80 Ry; per-rank PW counts sum
    # to 39,103, density G to 311,563; 46-48 and 182-184 rays x 48 ranks.
    pw = PWSetup(Cell.cubic(9.85 * ANG2BOHR), 80.0)
    assert pw.n == (108, 108, 108)
    assert pw.ngpw == 39103
    assert pw.ngrho == 311563
    assert pw.rays(pw.mill_pw) == 2217
    assert pw.rays(pw.mill_rho) == 8797


def test_backward_fft_matches_brute_force_dft():
    pw = PWSetup(Cell.cubic(6.0), 3.0, n=(12, 10, 8))
    rng = np.random.default_rng(0)
    c1 = rng.normal(size=pw.ngpw) + 1j * rng.normal(size=pw.ngpw)
    c2 = rng.normal(size=pw.ngpw) + 1j * rng.normal(size=pw.ngpw)
    c1[0] = c1[0].real
    c2[0] = c2[0].real
    got = O.fft_backward(O.pack_pair(c1, c2, pw.mill_pw, pw.n))
    n1, n2, n3 = pw.n
    z, y, x = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    psi1 = np.zeros(z.shape, complex)
    psi2 = np.zeros(z.shape, complex)
    for i, (a, b, d) in enumerate(pw.mill_pw):
        ph = np.exp(2j * np.pi * (a * x / n1 + b * y / n2 + d * z / n3))
        for c, psi in ((c1, psi1), (c2, psi2)):
            psi += c[i] * ph
            if (a, b, d) != (0, 0, 0):
                psi += np.conj(c[i]) * np.conj(ph)
    # both fields are real, and the packed transform is psi1 + i psi2
    assert np.abs(psi1.imag).max() < 1e-12 and np.abs(psi2.imag).max() < 1e-12
    np.testing.assert_allclose(got, psi1.real + 1j * psi2.real, atol=1e-11)
    # forward transform undoes it with the 1/N scale
    back = O.fft_forward(got)
    ip = O.grid_index(pw.mill_pw, pw.n)
    np.testing.assert_allclose(back[ip], c1 + 1j * c2, atol=1e-12)


def test_vpsi_free_electron_and_constant_potential():
    pw = small_setup()
    c = random_bands(pw, 5)
    occ = np.array([2.0, 2.0, 1.0, 0.0, 2.0])
    g2 = pw.g2_pw
    for v0 in (0.0, -0.7):
        v = np.full((pw.n[2], pw.n[1], pw.n[0]), v0)
        c2 = O.vpsi(c, occ, v, pw.mill_pw, pw.n, pw.cell.g)
        fi = np.where(occ == 0, 1.0, occ * 0.5)
        np.testing.assert_allclose(c2, -fi[:, None] * (g2 + 2 * v0) * c, atol=1e-12)


def test_vpsi_cosine_potential_couples_g_plus_minus_g0():
    # V(r) = 2 cos(G0.r) => (V psi)(G) = c(G - G0) + c(G + G0)
    pw = small_setup()
    c = random_bands(pw, 2, seed=3)
    n1, n2, n3 = pw.n
    z, y, x = np.meshgrid(np.arange(n3) / n3, np.arange(n2) / n2, np.arange(n1) / n1, indexing="ij")
    m0 = np.array([1, 2, 0])
    v = 2 * np.cos(2 * np.pi * (m0[0] * x + m0[1] * y + m0[2] * z))
    occ = np.array([2.0, 2.0])
    c2 = O.vpsi(c, occ, v, pw.mill_pw, pw.n, pw.cell.g)
    lut = {tuple(m): i for i, m in enumerate(pw.mill_pw)}

    def coef(b, m):
        m = tuple(m)
        if m in lut:
            return c[b, lut[m]]
        mm = tuple(-np.array(m))
        return np.conj(c[b, lut[mm]]) if mm in lut else 0.0

    g2 = pw.g2_pw
    for b in range(2):
        for i in range(0, pw.ngpw, 7):
            m = pw.mill_pw[i]
            vpsi = coef(b, m - m0) + coef(b, m + m0)
            ref = -(g2[i] * c[b, i] + 2 * vpsi)
            assert abs(c2[b, i] - ref) < 1e-12


def test_vpsi_hermitian():
    pw = small_setup()
    c = random_bands(pw, 2, seed=5)
    v = smooth_potential(pw)
    h = O.vpsi(c, np.array([2.0, 2.0]), v, pw.mill_pw, pw.n, pw.cell.g)  # = -H c
    a = O.dotp(c[0], -h[1])
    b = O.dotp(-h[0], c[1])
    assert abs(a - b) < 1e-12 * max(1.0, abs(a))


def test_density_integrates_to_sum_of_norms():
    pw = small_setup()
    c = random_bands(pw, 3, seed=7)
    occ = np.array([2.0, 2.0, 1.0])
    rho = O.eval_density(c, occ, pw.mill_pw, pw.n, pw.cell.omega)
    integral = rho.sum() * pw.cell.omega / pw.nr
    norms = sum(occ[i] * O.dotp(c[i], c[i]) for i in range(3))
    assert abs(integral - norms) < 1e-12 * norms
