"""CPU pins of the FP64 oracle and the host setup against closed forms,
library routines and the reference's recorded energies."""
import json
import math
import os

import numpy as np
import pytest
import scipy.interpolate as SI
import scipy.special as SS

from amdkiit_amd import psp, systems
from amdkiit_amd.pwgrid import Cell, PWSetup
from oracle import xc_oracle as X
from oracle.ks_oracle import KSOracle
from tests.common import random_bands

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "reference_energies.json")))


# -- numerics of the setup -------------------------------------------------------
def test_natural_spline_matches_scipy():
    x = np.linspace(0.0, 3.0, 40)
    y = np.sin(2 * x) * np.exp(-x)
    q = np.linspace(0.0, 2.999, 500)
    ref = SI.CubicSpline(x, y, bc_type="natural")(q)
    np.testing.assert_allclose(psp.NaturalSpline(x, y)(q), ref, atol=1e-13)


def test_simpsn_weights_integrate_polynomials():
    # SIMPSN drops point N and its end weights (109, -5, 63, 49)/48 make the
    # rule exact for cubics over [x_1 - h, x_{N-1} + h] (one step beyond the
    # first and the last point it sums); its integrands vanish at both ends.
    n = 101
    x = np.linspace(0.0, 1.0, n)
    h = x[1] - x[0]
    a, b = -h, x[n - 2] + h
    for p in range(4):
        got = psp.simpsn(x ** p) * h
        exact = (b ** (p + 1) - a ** (p + 1)) / (p + 1)
        assert abs(got - exact) < 1e-12


def test_spherical_bessel_matches_scipy():
    x = np.concatenate([np.linspace(1e-6, 0.99, 50), np.linspace(1.0, 40.0, 200)])
    for l in range(4):
        np.testing.assert_allclose(psp.sph_bessel(l, x), SS.spherical_jn(l, x), atol=1e-10)


def test_ylm_normalisation_and_order():
    rng = np.random.default_rng(0)
    G = rng.normal(size=(20000, 3))
    # Monte Carlo over directions: <Y_L Y_L'> = delta / (4 pi)
    Y = np.stack([psp.ylm(L, G) for L in range(1, 10)])
    M = Y @ Y.T / len(G) * 4 * np.pi
    assert np.abs(M - np.eye(9)).max() < 0.05
    g = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
    c = math.sqrt(3 / (4 * np.pi))
    assert np.allclose([psp.ylm(2, g)[0], psp.ylm(3, g)[1], psp.ylm(4, g)[2]], [c, c, c])


def test_gth_local_matches_numerical_fourier_transform():
    s = psp.gth("O", "PBE")
    r = np.linspace(1e-6, 30.0, 300001)
    dr = r[1] - r[0]
    rl = s.rloc
    c1, c2, c3, c4 = s.cs
    x = r / rl
    vloc = (-s.zv / r * SS.erf(r / (math.sqrt(2) * rl))
            + np.exp(-0.5 * x * x) * (c1 + c2 * x ** 2 + c3 * x ** 4 + c4 * x ** 6))
    short = vloc + s.zv * SS.erf(r / psp.RAGGIO) / r
    for g in (0.0, 0.7, 2.3, 5.0):
        j0 = np.sinc(g * r / np.pi)
        num = 4 * np.pi * np.sum(r * r * short * j0) * dr
        assert abs(s.local(np.array([g * g]))[0] - num) < 1e-6


def test_grid_rule_all_configs():
    sizes = {"h2o_32_pbe": ((100,) * 3, 31995), "h2o_64_pbe": ((144,) * 3, 78098)}
    for L, E, n, npw in [(18.63, 70.0, 100, 31995), (23.47, 80.0, 144, 78098), (29.57, 80.0, 180, 156277),
                         (18.9, 70.0, 108, 33438), (20.5224, 30.0, 72, 11936)]:
        pw = PWSetup(Cell.cubic(L), E)
        assert pw.n == (n, n, n) and pw.ngpw == npw


# -- exchange-correlation ------------------------------------------------------------
def test_lda_closed_forms():
    rho = np.array([1e-3, 0.02, 0.2, 0.5, 3.0])
    rs = (3 / (4 * np.pi * rho)) ** (1 / 3)
    f, v, _ = X.evaluate("PZ", rho, 0 * rho)
    ex = -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3)
    # Perdew-Zunger correlation: Pade branch for rs >= 1, log branch below
    hi = -0.1423 / (1 + 1.0529 * np.sqrt(rs) + 0.3334 * rs)
    lo = 0.0311 * np.log(rs) - 0.048 + 0.0020 * rs * np.log(rs) - 0.0116 * rs
    ec = np.where(rs >= 1, hi, lo)
    assert list(rs >= 1) == [True, True, True, False, False]
    np.testing.assert_allclose(f / rho, ex + ec, rtol=1e-13)
    # the two branches meet at rs = 1 to the published fit accuracy
    assert abs((-0.1423 / (1 + 1.0529 + 0.3334)) - (-0.048 - 0.0116)) < 1e-4
    # v = d(rho eps)/d rho against a central difference
    h = 1e-6 * rho
    fp, _, _ = X.evaluate("PZ", rho + h, 0 * rho)
    fm, _, _ = X.evaluate("PZ", rho - h, 0 * rho)
    np.testing.assert_allclose(v, (fp - fm) / (2 * h), rtol=1e-7)


def test_teter93_reproduces_pz_to_fit_accuracy():
    # the GTH Pade form is a fit to the PZ/CA LDA: agreement to ~1e-3 relative
    rho = np.logspace(-4, 0, 30)
    ft, _, _ = X.evaluate("TETER93", rho, 0 * rho)
    fp, _, _ = X.evaluate("PZ", rho, 0 * rho)
    assert np.abs(ft / fp - 1).max() < 5e-3


def test_pbe_limits_and_derivatives():
    rho = np.array([0.01, 0.1, 0.5])
    # sigma -> 0: PBE -> LDA exchange + PW92 correlation
    f0, _, _ = X.evaluate("PBE", rho, np.full(3, 1e-30))
    rs = (3 / (4 * np.pi * rho)) ** (1 / 3)
    ex = -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3)
    np.testing.assert_allclose(f0 / rho, ex + X.pw92_c(rs), rtol=1e-12)
    # large gradients: F_x -> 1 + kappa and eps_c + H -> 0 (H -> -eps_c)
    kf = (3 * np.pi ** 2 * rho) ** (1 / 3)
    sig = (2 * kf * rho * 1e4) ** 2
    ratio = X.pbe(rho, sig) / (rho * ex)
    assert np.all(np.abs(ratio - 1.804) < 1e-6)
    # complex-step derivatives against central differences
    sigma = np.array([1e-4, 3e-3, 0.2])
    f, vr, vs = X.evaluate("PBE", rho, sigma)
    for i in range(3):
        h = 1e-6 * rho[i]
        d = (X.pbe(rho[i] + h, sigma[i]) - X.pbe(rho[i] - h, sigma[i])) / (2 * h)
        assert abs(d - vr[i]) < 1e-7 * max(1, abs(d))
        h = 1e-6 * sigma[i]
        d = (X.pbe(rho[i], sigma[i] + h) - X.pbe(rho[i], sigma[i] - h)) / (2 * h)
        assert abs(d - vs[i]) < 1e-6 * max(1e-3, abs(d))


def test_smoothing_matches_reference_thresholds():
    r = np.array([-1.0, 5e-10, 1e-9, 2e-9, 4e-8, 4.1e-8])
    smoo, dsmoo, act = X.smooth(r)
    assert list(act) == [False, False, False, True, True, True]
    assert smoo[-1] == 1.0 and smoo[-2] < 1.0


# -- ions --------------------------------------------------------------------------
def test_structure_factor_and_ewald_direct():
    s = systems.water_box((1, 1, 1), 9.0, 8.0, "PZ", jitter=0.0)
    o = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc)
    G = s.pw.cell.gvec(s.pw.mill_rho[:50])
    for a in range(s.nat):
        np.testing.assert_allclose(o.eigr[:50, a], np.exp(-1j * G @ s.tau[a]), atol=1e-14)
    # e_sr: direct sum over the three pairs (boxes large enough for minimum image)
    esr, f = o.e_sr()
    z = s.tables.zv[s.sp]
    rc = math.sqrt(2) * 1.2
    e = 0
    for i in range(3):
        for j in range(i + 1, 3):
            r = np.linalg.norm(s.tau[i] - s.tau[j])
            e += z[i] * z[j] * math.erfc(r / rc) / r
    assert abs(esr - e) < 1e-14
    assert np.abs(f.sum(0)).max() < 1e-13  # Newton's third law


# -- the whole SCF against the reference's recorded run ---------------------------------
@pytest.mark.slow
def test_oracle_scf_reproduces_recorded_single_water():
    s = systems.reference_water(1)
    o = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc, workers=os.cpu_count() or 1)
    c = o.ortho(random_bands(s.pw, s.nbands, seed=1))
    for it in range(1, 300):
        c, e = o.scf_step(c, s.occ, it)
        if e["gemax"] < 1e-6:
            break
    ref = REF["h2o_1"]
    assert abs(e["etot"] - ref["etot"]) < 3e-6  # 1e-6 Ha/atom
    assert abs(e["eself"] - ref["eself"]) < 1e-9
    assert abs(e["esr"] - ref["esr"]) < 1e-9
    for k in ("ekin", "eloc", "enl", "exc", "estat"):
        assert abs(e[k] - ref[k]) < 1e-4, k


@pytest.mark.slow
def test_oracle_forces_are_minus_energy_gradient():
    s = systems.water_box((1, 1, 2), 10.0, 20.0, "PZ")
    def converge(o, c):
        for it in range(1, 2000):
            c, e = o.scf_step(c, s.occ, it)
            if e["gemax"] < 2e-8:
                break
        return c, e
    o = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc)
    c, e = converge(o, o.ortho(random_bands(s.pw, s.nbands, seed=2)))
    rho = o.density(c, s.occ)
    F = o.forces(c, s.occ, o.rhor2g(rho), o.fnl(c))
    h = 1e-3
    for a, k in [(0, 0), (2, 1)]:
        E = []
        for sgn in (1, -1):
            tau = s.tau.copy()
            tau[a, k] += sgn * h
            o2 = KSOracle(s.pw, s.tables, s.sp, tau, s.xc)
            E.append(converge(o2, c.copy())[1]["etot"])
        fd = -(E[0] - E[1]) / (2 * h)
        assert abs(F[a, k] - fd) < 1e-4, (a, k, F[a, k], fd)
