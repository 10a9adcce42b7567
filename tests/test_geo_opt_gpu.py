"""Geometry optimisation on the device SCF against the reference's recorded
runs (tests/{bulk_si,1_water,2_water}/geo_opt/atom_force.dat): the driver of
do_opt.f90:197-266 with the restated L-BFGS (More-Thuente line search) must
print the same geometries, energies and gradients step after step."""
import os
import re

import numpy as np
import pytest

from amdkiit_amd import drivers as D
from amdkiit_amd import systems
from amdkiit_amd.context import KSContext
from tests.common import random_bands

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def recorded_steps(name):
    """(energy, gradient[nat, 3]) records of a recorded atom_force.dat."""
    lines = open(os.path.join(HERE, "golden", "outputs", name)).read().split("\n")
    out, i = [], 0
    while i < len(lines) and lines[i].strip():
        nat = int(lines[i].split()[0])
        e = float(re.match(r"\s*ENERGY=\s*(\S+)", lines[i + 1]).group(1))
        g = np.array([[float(v) for v in ln.split()[1:4]] for ln in lines[i + 2:i + 2 + nat]])
        out.append((e, g))
        i += 2 + nat
    return out


CASES = {
    # system, recorded file, wavefunction GRAD_CUTOFF, geometry GRAD_CUTOFF, records compared
    "h2o_1": (lambda: systems.reference_water(1), "h2o_1_geo_atom_force.dat", 1e-7, 1e-4, 4),
    "h2o_2": (lambda: systems.reference_water(2), "h2o_2_geo_atom_force.dat", 1e-6, 1e-4, 2),
}
# How far a recorded path can be followed: every evaluation's gradient
# carries the noise of an SCF converged to GEMAX < GRAD_CUTOFF (~1e-6 to 1e-5
# Ha/bohr here, different in the reference's run and in this one), and the
# correction pairs y = g_new - g_old of the later iterations are differences
# of gradients that change by ~1e-4, so the L-BFGS directions drift apart by
# ~10 % from the third record of 2 H2O on.  1 H2O starts with |g| = 8e-5,
# below the geometry cutoff of 1e-4 (not tested before the first iteration),
# so where the path stops is decided by the same noise: in mixed precision the
# first accepted iterate has |g| within a few 1e-6 of the cutoff and the run
# ends after its first or second record, depending on rounding in the SCF;
# FP64 follows at least two.


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_lbfgs_reproduces_recorded_geometry_steps(name, prec):
    sysf, fname, wcut, gcut, nrec = CASES[name]
    s = sysf()
    rec = recorded_steps(fname)[:nrec]
    print(f"case=geo_opt_{name} prec={prec} n={s.pw.n} nb={s.nbands} records={len(rec)}", flush=True)
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c0 = ctx.ortho(random_bands(s.pw, s.nbands, seed=1))
    grads = []

    class Recorder(D.GPUForceProvider):
        def __call__(self, tau):
            e, f = super().__call__(tau)
            grads.append((e, -np.array(f)))
            return e, f

    prov = Recorder(ctx, s, c0, gcut=wcut, maxit=3000)
    _, hist, _ = D.lbfgs(prov, s.tau.copy(), gtol=gcut, maxiter=nrec - 1,
                         log=lambda i, e, gm: print(f"  step {i} E {e:.12f} max|g| {gm:.3e}", flush=True))
    ctx.close()
    assert len(hist) >= (2 if prec == "fp64" or name != "h2o_1" else 1)
    # printed records are the evaluations whose hist entry came from a print
    printed = [g for g in grads if any(abs(g[0] - h[0]) == 0.0 for h in hist)]
    for k, (e_ref, g_ref) in enumerate(rec[:len(printed)]):
        e, g = printed[k]
        dE, dG = abs(e - e_ref) / s.nat, np.abs(g - g_ref).max()
        print(f"  record {k}: E {e:.10f} ref {e_ref:.10f} |dE|/atom {dE:.1e} max|dg| {dG:.1e} "
              f"(max|g| {np.abs(g_ref).max():.1e})", flush=True)
        assert dE < 1e-6, k
        assert dG < 1e-4, k


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fcc_silicon_optimisation_end_point(prec):
    """tests/bulk_si/geo_opt: the first record (E, dE/dR) and the energy of
    the run's last accepted geometry against the recorded final energy table.
    The path between them is not compared: the start is the ideal diamond,
    whose gradient is the one-image E_SR pair force along (1, -1, -1), and
    which of x, y or z the first accepted step singles out is decided by
    1e-8-level rounding asymmetries of the gradient (the reference breaks y,
    the restated oracle and this build break x).  Neither is the reference's
    termination: its last two SCF runs converged in their first step, where
    CHK_TFOR (gradient.f90:121-125) leaves TFOR false, so no forces were
    formed, LBFGS saw |g| = 0 and stopped (amdkiit.out:3071 prints GNORM
    0.000); this driver always takes the forces of a converged step and
    goes on until its line search gives up on the E_SR minimum-image kinks.
    Both runs reach the same minimum."""
    import json
    ref = json.load(open(os.path.join(HERE, "golden", "reference_forces.json")))["si2_geo_final"]
    rec0 = recorded_steps("si_geo_atom_force.dat")[0]
    s = systems.reference_si2()
    print(f"case=geo_opt_si2_full prec={prec}", flush=True)
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c0 = ctx.ortho(random_bands(s.pw, s.nbands, seed=1))
    calls = []

    class Counter(D.GPUForceProvider):
        def __call__(self, tau):
            e, f = super().__call__(tau)
            calls.append((e, -np.array(f)))
            return e, f

    prov = Counter(ctx, s, c0, gcut=1e-7, maxit=3000)
    x, hist, acc = D.lbfgs(prov, s.tau.copy(), gtol=1e-4, maxiter=100)
    ctx.close()
    e0, g0 = calls[0]
    print(f"  record 0: E {e0:.10f} ref {rec0[0]:.10f} max|dg| {np.abs(g0 - rec0[1]).max():.1e}", flush=True)
    print(f"  iterations={len(acc)} evaluations={len(calls)} (reference {ref['iterations']}, {ref['evaluations']}) "
          f"E={acc[-1][0]:.10f} ref={ref['etot']:.10f} |g|={acc[-1][2]:.2e}", flush=True)
    assert abs(e0 - rec0[0]) < 1e-6 * s.nat
    assert np.abs(g0 - rec0[1]).max() < 1e-5
    assert abs(min(a[0] for a in acc) - ref["etot"]) < 1e-6 * s.nat
