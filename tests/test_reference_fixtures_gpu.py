"""Converged GPU runs against the forces, gradients and energies the
reference's recorded runs printed (tests/golden/reference_forces.json,
tests/golden/reference_energies.json): E_tot and its terms, and -dE/dR of
every atom, in both precision modes."""
import json
import os

import numpy as np
import pytest

from amdkiit_amd import systems
from amdkiit_amd.context import KSContext
from tests.common import random_bands

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FORCES = json.load(open(os.path.join(HERE, "golden", "reference_forces.json")))
ENERGIES = json.load(open(os.path.join(HERE, "golden", "reference_energies.json")))


def converge_with_forces(s, prec, gcut, maxit=2000, seed=1):
    """SCF from seeded random bands until GEMAX < gcut, then one more step with
    TFOR: energies and forces of that step's state (do_opt.f90:97-106)."""
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c0 = ctx.ortho(random_bands(s.pw, s.nbands, seed=seed))
    ctx.wfn_load(c0, s.occ)
    for it in range(1, maxit + 1):
        e = ctx.scf_step(it, s.xc)
        if e["gemax"] < gcut:
            break
    e, f = ctx.scf_step_tfor(it + 1, s.xc, True)
    ctx.close()
    return e, f, it


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fcc_silicon_energy_table_and_gradient(prec):
    """tests/bulk_si: E_tot and its terms from the single-point run, -F from
    the first geometry step (FCC cell, FCC minimum image of E_SR, whose
    one-image pair force is the whole 1.553e-3 Ha/bohr gradient)."""
    s = systems.reference_si2()
    print(f"case=si2_fcc prec={prec} n={s.pw.n} npw={s.pw.ngpw}", flush=True)
    e, f, it = converge_with_forces(s, prec, 1e-8 if prec == "fp64" else 1e-7)
    ref = FORCES["si2_single_point"]
    grad = np.array([a["grad"] for a in FORCES["si2_geo_step0"]["atoms"]])
    dF = np.abs(-f - grad).max()
    print(f"  steps={it} etot={e['etot']:.10f} ref={ref['etot']:.10f} max|dF|={dF:.2e}", flush=True)
    assert abs(e["etot"] - ref["etot"]) < 1e-6 * s.nat
    assert abs(e["esr"] - ref["esr"]) < 1e-10
    assert abs(e["eself"] - ref["eself"]) < 1e-9
    for k in ("ekin", "eloc", "enl", "exc"):
        assert abs(e[k] - ref[k]) < 2e-5, (k, e[k], ref[k])
    assert abs(e["estat"] - ref["estat"]) < 2e-5
    assert dF < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("nmol", [1, 2])
def test_water_energy_and_gradients(nmol, prec):
    """tests/{1,2}_water: the wf_opt energy table and the full-precision
    gradient of the first geometry step (the same coordinates)."""
    s = systems.reference_water(nmol)
    print(f"case=h2o_{nmol} prec={prec} n={s.pw.n} nb={s.nbands}", flush=True)
    e, f, it = converge_with_forces(s, prec, 1e-8 if prec == "fp64" else 1e-7)
    ref = ENERGIES[f"h2o_{nmol}"]
    geo = FORCES[f"h2o_{nmol}_geo_step0"]
    # gradient records list the atoms O first, then H, as reference_water orders them
    assert [a["sym"] for a in geo["atoms"]] == ["O" if i == 0 else "H" for i in s.sp]
    grad = np.array([a["grad"] for a in geo["atoms"]])
    dF = np.abs(-f - grad).max()
    print(f"  steps={it} etot={e['etot']:.10f} ref={ref['etot']:.10f} geo={geo['energy']:.10f} "
          f"max|F|={np.abs(grad).max():.2e} max|dF|={dF:.2e}", flush=True)
    assert abs(e["etot"] - ref["etot"]) < 1e-6 * s.nat
    assert abs(e["etot"] - geo["energy"]) < 1e-6 * s.nat
    for k in ("ekin", "eloc", "enl", "exc", "esr", "eself", "estat"):
        assert abs(e[k] - ref[k]) < 1e-5, (k, e[k], ref[k])
    assert dF < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_32_water_forces_of_the_md_start(prec):
    """tests/32_water/md: the 96 forces printed after the initial SCF (GEMAX <
    1e-6, 4 significant figures, up to 8e-2 Ha/bohr).  Atoms are matched by
    their printed coordinates (bohr, 5 decimals)."""
    s = systems.reference_water(32)
    print(f"case=h2o_32_md prec={prec} n={s.pw.n} nb={s.nbands}", flush=True)
    e, f, it = converge_with_forces(s, prec, 1e-6)
    rows = FORCES["h2o_32_md_step0"]["atoms"]
    xyz = np.array([r["coord_bohr"] for r in rows])
    fref = np.array([r["force"] for r in rows])
    idx = [int(np.argmin(np.linalg.norm(xyz - t, axis=1))) for t in s.tau]
    assert sorted(idx) == list(range(len(rows)))
    assert np.abs(xyz[idx] - s.tau).max() < 1e-4
    assert [rows[i]["sym"] for i in idx] == ["O" if i == 0 else "H" for i in s.sp]
    dF = np.abs(f - fref[idx])
    print(f"  steps={it} etot={e['etot']:.10f} max|F|={np.abs(fref).max():.3e} max|dF|={dF.max():.2e} "
          f"mean|dF|={dF.mean():.2e}", flush=True)
    assert dF.max() < 1e-4
