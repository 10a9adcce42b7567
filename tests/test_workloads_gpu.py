"""The driver workloads at their full sizes, through the C ABI:

* 64 H2O PBE 80 Ry (144^3, 256 bands): Born-Oppenheimer NVE MD (velocity
  Verlet, update_velver.f90) on the device SCF and the TFOR forces, in FP64
  and mixed precision from the same start;
* 128 H2O PBE 80 Ry (180^3, 512 bands): geometry optimisation with the
  restated lbfgs.f (M = 5, More-Thuente search) as geom_opt_lbfgs
  (do_opt.f90:197-266) drives it, mixed precision.
"""
import numpy as np
import pytest

from amdkiit_amd import drivers as D
from amdkiit_amd import systems
from amdkiit_amd.context import KSContext

pytestmark = pytest.mark.gpu


def seeded_bands(s, seed):
    rng = np.random.default_rng(seed)
    c0 = (rng.random((s.nbands, s.pw.ngpw)) - 0.5 + 1j * (rng.random((s.nbands, s.pw.ngpw)) - 0.5)) / (1.0 + s.pw.g2_pw)
    c0[:, 0] = c0[:, 0].real
    return c0


def start(s, prec, seed, gcut, maxit=3000):
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c = ctx.ortho(seeded_bands(s, seed))
    return ctx, D.GPUForceProvider(ctx, s, c, gcut=gcut, maxit=maxit)


def test_bomd_64_waters_conserves_energy_and_precisions_agree():
    """8 steps of 0.5 fs at 300 K: the constant of motion stays within
    1.5e-5 Ha/atom of its start in both precisions (1 % of the kinetic energy
    per atom, 1.5 kT = 1.4e-3 Ha; Verlet's own oscillation at this step, with
    the O-H stretch at ~9 fs, measured 5.8e-6), and the mixed-precision E_pot
    follows FP64 within 1e-6 Ha/atom at every step."""
    s = systems.CONFIGS["h2o_64_pbe"]()
    m = np.array([D.MASS_AMU[s.species[i].symbol] for i in s.sp]) * D.AMU_AU
    v0 = D.maxwell_boltzmann(m, 300.0, seed=2024)
    runs = {}
    for prec in ("fp64", "mixed"):
        print(f"case=bomd_64h2o prec={prec} n={s.pw.n} nb={s.nbands} nat={s.nat}", flush=True)
        ctx, prov = start(s, prec, seed=7, gcut=1e-6)
        h, drift, _, _ = D.bomd(prov, s.tau.copy(), m, v0.copy(), dt_fs=0.5, nsteps=8,
                                log=lambda i, e, ek, ec: print(f"  md {i} E_pot {e:.8f} E_cons {ec:.8f}", flush=True))
        ctx.close()
        dev = np.abs(h[:, 2] - h[0, 2]).max() / s.nat
        print(f"  {prec}: max |E_cons - E_cons(0)| = {dev:.2e} Ha/atom, SCF steps {prov.nscf}", flush=True)
        assert dev < 1.5e-5
        runs[prec] = h
    d = np.abs(runs["mixed"][:, 0] - runs["fp64"][:, 0]).max() / s.nat
    print(f"max |E_pot mixed - fp64| = {d:.2e} Ha/atom", flush=True)
    assert d < 1e-6


def test_lbfgs_128_waters_lowers_energy_and_forces():
    """Two printed L-BFGS steps from the seeded 128-water box (SCF to
    GEMAX < 3e-5 at every geometry): each accepted iterate lowers E_tot, and
    the last one has a smaller gradient than the start."""
    s = systems.CONFIGS["h2o_128_pbe"]()
    print(f"case=lbfgs_128h2o prec=mixed n={s.pw.n} nb={s.nbands} nat={s.nat}", flush=True)
    ctx, prov = start(s, "mixed", seed=9, gcut=3e-5, maxit=6000)
    _, hist, acc = D.lbfgs(prov, s.tau.copy(), gtol=1e-4, maxiter=2,
                           log=lambda i, e, fm: print(f"  lbfgs {i} E {e:.8f} max|g| {fm:.2e}", flush=True))
    ctx.close()
    print(f"  SCF steps per evaluation {prov.nscf}", flush=True)
    assert len(acc) >= 1
    es = [hist[0][0]] + [a[0] for a in acc]
    assert all(b < a for a, b in zip(es, es[1:]))
    assert acc[-1][2] < hist[0][2]
