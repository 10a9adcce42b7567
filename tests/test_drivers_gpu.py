"""BO-MD and L-BFGS driven by the device SCF and the adjoint forces."""
import numpy as np
import pytest

from amdkiit_amd import drivers as D
from amdkiit_amd import systems
from amdkiit_amd.context import KSContext
from tests.common import random_bands

pytestmark = pytest.mark.gpu


def water_pair():
    s = systems.water_box((1, 1, 2), 10.0, 20.0, "PZ")
    ctx = KSContext(s.pw, s.nbands, "fp64")
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c0 = ctx.ortho(random_bands(s.pw, s.nbands, seed=3))
    return s, ctx, c0


def test_tfor_forces_match_separate_force_call_at_convergence():
    s, ctx, c0 = water_pair()
    prov = D.GPUForceProvider(ctx, s, c0, gcut=1e-7)
    e, f = prov(s.tau)
    print(f"scf steps={prov.nscf[-1]} E={e:.10f}", flush=True)
    f2 = ctx.scf_forces()
    assert np.abs(f - f2).max() < 1e-5
    assert np.abs(f.sum(0)).max() < 1e-4  # no net force beyond grid aliasing
    ctx.close()


def test_born_oppenheimer_forces_match_central_differences():
    """On the converged surface the same forces are the total derivative."""
    from tests.autograd_harness import central_differences, energy_and_grad
    s, ctx, c0 = water_pair()
    prov = D.GPUForceProvider(ctx, s, c0, gcut=1e-8)
    e0, g0 = energy_and_grad(prov, s.tau)
    fd = central_differences(prov, s.tau, [0, 2], h=1e-3)
    for (a, k), v in fd.items():
        print(f"atom {a} dir {k}: fd={v:+.8f} adjoint={g0[a, k]:+.8f}", flush=True)
        assert abs(v - g0[a, k]) < 1e-4
    ctx.close()


def test_bomd_energy_error_is_second_order_in_dt():
    """Velocity Verlet on the GPU surface: over the same 2 fs, halving dt cuts
    the constant-of-motion error about fourfold (forces consistent with E)."""
    s, ctx, c0 = water_pair()
    prov = D.GPUForceProvider(ctx, s, c0, gcut=1e-8)
    m = np.array([D.MASS_AMU["O" if sp == 0 else "H"] for sp in s.sp]) * D.AMU_AU
    v = D.maxwell_boltzmann(m, 300.0, seed=11)
    err = []
    for dt, n in ((0.5, 4), (0.25, 8)):
        prov.ctx.wfn_load(c0, s.occ)
        h, _, _, _ = D.bomd(prov, s.tau.copy(), m, v.copy(), dt_fs=dt, nsteps=n,
                            log=lambda *a: print("md", *a, flush=True))
        err.append(np.abs(h[:, 2] - h[0, 2]).max())
        print(f"dt={dt} fs: max |E_cons - E_cons(0)| = {err[-1]:.3e} Ha", flush=True)
    assert 2.5 < err[0] / err[1] < 6.0
    ctx.close()


def test_lbfgs_lowers_energy_and_forces():
    s, ctx, c0 = water_pair()
    prov = D.GPUForceProvider(ctx, s, c0, gcut=1e-7)
    x, hist, acc = D.lbfgs(prov, s.tau.copy(), gtol=1e-3, maxiter=4,
                           log=lambda *a: print("opt", *a, flush=True))
    assert acc[-1][0] < hist[0][0]
    assert acc[-1][2] < hist[0][2]
    ctx.close()
