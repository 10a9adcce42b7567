"""Host drivers around the SCF: Born-Oppenheimer MD and L-BFGS geometry
optimisation (the reference's molecular_dynamics.f90 / update_velver.f90 and
do_opt.f90:197-266 + lbfgs.f), driven by the device-resident SCF of libamdk,
writing the reference's output files when given a RunOutputs.

The drivers only need a force provider `provider(tau) -> (E, F)` (Hartree,
Ha/bohr, F = -dE/dR), so they are tested on the CPU with analytic potentials
and on the GPU with `GPUForceProvider`.
"""
import math
import time

import numpy as np

from .systems import SplitMix64

AMU_AU = 1822.888485          # constants.f90:8
KBOLTZ = 0.316681534e-5       # Ha/K, constants.f90:7
AU_TO_FS = 2.418884326505e-2  # constants.f90:6
MASS_AMU = {"H": 1.00794, "O": 15.9994, "Si": 28.0855}


class GPUForceProvider:
    """E_tot and forces from a converged device SCF (wfn_opt semantics: the
    PCG history restarts at istep = 1 for every geometry, the previous
    wavefunction is the starting guess, forces are taken in the converged
    step before its wavefunction update, as TFOR does in do_opt.f90:97-106)."""

    def __init__(self, ctx, system, c0, gcut=1e-6, maxit=2000, outputs=None):
        self.ctx, self.s, self.gcut, self.maxit = ctx, system, gcut, maxit
        self.outputs = outputs  # RunOutputs: SP_ENERGY.dat lines per SCF step
        self.ctx.wfn_load(c0, system.occ)
        self.nscf = []

    def __call__(self, tau):
        s = self.s
        self.ctx.set_positions(s.sp, tau)
        self.ctx.ortho_resident()
        prev = math.inf
        f = None
        for it in range(1, self.maxit + 1):
            want = prev < 3.0 * self.gcut
            t0 = time.perf_counter()
            e, fi = self.ctx.scf_step_tfor(it, s.xc, want)
            if self.outputs is not None:
                self.outputs.sp_energy(it, e["etot"], 1e3 * (time.perf_counter() - t0))
            f = fi  # forces only of this step's state, never an earlier one
            prev = e["gemax"]
            if e["gemax"] < self.gcut and f is not None:
                self.nscf.append(it)
                return e["etot"], f
        raise RuntimeError(f"SCF did not converge in {self.maxit} steps (GEMAX {prev:.2e})")


def maxwell_boltzmann(masses_au, temperature, seed):
    """Seeded MB velocities (Box-Muller on SplitMix64 uniforms), zero total
    momentum, rescaled to `temperature` with ndof = 3N - 3 (init_vel,
    md_tools.f90:19-49, with an independent draw per atom)."""
    rng = SplitMix64(seed)
    n = len(masses_au)
    u = rng.uniform(6 * n).reshape(n, 3, 2)
    g = np.sqrt(-2.0 * np.log(1.0 - u[..., 0])) * np.cos(2 * math.pi * u[..., 1])
    v = g * np.sqrt(temperature * KBOLTZ / masses_au)[:, None]
    v -= (masses_au[:, None] * v).sum(0) / masses_au.sum()
    ndof = 3 * n - 3
    t = (masses_au[:, None] * v * v).sum() / (ndof * KBOLTZ)
    return v * math.sqrt(temperature / t)


def kinetic(masses_au, v):
    return 0.5 * float((masses_au[:, None] * v * v).sum())


def temperature(masses_au, v):
    """Instantaneous temperature with ndof = 3N - 3 (temp_ke, md_tools.f90:50-67;
    molecular_dynamics.f90:39)."""
    return 2.0 * kinetic(masses_au, v) / (KBOLTZ * (3 * len(v) - 3))


def bomd(provider, tau0, masses_au, v0, dt_fs, nsteps, log=None, outputs=None, symbols=None):
    """NVE velocity Verlet (velup_velver / posup_velver, update_velver.f90:14-48)
    with the centre of mass held fixed (shift_com).  Returns per-step
    (E_pot, E_kin, E_cons) and the drift of E_cons in Ha/atom/ps.  With
    `outputs` (RunOutputs) and `symbols`, MD_TRAJECTORY.xyz and MD_ENERGY.dat
    get one frame per step from step 0, as molecular_dynamics.f90:74-127."""
    dt = dt_fs / AU_TO_FS
    tau = np.array(tau0, np.float64)
    v = np.array(v0, np.float64)
    m = np.asarray(masses_au, np.float64)
    com0 = (m[:, None] * tau).sum(0) / m.sum()

    def record(step, e, ek, t0):
        if outputs is not None:
            outputs.md_trajectory(step, symbols, tau, v)
            outputs.md_energy(step, temperature(m, v), e, e + ek, time.perf_counter() - t0)

    t0 = time.perf_counter()
    e, f = provider(tau)
    hist = [(e, kinetic(m, v), e + kinetic(m, v))]
    record(0, e, hist[0][1], t0)
    for step in range(1, nsteps + 1):
        t0 = time.perf_counter()
        v += 0.5 * dt * f / m[:, None]
        tau += dt * v
        tau -= (m[:, None] * tau).sum(0) / m.sum() - com0
        e, f = provider(tau)
        v += 0.5 * dt * f / m[:, None]
        ek = kinetic(m, v)
        hist.append((e, ek, e + ek))
        record(step, e, ek, t0)
        if log:
            log(step, e, ek, e + ek)
    h = np.array(hist)
    t_ps = np.arange(len(h)) * dt_fs * 1e-3
    drift = np.polyfit(t_ps, h[:, 2], 1)[0] / len(tau) if len(h) > 2 else 0.0
    return h, drift, tau, v


def lbfgs(provider, x0, m=5, gtol=1e-4, maxiter=100, log=None, outputs=None, symbols=None):
    """Geometry optimisation as geom_opt_lbfgs (do_opt.f90:197-266) runs it:
    evaluate E and F at x, print the geometry when the previous LBFGS call
    completed an iteration (the first geometry included), hand (x, E, -F) to
    the reverse-communication L-BFGS of amdkiit_amd.lbfgs (M = 5, XTOL =
    1e-16, EPS = gtol on the 2-norm of the gradient), which either asks for
    the next geometry or stops; at most maxiter printed steps.  Note that a
    printed geometry is the first trial point of a line search (the point
    LBFGS asks for right after completing an iteration), not necessarily the
    accepted one.

    Returns (x, hist, accepted): x the last accepted geometry (the start if
    none), hist [(E, max|g|, |g|)] of the printed geometries, accepted the
    same for every accepted iterate.  With `outputs` and `symbols`, each
    printed geometry goes to atom_coord.xyz and its gradient dE/dR to
    atom_force.dat (PRINT_COORDINATE, do_opt.f90:244-249)."""
    from .lbfgs import LBFGS

    shape = np.shape(x0)
    x = np.array(x0, np.float64).reshape(-1)
    opt = LBFGS(x.size, m=m, eps=gtol, xtol=1e-16)
    hist, accepted = [], []
    x_acc = x.copy()
    print_next, iopt = True, 0
    while True:
        e, f = provider(x.reshape(shape))
        g = -np.asarray(f, np.float64).reshape(-1)
        rec = (e, float(np.abs(g).max()), float(np.sqrt(g @ g)))
        if print_next:
            if outputs is not None:
                outputs.atom_coord(iopt, e, symbols, x.reshape(-1, 3))
                outputs.atom_force(iopt, e, symbols, g.reshape(-1, 3))
            hist.append(rec)
            if log:
                log(iopt, e, rec[1])
            iopt += 1
        x_eval = x.copy()
        iflag = opt.step(x, e, g)
        if opt.accepted:
            accepted.append(rec)
            x_acc = x_eval
        print_next = opt.accepted
        if iflag <= 0 or iopt > maxiter:
            break
    return x_acc.reshape(shape), hist, accepted
