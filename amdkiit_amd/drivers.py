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


def lbfgs(provider, x0, m=5, gtol=1e-4, maxiter=20, log=None, outputs=None, symbols=None):
    """L-BFGS on E(R) with the two-loop recursion and M = 5 corrections, as the
    reference's geom_opt_lbfgs (do_opt.f90:197-266) calls lbfgs.f.  The step
    length comes from a strong-Wolfe line search (c1 = 1e-4, c2 = 0.9, Nocedal &
    Wright alg. 3.5/3.6) instead of lbfgs.f's Moré-Thuente routine.  With
    `outputs` and `symbols`, every accepted geometry (the first included) is
    appended to atom_coord.xyz and atom_force.dat, the latter holding the
    gradient dE/dR as the reference's PRINT_COORDINATE call does
    (do_opt.f90:244-249)."""
    x = np.array(x0, np.float64).reshape(-1)
    shape = np.shape(x0)

    def fg(xx):
        e, f = provider(xx.reshape(shape))
        return e, -np.asarray(f, np.float64).reshape(-1)

    def record(iopt, e, xx, g):
        if outputs is not None:
            outputs.atom_coord(iopt, e, symbols, xx.reshape(-1, 3))
            outputs.atom_force(iopt, e, symbols, g.reshape(-1, 3))

    e, g = fg(x)
    S, Y = [], []
    hist = [(e, float(np.abs(g).max()))]
    record(0, e, x, g)
    for it in range(1, maxiter + 1):
        if np.abs(g).max() < gtol:
            break
        q = g.copy()
        al = []
        for s_, y_ in reversed(list(zip(S, Y))):
            r = 1.0 / float(y_ @ s_)
            a = r * float(s_ @ q)
            al.append((a, r, s_, y_))
            q -= a * y_
        if S:
            q *= float(S[-1] @ Y[-1]) / float(Y[-1] @ Y[-1])
        else:
            q *= min(1.0, 0.1 / max(np.abs(g).max(), 1e-12))
        for a, r, s_, y_ in reversed(al):
            b = r * float(y_ @ q)
            q += (a - b) * s_
        d = -q
        if float(d @ g) >= 0:
            d = -g
        step, e1, g1 = _wolfe(fg, x, e, g, d)
        s_ = step * d
        y_ = g1 - g
        x = x + s_
        e, g = e1, g1
        if float(y_ @ s_) > 1e-12:
            S.append(s_)
            Y.append(y_)
            if len(S) > m:
                S.pop(0)
                Y.pop(0)
        hist.append((e, float(np.abs(g).max())))
        record(it, e, x, g)
        if log:
            log(it, e, float(np.abs(g).max()))
    return x.reshape(shape), hist


def _wolfe(fg, x, e0, g0, d, c1=1e-4, c2=0.9, amax=4.0, maxev=12):
    dg0 = float(g0 @ d)
    a_prev, e_prev, dg_prev = 0.0, e0, dg0
    a = 1.0
    for i in range(maxev):
        e, g = fg(x + a * d)
        dg = float(g @ d)
        if e > e0 + c1 * a * dg0 or (i > 0 and e >= e_prev):
            return _zoom(fg, x, e0, dg0, d, a_prev, e_prev, dg_prev, a, e, dg, c1, c2)
        if abs(dg) <= -c2 * dg0:
            return a, e, g
        if dg >= 0:
            return _zoom(fg, x, e0, dg0, d, a, e, dg, a_prev, e_prev, dg_prev, c1, c2)
        a_prev, e_prev, dg_prev = a, e, dg
        a = min(2.0 * a, amax)
    return a, e, g


def _zoom(fg, x, e0, dg0, d, alo, elo, dglo, ahi, ehi, dghi, c1, c2, maxev=12):
    for _ in range(maxev):
        # cubic interpolation between the bracket ends, safeguarded to the middle 80 %
        d1 = dglo + dghi - 3.0 * (elo - ehi) / (alo - ahi)
        rad = d1 * d1 - dglo * dghi
        if rad > 0:
            d2 = math.copysign(math.sqrt(rad), ahi - alo)
            a = ahi - (ahi - alo) * (dghi + d2 - d1) / (dghi - dglo + 2 * d2)
        else:
            a = 0.5 * (alo + ahi)
        lo, hi = min(alo, ahi), max(alo, ahi)
        a = min(max(a, lo + 0.1 * (hi - lo)), hi - 0.1 * (hi - lo))
        e, g = fg(x + a * d)
        dg = float(g @ d)
        if e > e0 + c1 * a * dg0 or e >= elo:
            ahi, ehi, dghi = a, e, dg
        else:
            if abs(dg) <= -c2 * dg0:
                return a, e, g
            if dg * (ahi - alo) >= 0:
                ahi, ehi, dghi = alo, elo, dglo
            alo, elo, dglo = a, e, dg
    return a, e, g
