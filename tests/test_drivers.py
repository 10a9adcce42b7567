"""Host drivers (BO-MD velocity Verlet, L-BFGS) on analytic potentials."""
import math

import numpy as np

from amdkiit_amd import drivers as D


def spring_provider(k, x0, d0=1.5):
    """Translation-invariant pair springs 0.5 k (|r_ij| - d0)^2 plus a weak
    anharmonic term, so the net force is zero as for a periodic cell."""
    n = len(x0)

    def p(tau):
        e = 0.0
        f = np.zeros_like(tau)
        for i in range(n):
            for j in range(i + 1, n):
                r = tau[i] - tau[j]
                d = float(np.linalg.norm(r))
                x = d - d0
                e += 0.5 * k * x * x + 0.1 * k * x ** 4
                g = (k * x + 0.4 * k * x ** 3) * r / d
                f[i] -= g
                f[j] += g
        return e, f
    return p


def test_maxwell_boltzmann_temperature_and_momentum():
    m = np.array([16.0, 1.0, 1.0, 16.0, 1.0, 1.0]) * D.AMU_AU
    v = D.maxwell_boltzmann(m, 300.0, seed=5)
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
    t = (m[:, None] * v * v).sum() / ((3 * len(m) - 3) * D.KBOLTZ)
    assert abs(t - 300.0) < 1e-9
    assert not np.allclose(v[0], v[3])  # independent draws per atom


def test_velocity_verlet_conserves_energy_and_is_reversible():
    rng = np.random.default_rng(0)
    x0 = rng.normal(size=(4, 3))
    m = np.array([1.0, 2.0, 3.0, 4.0]) * D.AMU_AU
    k = 0.3
    prov = spring_provider(k, x0)
    tau = x0 + 0.1 * rng.normal(size=(4, 3))
    v = D.maxwell_boltzmann(m, 300.0, seed=1)
    h, drift, tau1, v1 = D.bomd(prov, tau, m, v, dt_fs=0.1, nsteps=400)
    # second-order integrator: bounded energy error, no secular drift
    span = h[:, 0].max() - h[:, 0].min()
    assert span > 1e-4  # the system actually moves
    assert np.abs(h[:, 2] - h[0, 2]).max() < 1e-2 * span
    assert abs(drift) < 1e-3
    # time reversal returns to the start (COM shift keeps it fixed)
    _, _, tau2, v2 = D.bomd(prov, tau1, m, -v1, dt_fs=0.1, nsteps=400)
    com = lambda t: (m[:, None] * t).sum(0) / m.sum()
    assert np.abs((tau2 - com(tau2)) - (tau - com(tau))).max() < 1e-8


def test_lbfgs_minimises_quadratic_and_rosenbrock():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(6, 6))
    H = A @ A.T + 6 * np.eye(6)
    xs = rng.normal(size=6)

    def quad(x):
        d = x - xs
        return 0.5 * float(d @ H @ d), -(H @ d)

    x, hist, acc = D.lbfgs(quad, np.zeros(6), gtol=1e-8, maxiter=60)
    assert np.abs(x - xs).max() < 1e-6

    def rosen(x):
        a, b = x
        e = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        g = np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])
        return e, -g

    x, hist, acc = D.lbfgs(rosen, np.array([-1.2, 1.0]), gtol=1e-6, maxiter=200)
    assert np.abs(x - 1.0).max() < 1e-4
    # the line search's sufficient-decrease test: accepted energies decrease
    assert all(acc[i + 1][0] <= acc[i][0] + 1e-12 for i in range(len(acc) - 1))
    assert acc[-1][2] <= 1e-6


def test_drivers_write_the_reference_output_files(tmp_path):
    """MD writes MD_TRAJECTORY.xyz and MD_ENERGY.dat frames from step 0, the
    geometry optimiser atom_coord.xyz and atom_force.dat (gradient) per
    accepted geometry; the files read back to the driver's own numbers."""
    from amdkiit_amd.outputs import ANGS2BOHR, RunOutputs
    rng = np.random.default_rng(3)
    x0 = rng.normal(size=(3, 3)) * 1.5
    m = np.array([16.0, 1.0, 1.0]) * D.AMU_AU
    prov = spring_provider(0.3, x0)
    v = D.maxwell_boltzmann(m, 300.0, seed=2)
    out = RunOutputs(str(tmp_path / "md"))
    h, _, tau, v1 = D.bomd(prov, x0, m, v, dt_fs=0.2, nsteps=5, outputs=out, symbols=["O", "H", "H"])
    en = [ln.split() for ln in (tmp_path / "md" / "MD_ENERGY.dat").read_text().splitlines()]
    assert [int(r[0]) for r in en] == list(range(6))
    assert abs(float(en[0][1]) - 300.0) < 1e-9  # temperature with ndof = 3N - 3
    assert np.allclose([float(r[2]) for r in en], h[:, 0], rtol=0, atol=1e-14)
    assert np.allclose([float(r[3]) for r in en], h[:, 2], rtol=0, atol=1e-14)
    tr = (tmp_path / "md" / "MD_TRAJECTORY.xyz").read_text().splitlines()
    assert len(tr) == 6 * 5 and tr[1].split() == ["MD", "STEP=", "0"] and tr[-4].split()[-1] == "5"
    last = np.array([[float(x) for x in ln[2:].split()] for ln in tr[-3:]])
    assert np.abs(last[:, :3] * ANGS2BOHR - tau).max() < 1e-5
    assert np.abs(last[:, 3:] * ANGS2BOHR - v1).max() < 1e-5
    out = RunOutputs(str(tmp_path / "opt"))
    seen = []

    def prov_rec(t):
        seen.append(np.array(t))
        return prov(t)

    x, hist, acc = D.lbfgs(prov_rec, x0, gtol=1e-6, maxiter=30, outputs=out, symbols=["O", "H", "H"])
    fr = (tmp_path / "opt" / "atom_force.dat").read_text().splitlines()
    assert len(fr) == 5 * len(hist)
    e_last = float(fr[-4].split()[1])
    g_last = np.array([[float(v) for v in ln.split()[1:]] for ln in fr[-3:]])
    co = (tmp_path / "opt" / "atom_coord.xyz").read_text().splitlines()
    x_last = np.array([[float(v) for v in ln.split()[1:]] for ln in co[-3:]]) / 0.529177
    # the last printed geometry is one of the evaluated ones, with its E and dE/dR
    k = int(np.argmin([np.abs(t - x_last).max() for t in seen]))
    assert np.abs(seen[k] - x_last).max() < 1e-5
    e, f = prov(seen[k])
    assert abs(e_last - e) < 1e-14 and np.abs(g_last + f).max() < 1e-14  # gradient = -F
    assert abs(hist[-1][0] - e) < 1e-14
