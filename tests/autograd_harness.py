"""Test-only torch.autograd.Function around the device energy and its adjoint.

forward(R) evaluates E_tot on the GPU at positions R (bohr); backward returns
dE/dR = -F from the hand-written adjoint kernels (local iG V_loc S_I rho* and
nonlocal d<beta|psi>/dR parts with per-direction derivative projectors, plus
the Ewald short-range term).  Two energy maps are provided:
  * FixedPsiEnergy: E_tot[psi, R] at a fixed wavefunction; its R-gradient is
    the partial derivative the force kernels compute, exact for any psi;
  * drivers.GPUForceProvider: the Born-Oppenheimer surface (SCF converged at
    every R), where the same forces are the total derivative.
Only tests use this; production callers bind the C ABI.
"""
import numpy as np
import torch


class FixedPsiEnergy:
    """(E, F) at fixed bands c (caller order) for any positions."""

    def __init__(self, ctx, system, c):
        self.ctx, self.s, self.c = ctx, system, c

    def __call__(self, tau):
        self.ctx.set_positions(self.s.sp, tau)
        self.ctx.wfn_load(self.c, self.s.occ)
        e, f = self.ctx.scf_step_tfor(1, self.s.xc, True)
        return e["etot"], f


class ScfEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tau, energy_map):
        e, f = energy_map(tau.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.as_tensor(-np.asarray(f), dtype=torch.float64))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        (grad,) = ctx.saved_tensors
        return grad_e * grad, None


def energy_and_grad(energy_map, tau):
    """(E, dE/dR) through autograd."""
    t = torch.tensor(np.asarray(tau, np.float64), requires_grad=True)
    e = ScfEnergy.apply(t, energy_map)
    e.backward()
    return float(e), t.grad.numpy().copy()


def central_differences(energy_map, tau, atoms, h=1e-3):
    """dE/dR_{a,k} by (E(R + h e_ak) - E(R - h e_ak)) / 2h for the given atoms."""
    out = {}
    for a in atoms:
        for k in range(3):
            tp = np.array(tau, np.float64)
            tp[a, k] += h
            tm = np.array(tau, np.float64)
            tm[a, k] -= h
            out[a, k] = (energy_map(tp)[0] - energy_map(tm)[0]) / (2 * h)
    return out
