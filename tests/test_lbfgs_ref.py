"""amdkiit_amd.lbfgs (restated L-BFGS + More-Thuente search) against the
reference's own lbfgs.f, compiled from the reference tree into
oracle/_ref/ by oracle/Makefile: the same sequence of requested points,
IFLAG values and completed-iteration flags on test problems whose line
searches take the interpolation branches (overshooting first steps,
bracketing, extrapolation)."""
import numpy as np
import pytest

from amdkiit_amd.lbfgs import LBFGS
from oracle import ref_lbfgs

pytestmark = pytest.mark.skipif(not ref_lbfgs.available(), reason="oracle/_ref/liblbfgs_ref.so not built")


def rosenbrock(x):
    f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] += -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


def lennard_jones(x):
    r = x.reshape(-1, 3)
    d = r[:, None, :] - r[None, :, :]
    r2 = (d ** 2).sum(-1) + np.eye(len(r))
    inv6 = 1.0 / r2 ** 3
    e = 4 * (inv6 ** 2 - inv6)
    np.fill_diagonal(e, 0.0)
    w = 24 * (2 * inv6 ** 2 - inv6) / r2
    np.fill_diagonal(w, 0.0)
    g = -(w[:, :, None] * d).sum(1)
    return 0.5 * e.sum(), g.reshape(-1)


def soft_wells(x):
    """sum log(1 + (x - c)^2): the first step 1/|g| overshoots far."""
    c = np.linspace(-1.0, 1.0, x.size)
    d = x - c
    return float(np.sum(np.log1p(d * d))), 2 * d / (1 + d * d)


def lj_start():
    rng = np.random.default_rng(5)
    base = np.array([[0, 0, 0], [1.1, 0, 0], [0, 1.1, 0], [0, 0, 1.1], [1.1, 1.1, 0], [1.1, 0, 1.1], [0, 1.1, 1.1]], float)
    return (base + 0.08 * rng.normal(size=base.shape)).reshape(-1)


CASES = {
    "rosenbrock": (rosenbrock, np.array([-1.2, 1.0, -0.5, 0.8, 1.3, -0.7]), 1e-6, 40),
    "lennard_jones_7": (lennard_jones, lj_start(), 1e-5, 60),
    "soft_wells": (soft_wells, np.full(5, 7.0), 1e-8, 40),
}


@pytest.mark.parametrize("name", list(CASES))
def test_same_points_and_flags_as_the_reference(name):
    fun, x0, eps, ncalls = CASES[name]
    xa, xb = x0.copy(), x0.copy()
    mine, ref = LBFGS(x0.size, eps=eps), ref_lbfgs.RefLBFGS(x0.size, eps=eps)
    flags = []
    for k in range(ncalls):
        fa, ga = fun(xa)
        fb, gb = fun(xb)
        ra, rb = mine.step(xa, fa, ga), ref.step(xb, fb, gb)
        assert ra == rb, (k, ra, rb)
        assert mine.accepted == ref.accepted, k
        flags.append(mine.accepted)
        if ra != 1:
            break
        scale = max(1.0, np.abs(xb).max())
        assert np.abs(xa - xb).max() < 1e-9 * scale * (1 + k), (k, np.abs(xa - xb).max())
    # the path took line searches of more than one evaluation
    assert not all(flags[1:])
