"""GPU parity of the local H|psi> and density against the FP64 oracle."""
import numpy as np
import pytest

from amdkiit_amd.context import KSContext
from amdkiit_amd.pwgrid import Cell, PWSetup
from oracle import pw_oracle as O
from tests.common import random_bands, small_setup, smooth_potential

pytestmark = pytest.mark.gpu

CASES = [
    ("small_odd_bands", lambda: small_setup(), 5),
    ("noncubic_grid", lambda: PWSetup(Cell(np.diag([9.0, 10.0, 11.0])), 10.0), 4),
    ("generic_lengths_16_12_10", lambda: PWSetup(Cell(np.diag([8.0, 6.5, 5.2])), 5.0, n=(16, 12, 10)), 3),
    ("cfg1_shape_108", lambda: PWSetup(Cell.cubic(18.9), 70.0), 4),
    ("cfg3_shape_72", lambda: PWSetup(Cell.cubic(20.5224), 30.0), 3),
    ("cfg2_shape_100", lambda: PWSetup(Cell.cubic(18.63), 70.0), 2),
    ("cfg4_shape_144", lambda: PWSetup(Cell.cubic(23.47), 80.0), 2),
    ("cfg5_shape_180", lambda: PWSetup(Cell.cubic(29.57), 80.0), 2),
]


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name,mk,nb", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_vpsi_matches_oracle(name, mk, nb, prec, tol):
    pw = mk()
    print(f"case={name} n={pw.n} ngpw={pw.ngpw} nb={nb} prec={prec}", flush=True)
    c = random_bands(pw, nb, seed=11)
    occ = np.array([2.0] * (nb - 1) + [0.0])
    v = smooth_potential(pw)
    ctx = KSContext(pw, nb, prec)
    ctx.set_vks(v)
    got = ctx.vpsi(c, occ)
    ref = O.vpsi(c, occ, v, pw.mill_pw, pw.n, pw.cell.g)
    assert rel(got, ref) < tol
    ctx.close()


@pytest.mark.parametrize("name,mk,nb", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_density_matches_oracle(name, mk, nb, prec, tol):
    pw = mk()
    print(f"case={name} n={pw.n} ngpw={pw.ngpw} nb={nb} prec={prec}", flush=True)
    c = random_bands(pw, nb, seed=12)
    occ = np.array([2.0] * (nb - 1) + [1.0])
    ctx = KSContext(pw, nb, prec)
    got = ctx.eval_density(c, occ)
    ref = O.eval_density(c, occ, pw.mill_pw, pw.n, pw.cell.omega)
    assert rel(got, ref) < tol
    ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_chunked_pairs_match_single_pass(prec, monkeypatch):
    """A small work-array budget splits the band pairs into several chunks
    (W1/W2 reused per chunk); H|psi> and rho must not change."""
    pw = PWSetup(Cell.cubic(18.63), 70.0)
    nb = 11
    c = random_bands(pw, nb, seed=5)
    occ = np.full(nb, 2.0)
    v = smooth_potential(pw)
    out = {}
    for budget in (None, "20"):
        if budget:
            monkeypatch.setenv("AMDK_PAIR_BUDGET_MB", budget)
        ctx = KSContext(pw, nb, prec)
        ctx.set_vks(v)
        out[budget] = (ctx.vpsi(c, occ), ctx.eval_density(c, occ))
        ctx.close()
        monkeypatch.delenv("AMDK_PAIR_BUDGET_MB", raising=False)
    assert np.array_equal(out[None][0], out["20"][0])
    assert rel(out["20"][1], out[None][1]) < 1e-13
