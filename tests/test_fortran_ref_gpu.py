"""The reference's routines under their own names (Fortran module amdk_ref)
on the reference's array layouts, driven in the wfn_opt order by
amdk_wfopt, against the device-resident C path (amdk_scf_step_tfor) from
the same start: energy terms of every step, GEMAX/CNORM, the TFOR forces in
force(3, NA_MAX, sp_t) and the final C_0."""
import os
import subprocess

import numpy as np
import pytest

from amdkiit_amd import lib as L
from amdkiit_amd import systems
from amdkiit_amd.context import KSContext
from oracle.ks_oracle import KSOracle
from tests.common import random_bands

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "amdkiit_amd", "fortran", "amdk_wfopt")

CASES = {
    "h2o2_pbe": lambda: systems.water_box((1, 1, 2), 10.0, 20.0, "PBE"),
    "h2_upf": lambda: systems.h2_upf_small(),
    "si2_fcc": lambda: systems.reference_si2(),
}


def write_setup(path, s, c0, prec, nstep):
    T = s.tables
    nsp = T.nsp
    counts = np.bincount(s.sp, minlength=nsp)
    na_max = int(counts.max()) + 1  # a spare slot: the layout need not be tight
    ngh_max = max(1, T.maxproj)
    atco = np.zeros((nsp, na_max, 3))
    for sp in range(nsp):
        atco[sp, :counts[sp]] = s.tau[s.sp == sp]
    hdr = np.array([*s.pw.n, s.pw.cell.symmetry, prec, L.XC[s.xc], s.nbands, nsp, na_max, ngh_max,
                    s.pw.ngpw, s.pw.ngrho, nstep], np.int32)

    def tab(x, shape):
        out = np.zeros(shape)
        x = np.asarray(x, np.float64)
        out[tuple(slice(0, d) for d in x.shape)] = x
        return out

    with open(path, "wb") as f:
        hdr.tofile(f)
        np.asarray(s.pw.cell.a, np.float64).ravel(order="F").tofile(f)
        np.ascontiguousarray(s.pw.mill_pw, np.int32).tofile(f)
        np.ascontiguousarray(s.pw.mill_rho, np.int32).tofile(f)
        s.occ.astype(np.float64).tofile(f)
        counts.astype(np.int32).tofile(f)
        atco.tofile(f)
        np.asarray(T.zv, np.float64).tofile(f)
        np.asarray(T.vps, np.float64).tofile(f)
        np.asarray(T.rhos, np.float64).tofile(f)
        np.asarray(T.nproj, np.int32).tofile(f)
        tab(np.asarray(T.proj_l).reshape(nsp, -1), (nsp, ngh_max)).astype(np.int32).tofile(f)
        tab(np.asarray(T.twnl).reshape(nsp, -1, s.pw.ngpw), (nsp, ngh_max, s.pw.ngpw)).tofile(f)
        tab(np.asarray(T.dmat).reshape(nsp, T.maxproj, T.maxproj) if T.maxproj else np.zeros((nsp, 0, 0)),
            (nsp, ngh_max, ngh_max)).tofile(f)
        # WSG(nsp, nghmax) in Fortran order: species fastest
        np.ascontiguousarray(tab(T.wsg, (nsp, ngh_max)).T).tofile(f)
        np.ascontiguousarray(c0, np.complex128).tofile(f)
    return counts, na_max


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_reference_routines_match_the_resident_path(name, prec, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("amdk_wfopt not built (amdflang missing)")
    s = CASES[name]()
    nstep = 4
    orc = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc)
    c0 = orc.ortho(random_bands(s.pw, s.nbands, seed=17))
    p = L.PREC_FP64 if prec == "fp64" else L.PREC_MIXED
    counts, na_max = write_setup(str(tmp_path / "setup.bin"), s, c0, p, nstep)
    print(f"case=fortran_ref_{name} prec={prec} n={s.pw.n} nb={s.nbands} nat={s.nat}", flush=True)
    r = subprocess.run([EXE, str(tmp_path / "setup.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    raw = np.fromfile(str(tmp_path / "out.bin"), np.float64)
    steps = raw[:10 * nstep].reshape(nstep, 10)
    nsp = s.tables.nsp
    force = raw[10 * nstep:10 * nstep + 3 * na_max * nsp].reshape(nsp, na_max, 3)
    cf = raw[10 * nstep + 3 * na_max * nsp:].view(np.complex128).reshape(s.nbands, s.pw.ngpw)

    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    ctx.wfn_load(c0, s.occ)
    tol = 1e-10 if prec == "fp64" else 2e-6
    omega = s.pw.cell.omega
    for it in range(1, nstep + 1):
        e, f = ctx.scf_step_tfor(it, s.xc, it == nstep)
        got = dict(zip(("etot", "ekin", "eh", "eloc", "enl", "exc", "esr", "eself", "gemax", "cnorm"), steps[it - 1]))
        want = dict(e)
        want["eh"] = e["eh"] * omega
        for k, v in got.items():
            assert abs(v - want[k]) < tol * max(1.0, abs(want[k])), (it, k, v, want[k])
    # force(3, NA_MAX, sp_t): atom ia of species sp; the spare slots stay zero
    idx = np.zeros(s.nat, int)
    for sp in range(nsp):
        idx[s.sp == sp] = np.arange(counts[sp])
    fr = force[s.sp, idx]
    assert np.abs(fr - f).max() < (1e-9 if prec == "fp64" else 1e-6)
    for sp in range(nsp):
        assert np.all(force[sp, counts[sp]:] == 0.0)
    c_res = ctx.wfn_store()
    ctx.close()
    assert np.abs(cf - c_res).max() / np.abs(c_res).max() < (1e-9 if prec == "fp64" else 1e-5)
