"""UPF v2 pseudopotentials (readupf.f90, form_factor_upf, upf_nlppgrid /
upf_nlpro): the parser, the projector bookkeeping the reference uses, and the
oracle's SCF against the recorded tests/H2_upf run."""
import json
import os

import numpy as np
import pytest

from amdkiit_amd import psp, systems
from oracle.ks_oracle import KSOracle
from tests.common import random_bands

HERE = os.path.dirname(os.path.abspath(__file__))
UPF = os.path.join(HERE, "golden", "psp", "H_ONCV.upf")
REF = json.load(open(os.path.join(HERE, "golden", "reference_energies.json")))


def test_upf_header_mesh_and_dij():
    d = psp.read_upf(UPF)
    assert d["zv"] == 1.0 and d["lmax"] == 1 and d["mesh"] == 1180
    assert len(d["r"]) == len(d["rab"]) == len(d["vloc"]) == 1180
    assert [l for l, _ in d["betas"]] == [0, 0, 1]
    assert np.allclose(np.diag(d["dij"]), [-3.3306769125, -1.0390204702, -0.55014255401])
    assert np.count_nonzero(d["dij"] - np.diag(np.diag(d["dij"]))) == 0


def test_upf_projector_list_and_coupling():
    s = psp.UpfSpecies(UPF, "H", 80.0, 320.0)
    # iv over l, then m, then the projectors of that l (readupf.f90:296-309)
    assert s.ivs == [(0, 1, 1), (0, 1, 2), (1, 2, 1), (1, 3, 1), (1, 4, 1)]
    # hlsg = D / 2 (Ry -> Ha) on the diagonal of each (l, m) block
    assert np.allclose(np.diag(s.dmat()), [-1.66533845625, -0.5195102351, -0.275071277005, -0.275071277005,
                                           -0.275071277005])
    assert np.count_nonzero(s.dmat() - np.diag(np.diag(s.dmat()))) == 0
    g2 = np.linspace(0.0, 80.0, 7)
    rad = [p[2](g2) for p in s.projectors()]
    # the second l = 0 projector copies the first one's table (upf_nlppgrid)
    assert np.array_equal(rad[0], rad[1])
    assert np.array_equal(rad[2], rad[3]) and np.array_equal(rad[2], rad[4])
    # the l = 1 table is the transform of the file's second beta function
    d = psp.read_upf(UPF)
    r, rab = d["r"], d["rab"]
    w = psp.simpsn_weights(len(r))
    g = 3.0
    direct = 4 * np.pi * np.sum(w * rab * r * d["betas"][1][1] * psp.sph_bessel(1, g * r))
    assert abs(s.projectors()[2][2](g * g) - direct) < 1e-6 * abs(direct)
    # j_1(0) = 0: no l = 1 projector at G = 0; the l = 0 table at G = 0 is 4 pi int r beta
    assert abs(s.projectors()[2][2](0.0)) < 1e-12
    direct0 = 4 * np.pi * np.sum(w * rab * r * d["betas"][0][1])
    assert abs(s.projectors()[0][2](0.0) - direct0) < 1e-9 * abs(direct0)


def test_upf_local_form_factor_long_range_part_removed():
    """V_loc(r) -> -Z/r outside the core, so r^2 V + Z r erf(r / 1.2) decays and
    the G = 0 form factor is finite; its value is the SIMPSN sum itself."""
    s = psp.UpfSpecies(UPF, "H", 80.0, 320.0)
    d = psp.read_upf(UPF)
    r, rab = d["r"], d["rab"]
    tail = r > 8.0
    v = 0.5 * d["vloc"][tail]
    assert np.abs(v * r[tail] + 1.0).max() < 5e-6  # the file holds 8-digit values
    fint = r * r * 0.5 * d["vloc"] + psp.erf(r / psp.RAGGIO) * r
    direct = 4 * np.pi * np.sum(psp.simpsn_weights(len(r)) * fint * rab)
    assert abs(s.local(0.0) - direct) < 1e-12 * abs(direct)


def test_upf_off_diagonal_dij_is_rejected(tmp_path):
    txt = open(UPF).read()
    i = txt.index("<PP_DIJ")
    j = txt.index(">", i) + 1
    k = txt.index("</PP_DIJ>")
    vals = np.array(txt[j:k].split(), float)
    vals[1] = 0.25
    bad = txt[:j] + "\n" + " ".join(f"{v:.10E}" for v in vals) + "\n" + txt[k:]
    p = tmp_path / "bad.upf"
    p.write_text(bad)
    with pytest.raises(ValueError, match="non-diagonal"):
        psp.UpfSpecies(str(p), "H", 80.0, 320.0)


@pytest.mark.slow
def test_oracle_scf_reproduces_recorded_h2_upf():
    """tests/H2_upf: the last SP_ENERGY.dat line is -1.024460 (F16.6).  Giving
    each projector its own radial function instead moves E_tot to -1.1369, so
    this pins the reference's table indexing as well as the units and
    quadratures."""
    s = systems.reference_h2_upf()
    o = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc, workers=os.cpu_count() or 1)
    c = o.ortho(random_bands(s.pw, s.nbands, seed=1))
    for it in range(1, 400):
        c, e = o.scf_step(c, s.occ, it)
        if e["gemax"] < 1e-7:
            break
    ref = REF["h2_upf"]
    assert abs(e["etot"] - ref["etot"]) < ref["etot_halfwidth"]
