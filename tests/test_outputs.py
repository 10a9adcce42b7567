"""Output files in the reference's layouts, checked line for line against
excerpts of its recorded runs (tests/golden/outputs)."""
import os

import numpy as np

from amdkiit_amd.outputs import ANGS2BOHR, BOHR_COORD_PRINT, RunOutputs, list_directed

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "outputs")


def lines(name):
    with open(os.path.join(GOLD, name)) as f:
        return f.read().split("\n")[:-1]


def test_sp_energy_layout(tmp_path):
    out = RunOutputs(str(tmp_path))
    want = lines("h2o_1_wf_SP_ENERGY.dat")
    for ln in want:
        t = ln.split()
        out.sp_energy(int(t[0]), float(t[1]), float(t[2]))
    assert (tmp_path / "SP_ENERGY.dat").read_text().split("\n")[:-1] == want


def test_md_energy_layout(tmp_path):
    for name in ("h2o_1_md_MD_ENERGY.dat", "h2o_32_md_MD_ENERGY.dat"):
        out = RunOutputs(str(tmp_path / name))
        want = lines(name)
        for ln in want:
            t = ln.split()
            out.md_energy(int(t[0]), *[float(x) for x in t[1:]])
        assert (tmp_path / name / "MD_ENERGY.dat").read_text().split("\n")[:-1] == want


def test_md_trajectory_layout(tmp_path):
    out = RunOutputs(str(tmp_path))
    want = lines("h2o_1_md_MD_TRAJECTORY.xyz")
    i = 0
    while i < len(want):
        n = int(want[i].split()[0])
        step = int(want[i + 1].split()[-1])
        rows = [want[i + 2 + k] for k in range(n)]
        sym = [r[:2].strip() for r in rows]
        vals = np.array([[float(x) for x in r[2:].split()] for r in rows])
        out.md_trajectory(step, sym, vals[:, :3] * ANGS2BOHR, vals[:, 3:] * ANGS2BOHR)
        i += 2 + n
    assert (tmp_path / "MD_TRAJECTORY.xyz").read_text().split("\n")[:-1] == want


def _frames(want):
    i = 0
    while i < len(want):
        n = int(want[i].split()[0])
        t = want[i + 1].split()
        rows = [want[i + 2 + k].split() for k in range(n)]
        yield int(t[2]), float(t[1]), [r[0] for r in rows], np.array([[float(x) for x in r[1:]] for r in rows])
        i += 2 + n


def test_atom_force_layout(tmp_path):
    for name in ("h2o_1_geo_atom_force.dat", "h2o_2_geo_atom_force.dat", "si_geo_atom_force.dat"):
        out = RunOutputs(str(tmp_path / name))
        want = lines(name)
        for step, e, sym, f in _frames(want):
            out.atom_force(step, e, sym, f)
        assert (tmp_path / name / "atom_force.dat").read_text().split("\n")[:-1] == want


def test_atom_coord_layout(tmp_path):
    for name in ("h2o_1_geo_atom_coord.xyz", "si_geo_atom_coord.xyz"):
        want = lines(name)
        got = []
        for step, e, sym, x in _frames(want):
            got.append(list_directed(len(sym)))
            got.append(list_directed("ENERGY=", e, step))
            got += [list_directed(f"{s:<2}", *row) for s, row in zip(sym, x)]
        assert got == want
    # the writer prints bohr positions times 0.529177
    out = RunOutputs(str(tmp_path / "c"))
    out.atom_coord(3, -1.5, ["O"], [[1.0, 2.0, -0.5]])
    row = (tmp_path / "c" / "atom_coord.xyz").read_text().split("\n")[2]
    assert np.allclose([float(v) for v in row.split()[1:]], [BOHR_COORD_PRINT, 2 * BOHR_COORD_PRINT,
                                                              -0.5 * BOHR_COORD_PRINT], rtol=1e-15)


def test_list_directed_edge_values():
    assert list_directed(0) == "           0"
    assert list_directed(1, 0.5) == "           1  0.50000000000000000     "
    assert list_directed(1, 0.0) == "           1   0.0000000000000000     "
    assert list_directed("X ", -1.25e-3) == " X   -1.2500000000000000E-003"
    assert list_directed("X ", 9.999999999999999999) == " X    10.000000000000000     "
