"""Input structures: the reference's recorded test systems and the seeded
synthetic configurations (water boxes, displaced diamond Si)."""
import math
import os

import numpy as np

from . import psp
from .pwgrid import ANG2BOHR, Cell, PWSetup

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def read_xyz(path):
    """Coordinates as readcoordinate (yamlread.f90:358-424) reads them: two
    header lines skipped, then one atom per line to the end of the file."""
    with open(path) as f:
        lines = f.read().split("\n")
    sym, pos = [], []
    for ln in lines[2:]:
        t = ln.split()
        if len(t) < 4:
            continue
        sym.append(t[0])
        pos.append([float(x) for x in t[1:4]])
    return sym, np.array(pos)


class System:
    def __init__(self, pw, species, sp_of_atom, tau, nbands, occ, xc, name):
        self.pw, self.species, self.sp, self.tau = pw, species, np.asarray(sp_of_atom, np.int32), np.asarray(tau)
        self.nbands, self.occ, self.xc, self.name = nbands, np.asarray(occ, np.float64), xc, name
        self.tables = psp.Tables(pw, species)

    @property
    def nat(self):
        return len(self.tau)


def reference_water(nmol, ecut=80.0, box_ang=9.85):
    """Recorded MT/Teter93 water runs (tests/{1,2,32}_water/wf_opt)."""
    sym, pos = read_xyz(os.path.join(GOLDEN, "structures", f"h2o_{nmol}.xyz"))
    pw = PWSetup(Cell.cubic(box_ang * ANG2BOHR), ecut)
    sO = psp.CpmdNumericSpecies(os.path.join(GOLDEN, "psp", "O_MT_LDA.psp"), "O", 1, True, pw.ecut, pw.ecut_rho)
    sH = psp.CpmdNumericSpecies(os.path.join(GOLDEN, "psp", "H_MT_LDA.psp"), "H", 0, False, pw.ecut, pw.ecut_rho)
    # species order O then H, atoms grouped by species like the reference's atco
    order = [i for i, s in enumerate(sym) if s == "O"] + [i for i, s in enumerate(sym) if s == "H"]
    sp = [0 if sym[i] == "O" else 1 for i in order]
    tau = pos[order] * ANG2BOHR
    nel = 6 * nmol + 2 * nmol
    nb = nel // 2
    return System(pw, [sO, sH], sp, tau, nb, np.full(nb, 2.0), "TETER93", f"h2o_{nmol}_ref")


def reference_si2():
    """Recorded FCC silicon runs (tests/bulk_si/{single_point,geo_opt}): two
    atoms in the FCC primitive cell of a = 5.43103 A (set_vec's vectors,
    set_cell.f90:49-52; symmetry 2, whose minimum image E_SR uses), MT Si
    (KB, LMAX P: s projector, p local), XC_LDA_XC_TETER93, CUTOFF absent ->
    80 Ry (the WAVEFUNCTION_CUTOFF key is not read, yamlread.f90:132-138),
    4 doubly occupied states."""
    sym, pos = read_xyz(os.path.join(GOLDEN, "structures", "si2_fcc.xyz"))
    pw = PWSetup(Cell.fcc(5.43103 * ANG2BOHR), 80.0)
    sSi = psp.CpmdNumericSpecies(os.path.join(GOLDEN, "psp", "Si_MT_LDA.psp"), "Si", 1, True, pw.ecut, pw.ecut_rho)
    return System(pw, [sSi], [0] * len(sym), pos * ANG2BOHR, 4, np.full(4, 2.0), "TETER93", "si2_ref")


def reference_h2_upf():
    """Recorded UPF run (tests/H2_upf): H2 in a 15.1178 bohr cubic cell,
    ONCV H UPF v2 (input LMAX S, LOCAL 0 are not read for UPF species),
    CUTOFF absent -> 80 Ry (yamlread.f90:132-138), XC_LDA_XC_TETER93,
    coordinates in bohr, one doubly occupied state."""
    sym, pos = read_xyz(os.path.join(GOLDEN, "structures", "h2_upf.xyz"))
    pw = PWSetup(Cell.cubic(15.1178), 80.0)
    sH = psp.UpfSpecies(os.path.join(GOLDEN, "psp", "H_ONCV.upf"), "H", pw.ecut, pw.ecut_rho)
    return System(pw, [sH], [0] * len(sym), pos, 1, np.array([2.0]), "TETER93", "h2_upf_ref")


def h2_upf_small(ecut=30.0, nbands=4):
    """The tests/H2_upf structure and UPF species at a lower cutoff with
    several occupied bands: a small case for per-function parity tests."""
    sym, pos = read_xyz(os.path.join(GOLDEN, "structures", "h2_upf.xyz"))
    pw = PWSetup(Cell.cubic(15.1178), ecut)
    sH = psp.UpfSpecies(os.path.join(GOLDEN, "psp", "H_ONCV.upf"), "H", pw.ecut, pw.ecut_rho)
    return System(pw, [sH], [0] * len(sym), pos, nbands, np.full(nbands, 2.0), "PZ", "h2_upf_small")


class SplitMix64:
    """Portable seeded generator: SplitMix64 -> 53-bit uniforms in [0, 1)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        return (z >> 11) * (1.0 / (1 << 53))

    def uniform(self, n):
        return np.array([self.next() for _ in range(n)])


def water_box(lattice, box_bohr, ecut, xc="PBE", seed=2024, jitter=0.3, cell=None):
    """Seeded water box: O on an m1 x m2 x m3 lattice with U(-jitter, jitter)
    displacement, rigid monomer (O-H 0.9572 A, HOH 104.52 deg) with a uniform
    random orientation (random unit quaternion).  `cell` (lattice vectors as
    columns, bohr) replaces the cubic box; the O lattice is then laid out in
    fractional coordinates."""
    rng = SplitMix64(seed)
    m1, m2, m3 = lattice
    A = np.eye(3) * box_bohr if cell is None else np.asarray(cell, np.float64)
    L = 1.0
    roh = 0.9572 * ANG2BOHR
    th = math.radians(104.52) / 2
    h1 = np.array([roh * math.sin(th), 0.0, roh * math.cos(th)])
    h2 = np.array([-roh * math.sin(th), 0.0, roh * math.cos(th)])
    O, H = [], []
    for i in range(m1):
        for j in range(m2):
            for k in range(m3):
                o = A @ ((np.array([i / m1, j / m2, k / m3]) + 0.5 / np.array([m1, m2, m3])) * L)
                o = o + (rng.uniform(3) * 2.0 - 1.0) * jitter
                u1, u2, u3 = rng.uniform(3)
                q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                              math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
                x, y, z, w = q
                R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                              [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                              [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
                O.append(o)
                H.append(o + R @ h1)
                H.append(o + R @ h2)
    nmol = m1 * m2 * m3
    pw = PWSetup(Cell(A), ecut)
    species = [psp.gth("O", xc), psp.gth("H", xc)]
    tau = np.array(O + H)
    sp = [0] * nmol + [1] * (2 * nmol)
    nb = 4 * nmol
    return System(pw, species, sp, tau, nb, np.full(nb, 2.0), xc, f"h2o_{nmol}_{xc.lower()}")


def si64(ecut=30.0, seed=7, sigma=0.05, xc="PZ"):
    """2x2x2 conventional diamond Si (a = 5.43 A), Gaussian displacements."""
    return si64_like(cells=2, ecut=ecut, seed=seed, sigma=sigma, xc=xc)


def si64_like(cells=2, ecut=30.0, seed=7, sigma=0.05, xc="PZ"):
    """cells^3 conventional diamond cells of Si, Gaussian displacements."""
    rng = SplitMix64(seed)
    a = 5.43 * ANG2BOHR
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    pos = []
    for i in range(cells):
        for j in range(cells):
            for k in range(cells):
                for b in basis:
                    pos.append((np.array([i, j, k]) + b) * a)
    pos = np.array(pos)
    u = rng.uniform(2 * pos.size).reshape(-1, 2)
    gauss = np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(2 * math.pi * u[:, 1])
    pos = pos + sigma * gauss.reshape(pos.shape)
    pw = PWSetup(Cell.cubic(cells * a), ecut)
    nb = 2 * len(pos)
    return System(pw, [psp.gth("Si", "LDA")], [0] * len(pos), pos, nb, np.full(nb, 2.0), xc, f"si{len(pos)}")


CONFIGS = {
    "h2o_32_pbe": lambda: water_box((4, 4, 2), 18.63, 70.0, "PBE"),
    "h2o_64_pbe": lambda: water_box((4, 4, 4), 23.47, 80.0, "PBE"),
    "h2o_128_pbe": lambda: water_box((8, 4, 4), 29.57, 80.0, "PBE"),
    "h2o_1_pz": lambda: water_box((1, 1, 1), 18.9, 70.0, "PZ", jitter=0.0),
    "si64": lambda: si64(),
}
