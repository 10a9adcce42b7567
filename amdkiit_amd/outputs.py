"""The reference's run output files, in its column layouts.

* SP_ENERGY.dat      do_opt.f90:119-124, '(I4,3X,F16.6,8X,F10.1)': step,
                     E_tot, wall time of the step in ms.
* MD_ENERGY.dat      md_tools.f90:171-181, list-directed: step,
                     temperature, E_KS, E_KS + E_kin(ions), cpu seconds.
* MD_TRAJECTORY.xyz  md_tools.f90:152-169: atom count and "MD STEP=" step
                     list-directed, then '(A,6f16.6)' per atom: symbol,
                     position and velocity divided by angs2bohr.
* atom_force.dat, atom_coord.xyz  nuclear_grad.f90:269-312 (PRINT_COORDINATE),
                     list-directed: atom count, "ENERGY=" E step, then per
                     atom the symbol and the forces (Ha/bohr) or the
                     positions times 0.529177.

List-directed records follow the layout gfortran writes (the reference's
recorded files): one leading blank; CHARACTER items as they are; every
numeric item after the first preceded by one blank; INTEGER in 11 columns;
REAL(8) in 25 columns with 17 significant digits, as F (20 columns, then 5
blanks) for 0.1 <= |x| < 1e17 or zero, else as E with a 3-digit exponent.
Symbols are CHARACTER(2), as the reference's element table holds them.
"""
import math
import os

ANGS2BOHR = 1.0 / 0.529177210859   # constants.f90:17
BOHR_COORD_PRINT = 0.529177        # nuclear_grad.f90 (atom_coord.xyz factor)


def _ld_real(x):
    x = float(x)
    ax = abs(x)
    if x == 0.0:
        return f"{'0.0000000000000000':>20}" + " " * 5
    if 0.1 <= ax < 1e17:
        e = math.floor(math.log10(ax))
        s = f"{x:.{max(0, 16 - e)}f}"
        if abs(float(s)) >= 10.0 ** (e + 1):  # rounding carried into a new digit
            s = f"{x:.{max(0, 15 - e)}f}"
        return f"{s:>20}" + " " * 5
    m, ex = f"{x:.16E}".split("E")
    return f"{m}E{int(ex):+04d}".rjust(25)


def _ld_int(i):
    return f"{int(i):>11d}"


def list_directed(*items):
    """One list-directed record (without the newline)."""
    out = " "
    first = True
    for it in items:
        if isinstance(it, str):
            out += it
        else:
            if not first:
                out += " "
            out += _ld_int(it) if isinstance(it, (int,)) and not isinstance(it, bool) else _ld_real(it)
        first = False
    return out


def _sym2(s):
    return f"{s:<2}"[:2]


class RunOutputs:
    """Appends to the reference's output files in `directory` (created if
    needed), as the reference opens them with position="APPEND"."""

    def __init__(self, directory="."):
        self.dir = directory
        os.makedirs(directory, exist_ok=True)

    def _append(self, name, text):
        with open(os.path.join(self.dir, name), "a") as f:
            f.write(text)

    def sp_energy(self, step, etot, ms):
        self._append("SP_ENERGY.dat", f"{step:4d}   {etot:16.6f}        {ms:10.1f}\n")

    def md_energy(self, step, temperature, e_ks, e_cons, cpu_s):
        self._append("MD_ENERGY.dat", list_directed(int(step), temperature, e_ks, e_cons, cpu_s) + "\n")

    def md_trajectory(self, step, symbols, tau_bohr, vel_au):
        lines = [list_directed(len(symbols)), list_directed("MD STEP=", int(step))]
        for s, r, v in zip(symbols, tau_bohr, vel_au):
            vals = [c / ANGS2BOHR for c in r] + [c / ANGS2BOHR for c in v]
            lines.append(_sym2(s) + "".join(f"{c:16.6f}" for c in vals))
        self._append("MD_TRAJECTORY.xyz", "\n".join(lines) + "\n")

    def _frame(self, name, step, energy, symbols, rows):
        lines = [list_directed(len(symbols)), list_directed("ENERGY=", energy, int(step))]
        for s, row in zip(symbols, rows):
            lines.append(list_directed(_sym2(s), *[float(c) for c in row]))
        self._append(name, "\n".join(lines) + "\n")

    def atom_force(self, step, energy, symbols, forces):
        self._frame("atom_force.dat", step, energy, symbols, forces)

    def atom_coord(self, step, energy, symbols, tau_bohr):
        self._frame("atom_coord.xyz", step, energy, symbols, [[c * BOHR_COORD_PRINT for c in r] for r in tau_bohr])
