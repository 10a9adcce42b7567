"""Extract the force/gradient fixtures of the reference's recorded runs into
reference_forces.json (numbers only, each with the file and lines it comes
from).  Run once in a checkout that has the reference tree:

    python tests/golden/make_force_fixtures.py /root/reference

* 32 H2O, step-0 forces of the MD run (printed with 4 significant figures
  after the converged initial SCF, GRAD_CUTOFF 1e-6), with the printed
  coordinates (bohr) so that atoms can be matched to the structure file;
* 1 H2O, 2 H2O and Si2 FCC: the first record of the geometry optimisations'
  atom_force.dat (full-precision dE/dR and the energy of the starting
  geometry, which is the wf_opt / single_point geometry);
* Si2 FCC: the final energy table of the single-point run.
"""
import json
import os
import re
import sys


def md_forces(path):
    lines = open(path).read().split("\n")
    start = next(i for i, l in enumerate(lines) if "FORCES (-GRADIENTS)" in l) + 1
    rows = []
    i = start
    while i < len(lines) and re.match(r"\s*\d+\s+[A-Z][a-z]?\s", lines[i]):
        t = lines[i].split()
        rows.append({"sym": t[1], "coord_bohr": [float(x) for x in t[2:5]], "force": [float(x) for x in t[5:8]]})
        i += 1
    return rows, start + 1, i


def first_force_record(path):
    lines = open(path).read().split("\n")
    nat = int(lines[0].split()[0])
    m = re.match(r"\s*ENERGY=\s*(\S+)\s+(\d+)", lines[1])
    grads = []
    for l in lines[2:2 + nat]:
        t = l.split()
        grads.append({"sym": t[0], "grad": [float(x) for x in t[1:4]]})
    return float(m.group(1)), grads, 2 + nat


def energy_table(path):
    txt = open(path).read().split("\n")
    keys = {"TOTAL ENERGY": "etot", "KINETIC ENERGY": "ekin", "ELECTROSTATIC ENERGY": "estat",
            "SELF ENERGY CONTRIBUTION": "eself", "OVERLAP ENERGY CONTRIBUTION": "esr",
            "LOCAL PP ENERGY": "eloc", "NON LOCAL PP ENERGY": "enl", "EXCHANGE CORRELATION ENERGY": "exc"}
    out, lo, hi = {}, None, None
    for i, l in enumerate(txt):
        for k, v in keys.items():
            if re.search(r"#\s+" + k + r"\s+=", l):
                out[v] = float(l.split("=")[1].split()[0])
                lo = i + 1 if lo is None else lo
                hi = i + 1
    return out, lo, hi


def main(ref):
    t = os.path.join(ref, "tests")
    data = {"_source": "Forces and gradients printed by the reference's recorded runs (Hartree, bohr). "
                       "'force' = -dE/dR; 'grad' = dE/dR as atom_force.dat stores it."}
    rows, a, b = md_forces(os.path.join(t, "32_water", "md", "amdkiit.out"))
    data["h2o_32_md_step0"] = {"cite": f"tests/32_water/md/amdkiit.out:{a}-{b}", "grad_cutoff": 1e-6,
                               "printed_digits": 4, "atoms": rows}
    for key, sub in (("h2o_1_geo_step0", "1_water"), ("h2o_2_geo_step0", "2_water"), ("si2_geo_step0", "bulk_si")):
        e, g, n = first_force_record(os.path.join(t, sub, "geo_opt", "atom_force.dat"))
        data[key] = {"cite": f"tests/{sub}/geo_opt/atom_force.dat:1-{n}", "energy": e, "atoms": g}
    tab, a, b = energy_table(os.path.join(t, "bulk_si", "single_point", "amdkiit.out"))
    tab["cite"] = f"tests/bulk_si/single_point/amdkiit.out:{a}-{b}"
    data["si2_single_point"] = tab
    # end of the Si2 geometry optimisation: final energy table and the last
    # LB1 line (ITER, NFN = energy/gradient evaluations)
    path = os.path.join(t, "bulk_si", "geo_opt", "amdkiit.out")
    tab, a, b = energy_table(path)
    lines = open(path).read().split("\n")
    k = max(i for i, l in enumerate(lines) if re.match(r"\s*\d+\s+\d+\s+-?\d\.\d+D[+-]\d+", l))
    it, nfn = (int(v) for v in lines[k].split()[:2])
    tab.update({"cite": f"tests/bulk_si/geo_opt/amdkiit.out:{k + 1},{a}-{b}", "iterations": it, "evaluations": nfn})
    data["si2_geo_final"] = tab
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_forces.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
