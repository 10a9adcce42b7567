"""Writes twiddle_const.h: exp(-2 pi i m / L) for the in-register DFT lengths.

Run from the repository root: python amdkiit_amd/csrc/gen_twiddle.py
"""
import math
import os

LENGTHS = [2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20]


def main():
    out = ["// Twiddle constants exp(-2 pi i m / L) for the in-register DFTs (written by",
           "// gen_twiddle.py from the closed form in double precision).",
           "#pragma once", "", "namespace amdk {", "", "template <int L> struct Tw;"]
    for L in LENGTHS:
        c = ", ".join("%.20e" % math.cos(2 * math.pi * m / L) for m in range(L))
        s = ", ".join("%.20e" % (-math.sin(2 * math.pi * m / L)) for m in range(L))
        out += [f"template <> struct Tw<{L}> {{",
                f"  static constexpr double re[{L}] = {{{c}}};",
                f"  static constexpr double im[{L}] = {{{s}}};", "};"]
    out += ["", "}  // namespace amdk"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twiddle_const.h")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
