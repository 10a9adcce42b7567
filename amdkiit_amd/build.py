"""Build libamdk.so (HIP, gfx950) and the Fortran binding module in-tree.

Usage: python -m amdkiit_amd.build [--check]
The library is written next to this file (amdkiit_amd/libamdk.so) so that it
travels with the source tree; nothing is installed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libamdk.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("AMDK_FC", "/opt/rocm/bin/amdflang")

SOURCES = ["amdk_api.hip"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build_lib(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs + [
        "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=True):
    """Compile the ISO_C_BINDING modules (amdk_hip interfaces, amdk_ref with
    the reference's routine names) and the two drivers (host code only)."""
    fdir = os.path.join(HERE, "fortran")
    mods = [os.path.join(fdir, f) for f in ("amdk_hip.f90", "amdk_ref.f90")]
    progs = {"amdk_smoke": "amdk_smoke.f90", "amdk_wfopt": "amdk_wfopt.f90"}
    if not all(os.path.exists(s) for s in mods) or not shutil.which(FC):
        return None
    moddir = os.path.join(fdir, "_mod")
    os.makedirs(moddir, exist_ok=True)
    outs = []
    for exe, src in progs.items():
        out = os.path.join(fdir, exe)
        cmd = [FC, "-O2", "-J", moddir, "-o", out] + mods + [os.path.join(fdir, src)] + [
            "-L" + HERE, "-lamdk", "-Wl,-rpath," + HERE]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        outs.append(out)
    return outs


def build_oracle_ref(verbose=True):
    """Build the reference's own L-BFGS into oracle/_ref (a checker for the
    restated one) when the reference tree is present; never on a box
    without it."""
    ref = os.environ.get("AMDK_REFERENCE", "/root/reference")
    if not os.path.isdir(os.path.join(ref, "source")) or not shutil.which("make"):
        return None
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "REF=" + ref]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return os.path.join(ROOT, "oracle", "_ref", "liblbfgs_ref.so")


def main():
    build_lib()
    build_fortran()
    build_oracle_ref()


if __name__ == "__main__":
    sys.exit(main())
