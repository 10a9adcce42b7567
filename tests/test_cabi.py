"""CPU checks of the C ABI: every declaration of include/amdk.h is exported
by libamdk.so with the binding's signature table, and argument/device
validation fails cleanly without touching a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from amdkiit_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "amdk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(amdk_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported_and_bound():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared in amdk.h but not exported"
        assert n in L.SIGNATURES, f"{n} missing from the ctypes signature table"


def test_nm_lists_extern_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIBPATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (amdk_[a-z0-9_]+)$", out, flags=re.M))
    assert set(header_functions()) <= exported


def make_desc(prec=L.PREC_FP64, nbmax=4):
    from amdkiit_amd.pwgrid import Cell, PWSetup
    pw = PWSetup(Cell.cubic(8.0), 6.0)
    d = L.Desc()
    d.n1, d.n2, d.n3 = pw.n
    for i in range(3):
        d.a_cell[4 * i] = 8.0
    mp = np.ascontiguousarray(pw.mill_pw, np.int32).reshape(-1)
    mr = np.ascontiguousarray(pw.mill_rho, np.int32).reshape(-1)
    d.ngpw, d.miller_pw = pw.ngpw, L.iptr(mp)
    d.ngrho, d.miller_rho = pw.ngrho, L.iptr(mr)
    d.nbmax, d.precision = nbmax, prec
    return d, (mp, mr)


def create(d):
    lib = L.load()
    h = C.c_void_p()
    rc = lib.amdk_ctx_create(C.byref(d), C.byref(h))
    msg = lib.amdk_last_error(h).decode() if h else ""
    lib.amdk_ctx_destroy(h)
    return rc, msg


def test_bad_arguments_rejected_before_device_use():
    d, keep = make_desc(prec=7)
    rc, msg = create(d)
    assert rc == L.ERR_ARG and "precision" in msg
    d, keep = make_desc(nbmax=0)
    rc, msg = create(d)
    assert rc == L.ERR_ARG and "nbmax" in msg
    lib = L.load()
    assert lib.amdk_vpsi(None, None, 0, None, None, 0) == L.ERR_ARG
    assert lib.amdk_ldc(None) == -1


def test_no_device_is_a_clean_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    d, keep = make_desc()
    rc, msg = create(d)
    assert rc == L.ERR_DEVICE
    assert "device" in msg.lower()


def test_fortran_binding_compiled():
    exe = os.path.join(ROOT, "amdkiit_amd", "fortran", "amdk_smoke")
    if not os.path.exists(exe):
        pytest.skip("amdflang not available; Fortran smoke driver not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present: the GPU test runs the driver")
    except ImportError:
        pass
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    # without a device the context creation must fail cleanly through amdk_check
    assert r.returncode != 0
    assert "amdk_ctx_create" in r.stdout and "amdk error" in r.stdout
