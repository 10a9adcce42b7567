"""The Fortran ISO_C_BINDING smoke driver on the GPU (VPSI closed form and
ortho_gs through the same C ABI the reference's call sites would use)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fortran_smoke_driver():
    exe = os.path.join(ROOT, "amdkiit_amd", "fortran", "amdk_smoke")
    if not os.path.exists(exe):
        pytest.skip("Fortran smoke driver not built (amdflang missing)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
