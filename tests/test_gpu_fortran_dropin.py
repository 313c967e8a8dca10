"""Runs the Fortran drop-in test program (tests/fortran/test_fortfem_hip_dropin.f90)
on the GPU: the reference-named Fortran modules calling the HIP library through
ISO_C_BINDING."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "fortran", "test_fortfem_hip_dropin")


@pytest.mark.gpu
def test_fortran_dropin_program():
    assert os.path.exists(BIN), "build the Fortran drop-in (fortfem_amd/build.py) first"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout
    assert "drop-in Fortran checks passed" in out.stdout
