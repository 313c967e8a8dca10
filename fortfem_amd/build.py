"""Build recipe of the fortfem_amd native code (in-tree, for gfx950).

* ``libfortfem_hip.so`` -- HIP kernels + C-ABI (``make -C fortfem_amd/csrc``)
* ``fortran/libfortfem_hip_f.so`` -- Fortran drop-in modules over the C-ABI
  (AMD flang), when flang is available
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_hip(jobs: int = 8) -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"], check=True)
    return os.path.join(HERE, "libfortfem_hip.so")


def flang() -> str | None:
    for cand in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang"), shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran() -> str | None:
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    fc = flang()
    if fc is None:
        return None
    subprocess.run(["make", "-s", "-C", fdir, f"FLANG={fc}"], check=True)
    return os.path.join(fdir, "libfortfem_hip_f.so")


def build_all() -> None:
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_all()
