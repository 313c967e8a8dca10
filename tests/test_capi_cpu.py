"""CPU checks of the C-ABI library and the Python host side (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char \*)\s*(ffh_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    from fortfem_amd import _capi

    lib = _capi.load()
    declared = header_symbols()
    assert len(declared) >= 28
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_capi.EXPORTS) == declared


def test_library_is_gfx950_code_object():
    path = os.path.join(ROOT, "fortfem_amd", "libfortfem_hip.so")
    data = open(path, "rb").read()
    assert b"gfx950" in data


def test_errors_without_device_are_reported_not_faked():
    """With no GPU visible the product path fails loudly (-4), never falls back."""
    import fortfem_amd as ff

    lib = ff.capi.load()
    try:
        import torch  # noqa: F401

        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert lib.ffh_init(0, C.byref(h)) == ff.capi.FFH_ERR_DEVICE
    assert lib.ffh_last_error(None)


def test_host_mesh_entry_points_match_oracle():
    """ffh_box_mesh / ffh_unit_square_mesh are host routines of the library."""
    import fortfem_amd as ff
    import oracle as O

    for counts in ([2, 1, 3], [4, 4, 4]):
        v, t, st = ff.generate_structured_tetra_box_mesh([[-1, 2, 4], [1, 5, 10]], counts)
        vo, to = O.box_mesh([[-1, 2, 4], [1, 5, 10]], counts)
        assert st == 0 and np.array_equal(v, vo) and np.array_equal(t, to)
    v, t, b = ff.unit_square_mesh(9)
    vo, to, bo = O.unit_square_mesh(9)
    assert np.array_equal(v, vo) and np.array_equal(t, to) and np.array_equal(b, bo)
    v, t, st = ff.generate_structured_tetra_box_mesh([[0, 0, 0], [1, 1, 1]], [0, 1, 1])
    assert st != 0


def test_pcg_launch_form_rule():
    """Three launches per iteration in both precisions at every size unless
    the process runs with FFH_PCG_LAUNCHES=2; a host-side rule (no GPU)."""
    import fortfem_amd as ff

    if os.environ.get("FFH_PCG_LAUNCHES"):
        pytest.skip("launch form forced by FFH_PCG_LAUNCHES")
    for n in (274625, (1 << 20) - 1, 1 << 20, 2048383):
        assert ff.pcg_launches_per_iteration(n, "fp64") == 3
        assert ff.pcg_launches_per_iteration(n, "fp32") == 3
    lib = ff.capi.load()
    assert lib.ffh_pcg_launches_per_iteration(-1, ff.capi.FP64) == ff.capi.FFH_ERR_INVALID
    assert lib.ffh_pcg_launches_per_iteration(10, 7) == ff.capi.FFH_ERR_INVALID


def test_solver_options_defaults_mirror_reference():
    """solver_options() defaults (advanced_solvers.f90:1441-1452)."""
    import fortfem_amd as ff

    o = ff.solver_options()
    assert (o.method, o.tolerance, o.max_iterations, o.preconditioner, o.tolerance_type) == \
        ("auto", 1e-6, 1000, "none", "relative")
    c = o.to_c()
    assert c.method == 1 and c.preconditioner == 0 and c.tolerance_type == 0 and c.max_iterations == 1000
    c = ff.solver_options(method="cg", preconditioner="jacobi", tolerance_type="absolute").to_c()
    assert c.method == 0 and c.preconditioner == 1 and c.tolerance_type == 1
    with pytest.raises(TypeError):
        ff.solver_options(bogus=1)


def test_workload_rng_is_reproducible():
    from fortfem_amd import workloads

    a = workloads.lognormal_kappa(1000, 2)
    b = workloads.lognormal_kappa(1000, 2)
    assert np.array_equal(a, b)
    z = workloads.normal(12345, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    u = workloads.uniform(99, 1000)
    assert u.min() >= 0 and u.max() < 1
    # SplitMix64 reference value: first output of seed 0
    assert int(workloads.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


def test_fortran_host_dropin_program():
    """Host-only Fortran drop-in checks (mesh, Duffy rule through the C-ABI);
    also the regression test for the Duffy size query, which once wrote all
    points into a one-element array."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "fortran", "test_fortfem_hip_host")
    if not os.path.exists(exe):
        pytest.skip("Fortran drop-in not built (no flang)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "host Fortran checks passed" in out.stdout


def test_fortran_reference_import_lists_build_and_run():
    """A Fortran program whose modules import exactly what the reference's
    callers import from the replaced modules (fortfem_api.f90:1780-1785,
    fortfem_api_solvers.f90:9-17, fortfem_solvers_laplacian.f90:6-7,
    bem/laplace_fem_bem_coupling_3d.f90:8-9, solvers/tetra_lagrange_state_3d.f90:
    3-6, and every other public name of those modules) builds against the
    drop-in, and the forwarded host helpers (sparse_from_dense, solver_options,
    the preconditioner names) behave as the reference's."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "fortran", "test_reference_imports")
    if not os.path.exists(exe):
        pytest.skip("Fortran drop-in not built (no flang)")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "failures 0" in out.stdout


@pytest.mark.parametrize("case", ["gmres", "bicgstab", "ilu"])
def test_fortran_solve_sparse_refuses_cpu_only_methods(case):
    """solve_sparse with gmres, bicgstab or an ILU preconditioner goes to
    fortfem's CPU solve_sparse (advanced_solvers.f90:172-186); without
    fortfem's CPU modules the call stops naming the method, instead of
    running PCG with status OK."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "fortran", "test_gmres_refusal")
    if not os.path.exists(exe):
        pytest.skip("Fortran drop-in not built (no flang)")
    out = subprocess.run([exe, case], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "solve_sparse returned" not in out.stdout
    text = out.stdout + out.stderr
    assert "runs on fortfem's CPU solvers" in text
    assert (case if case != "ilu" else "'ilu'") in text
