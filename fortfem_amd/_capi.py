"""ctypes binding of libfortfem_hip.so (the C-ABI declared in include/fortfem_hip.h).

The library is built in-tree by ``fortfem_amd.build`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfortfem_hip.so")

FFH_OK = 0
FFH_ERR_INVALID = -1
FFH_ERR_CAPACITY = -2
FFH_ERR_HANDLE = -3
FFH_ERR_DEVICE = -4
FFH_ERR_ADMISSION = -5

FAMILY_H1 = 1
FAMILY_NEDELEC = 2
FAMILY_P1_2D = 3

COEFF_SCALAR = 0
COEFF_CELL = 1
COEFF_CELL_TENSOR = 2
COEFF_POINT_TENSOR = 3

SOURCE_CONSTANT = 0
SOURCE_SIN_PRODUCT = 1
SOURCE_SAMPLES = 2
SOURCE_DOF_VECTOR = 3

FP64 = 0
MIXED = 1
FP32 = 1

FULL = 0
REDUCED = 1

# every symbol include/fortfem_hip.h declares
EXPORTS = (
    "ffh_init", "ffh_finalize", "ffh_last_error", "ffh_solver_options_default",
    "ffh_box_mesh_size", "ffh_box_mesh", "ffh_unit_square_mesh", "ffh_tetra_duffy",
    "ffh_problem_create", "ffh_problem_free", "ffh_pattern_build", "ffh_problem_sizes",
    "ffh_global_dofs_download", "ffh_assemble", "ffh_load", "ffh_dirichlet",
    "ffh_constrained_download", "ffh_set_rhs", "ffh_rhs_download", "ffh_problem_matrix",
    "ffh_problem_solve", "ffh_matrix_upload", "ffh_matrix_free", "ffh_matrix_download",
    "ffh_spmv", "ffh_solve_sparse", "ffh_bench_spmv", "ffh_set_profiling",
    "ffh_last_solve_profile", "ffh_matrix_spmv_variant", "ffh_matrix_spmv_bytes",
    "ffh_pcg_launches_per_iteration", "ffh_nedelec_dof_map", "ffh_nedelec_basis",
    "ffh_nedelec_load_points", "ffh_spmv_jvp", "ffh_spmv_vjp", "ffh_solve_sparse_jvp",
    "ffh_solve_sparse_vjp", "ffh_assemble_jvp", "ffh_assemble_vjp", "ffh_nedelec_pml_values",
    "ffh_nedelec_pml_values_jvp", "ffh_nedelec_pml_values_vjp", "ffh_zmatrix_upload",
    "ffh_zmatrix_solve", "ffh_zmatrix_free",
    "ffh_server_create", "ffh_server_free", "ffh_server_prepare", "ffh_server_submit", "ffh_server_step",
    "ffh_server_poll", "ffh_server_wait", "ffh_server_stats", "ffh_server_progress",
    "ffh_problem_solution", "ffh_server_set_precision", "ffh_server_use_graph",
    "ffh_server_set_tiling", "ffh_server_tiling_stats", "ffh_server_product_forms",
    "ffh_solve_sparse_device_noalloc", "ffh_noalloc_spmv_variant", "ffh_problem_colouring",
)

NOALLOC_MAX_ROWS = 1 << 21


class SolverOptions(C.Structure):
    """ffh_solver_options == solver_options_t (advanced_solvers.f90:32-44) + additions."""

    _fields_ = [
        ("method", C.c_int),
        ("preconditioner", C.c_int),
        ("tolerance_type", C.c_int),
        ("max_iterations", C.c_int),
        ("tolerance", C.c_double),
        ("precision", C.c_int),
        ("breakdown_mode", C.c_int),
        ("inner_tolerance", C.c_double),
        ("max_refinements", C.c_int),
        ("check_interval", C.c_int),
        ("use_graph", C.c_int),
    ]


class SolverStats(C.Structure):
    """ffh_solver_stats == solver_stats_t (advanced_solvers.f90:47-57) + additions."""

    _fields_ = [
        ("converged", C.c_int),
        ("iterations", C.c_int),
        ("final_residual", C.c_double),
        ("solve_time", C.c_double),
        ("refinements", C.c_int),
        ("true_residual", C.c_double),
        ("initial_residual", C.c_double),
    ]


_lib = None

vp = C.c_void_p
ip = C.POINTER(C.c_int)
i64p = C.POINTER(C.c_int64)
dp = C.POINTER(C.c_double)


def load() -> C.CDLL:
    """Load libfortfem_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback for the fortfem_amd path)")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "ffh_init": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "ffh_finalize": (C.c_int, [vp]),
        "ffh_last_error": (C.c_char_p, [vp]),
        "ffh_solver_options_default": (C.c_int, [C.POINTER(SolverOptions)]),
        "ffh_box_mesh_size": (C.c_int, [vp, ip, ip]),
        "ffh_box_mesh": (C.c_int, [vp, vp, vp, vp]),
        "ffh_unit_square_mesh": (C.c_int, [C.c_int, vp, vp, vp]),
        "ffh_tetra_duffy": (C.c_int, [C.c_int, ip, vp, vp, vp, vp]),
        "ffh_problem_create": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int, vp, C.POINTER(vp)]),
        "ffh_problem_free": (C.c_int, [vp]),
        "ffh_pattern_build": (C.c_int, [vp, C.c_int, C.c_int, ip, i64p]),
        "ffh_problem_sizes": (C.c_int, [vp, ip, i64p, ip, i64p, ip, ip]),
        "ffh_global_dofs_download": (C.c_int, [vp, C.c_int, vp]),
        "ffh_assemble": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_double, C.c_int]),
        "ffh_load": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int64]),
        "ffh_dirichlet": (C.c_int, [vp, C.c_int, vp]),
        "ffh_constrained_download": (C.c_int, [vp, vp]),
        "ffh_set_rhs": (C.c_int, [vp, C.c_int, vp]),
        "ffh_rhs_download": (C.c_int, [vp, C.c_int, vp]),
        "ffh_problem_matrix": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "ffh_problem_solve": (C.c_int, [vp, C.POINTER(SolverOptions), vp, C.POINTER(SolverStats)]),
        "ffh_matrix_upload": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int, vp, vp, vp, C.POINTER(vp)]),
        "ffh_matrix_free": (C.c_int, [vp]),
        "ffh_matrix_download": (C.c_int, [vp, C.c_int, C.c_int64, vp, vp, vp, ip, i64p]),
        "ffh_spmv": (C.c_int, [vp, vp, vp]),
        "ffh_solve_sparse": (C.c_int, [vp, vp, vp, C.POINTER(SolverOptions), C.POINTER(SolverStats)]),
        "ffh_bench_spmv": (C.c_int, [vp, C.c_int, C.c_int, dp, dp]),
        "ffh_nedelec_pml_values": (C.c_int, [vp, C.c_int, vp, C.c_double, vp]),
        "ffh_nedelec_pml_values_jvp": (C.c_int, [vp, C.c_int, vp, vp, C.c_double, C.c_double, vp, vp]),
        "ffh_nedelec_pml_values_vjp": (C.c_int, [vp, C.c_int, vp, C.c_double, vp, vp, vp, vp]),
        "ffh_zmatrix_upload": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int, vp, vp, vp, C.POINTER(vp)]),
        "ffh_zmatrix_solve": (C.c_int, [vp, vp, vp, C.c_double, C.c_int, ip, C.POINTER(C.c_double), ip]),
        "ffh_zmatrix_free": (C.c_int, [vp]),
        "ffh_assemble_jvp": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_double, vp, vp, C.c_double, vp]),
        "ffh_assemble_vjp": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_double, vp, vp, vp, dp]),
        "ffh_spmv_jvp": (C.c_int, [vp, vp, vp, vp, vp]),
        "ffh_spmv_vjp": (C.c_int, [vp, vp, vp, vp, vp]),
        "ffh_solve_sparse_jvp": (C.c_int, [vp, vp, vp, vp, C.POINTER(SolverOptions), vp,
                                           C.POINTER(SolverStats)]),
        "ffh_solve_sparse_vjp": (C.c_int, [vp, vp, vp, C.POINTER(SolverOptions), vp, vp,
                                           C.POINTER(SolverStats)]),
        "ffh_set_profiling": (C.c_int, [vp, C.c_int]),
        "ffh_last_solve_profile": (C.c_int, [vp, C.c_int, ip, dp]),
        "ffh_matrix_spmv_variant": (C.c_int, [vp, C.c_int, C.c_int, ip]),
        "ffh_matrix_spmv_bytes": (C.c_int, [vp, C.c_int, dp]),
        "ffh_pcg_launches_per_iteration": (C.c_int, [C.c_int64, C.c_int]),
        "ffh_nedelec_dof_map": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, ip]),
        "ffh_nedelec_basis": (C.c_int, [C.c_int, C.c_int, vp, vp, vp]),
        "ffh_nedelec_load_points": (C.c_int, [C.c_int, ip, vp]),
        "ffh_server_create": (C.c_int, [vp, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]),
        "ffh_server_free": (C.c_int, [vp]),
        "ffh_server_prepare": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, ip, i64p]),
        "ffh_server_submit": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, C.c_double, C.c_int, vp, C.c_double,
                                        C.c_int, i64p]),
        "ffh_server_step": (C.c_int, [vp, ip, ip, ip]),
        "ffh_server_poll": (C.c_int, [vp, C.c_int64, vp, vp, vp, vp]),
        "ffh_server_wait": (C.c_int, [vp, C.c_int64, C.c_int, vp, vp, vp, vp]),
        "ffh_server_stats": (C.c_int, [vp, i64p, i64p]),
        "ffh_server_progress": (C.c_int, [vp, C.c_int64, vp, vp, vp]),
        "ffh_problem_solution": (C.c_int, [vp, vp]),
        "ffh_server_set_precision": (C.c_int, [vp, C.c_int, i64p]),
        "ffh_server_use_graph": (C.c_int, [vp, C.c_int]),
        "ffh_server_set_tiling": (C.c_int, [vp, C.c_int]),
        "ffh_server_tiling_stats": (C.c_int, [vp, ip, ip]),
        "ffh_server_product_forms": (C.c_int, [vp, ip, ip, ip]),
        "ffh_solve_sparse_device_noalloc": (C.c_int, [vp, C.c_int, C.c_int64, vp, vp, vp, vp, vp,
                                                      C.POINTER(SolverOptions), C.POINTER(SolverStats)]),
        "ffh_noalloc_spmv_variant": (C.c_int, [C.c_int, C.c_int64]),
        "ffh_problem_colouring": (C.c_int, [vp, C.c_int, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class FFHError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def check(code: int, ctx=None) -> None:
    if code != FFH_OK:
        msg = load().ffh_last_error(ctx)
        raise FFHError(code, (msg or b"").decode(errors="replace"))
