! fortfem_hip_capi -- ISO_C_BINDING interface to libfortfem_hip.so.
!
! One interface per extern "C" entry point of include/fortfem_hip.h, with the
! same argument order.  Handles are type(c_ptr); arrays are passed by address
! (Fortran column-major storage is exactly the layout the C-ABI documents:
! vertices(3,nv), tetrahedra(4,T), col_ptr(n+1), ...).
module fortfem_hip_capi
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, &
        c_char, c_null_ptr, c_double_complex
    implicit none
    private

    integer(c_int), parameter, public :: FFH_OK = 0
    integer(c_int), parameter, public :: FFH_ERR_INVALID = -1
    integer(c_int), parameter, public :: FFH_ERR_CAPACITY = -2
    integer(c_int), parameter, public :: FFH_ERR_HANDLE = -3
    integer(c_int), parameter, public :: FFH_ERR_DEVICE = -4
    integer(c_int), parameter, public :: FFH_ERR_ADMISSION = -5
    integer(c_int), parameter, public :: FFH_FAMILY_H1 = 1
    integer(c_int), parameter, public :: FFH_FAMILY_NEDELEC = 2
    integer(c_int), parameter, public :: FFH_FAMILY_P1_2D = 3
    integer(c_int), parameter, public :: FFH_COEFF_SCALAR = 0
    integer(c_int), parameter, public :: FFH_COEFF_CELL = 1
    integer(c_int), parameter, public :: FFH_COEFF_CELL_TENSOR = 2
    integer(c_int), parameter, public :: FFH_COEFF_POINT_TENSOR = 3
    integer(c_int), parameter, public :: FFH_SOURCE_CONSTANT = 0
    integer(c_int), parameter, public :: FFH_SOURCE_SIN_PRODUCT = 1
    integer(c_int), parameter, public :: FFH_SOURCE_SAMPLES = 2
    integer(c_int), parameter, public :: FFH_SOURCE_DOF_VECTOR = 3
    integer(c_int), parameter, public :: FFH_FP64 = 0
    integer(c_int), parameter, public :: FFH_MIXED = 1
    integer(c_int), parameter, public :: FFH_FP32 = 1
    integer(c_int), parameter, public :: FFH_FULL = 0
    integer(c_int), parameter, public :: FFH_REDUCED = 1

    ! ffh_solver_options (mirrors solver_options_t + additions)
    type, bind(C), public :: ffh_solver_options
        integer(c_int) :: method = 1
        integer(c_int) :: preconditioner = 0
        integer(c_int) :: tolerance_type = 0
        integer(c_int) :: max_iterations = 1000
        real(c_double) :: tolerance = 1.0e-6_c_double
        integer(c_int) :: precision = 0
        integer(c_int) :: breakdown_mode = 0
        real(c_double) :: inner_tolerance = 1.0e-4_c_double
        integer(c_int) :: max_refinements = 50
        integer(c_int) :: check_interval = 0
        integer(c_int) :: use_graph = 0
    end type ffh_solver_options

    ! ffh_solver_stats (mirrors solver_stats_t + additions)
    type, bind(C), public :: ffh_solver_stats
        integer(c_int) :: converged = 0
        integer(c_int) :: iterations = 0
        real(c_double) :: final_residual = 0.0_c_double
        real(c_double) :: solve_time = 0.0_c_double
        integer(c_int) :: refinements = 0
        real(c_double) :: true_residual = 0.0_c_double
        real(c_double) :: initial_residual = 0.0_c_double
    end type ffh_solver_stats

    public :: ffh_init, ffh_finalize, ffh_last_error, ffh_solver_options_default
    public :: ffh_box_mesh_size, ffh_box_mesh, ffh_unit_square_mesh, ffh_tetra_duffy
    public :: ffh_nedelec_dof_map, ffh_nedelec_basis, ffh_nedelec_load_points
    public :: ffh_problem_create, ffh_problem_free, ffh_pattern_build
    public :: ffh_problem_sizes, ffh_global_dofs_download, ffh_assemble, ffh_load
    public :: ffh_dirichlet, ffh_constrained_download, ffh_set_rhs, ffh_rhs_download
    public :: ffh_problem_matrix, ffh_problem_solve, ffh_matrix_upload
    public :: ffh_matrix_free, ffh_matrix_download, ffh_spmv, ffh_solve_sparse
    public :: ffh_bench_spmv, ffh_set_profiling, ffh_last_solve_profile
    public :: ffh_spmv_jvp, ffh_spmv_vjp, ffh_solve_sparse_jvp, ffh_solve_sparse_vjp
    public :: ffh_assemble_jvp, ffh_assemble_vjp, ffh_nedelec_pml_values
    public :: ffh_nedelec_pml_values_jvp, ffh_nedelec_pml_values_vjp

    interface
        integer(c_int) function ffh_init(device, ctx) bind(C, name="ffh_init")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function ffh_init

        integer(c_int) function ffh_finalize(ctx) bind(C, name="ffh_finalize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function ffh_finalize

        type(c_ptr) function ffh_last_error(ctx) bind(C, name="ffh_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function ffh_last_error

        integer(c_int) function ffh_solver_options_default(opts) &
                bind(C, name="ffh_solver_options_default")
            import :: c_int, ffh_solver_options
            type(ffh_solver_options), intent(out) :: opts
        end function ffh_solver_options_default

        integer(c_int) function ffh_box_mesh_size(counts, nvert, ntet) &
                bind(C, name="ffh_box_mesh_size")
            import :: c_int
            integer(c_int), intent(in) :: counts(3)
            integer(c_int), intent(out) :: nvert, ntet
        end function ffh_box_mesh_size

        integer(c_int) function ffh_box_mesh(bounds, counts, vertices, &
                tetrahedra) bind(C, name="ffh_box_mesh")
            import :: c_int, c_double
            real(c_double), intent(in) :: bounds(3, 2)
            integer(c_int), intent(in) :: counts(3)
            real(c_double), intent(out) :: vertices(*)
            integer(c_int), intent(out) :: tetrahedra(*)
        end function ffh_box_mesh

        integer(c_int) function ffh_unit_square_mesh(n, vertices, triangles, &
                is_boundary_vertex) bind(C, name="ffh_unit_square_mesh")
            import :: c_int, c_double
            integer(c_int), value :: n
            real(c_double), intent(out) :: vertices(*)
            integer(c_int), intent(out) :: triangles(*), is_boundary_vertex(*)
        end function ffh_unit_square_mesh

        integer(c_int) function ffh_tetra_duffy(p, nq, x, y, z, w) &
                bind(C, name="ffh_tetra_duffy")
            import :: c_int, c_double
            integer(c_int), value :: p
            integer(c_int), intent(out) :: nq
            real(c_double), intent(out) :: x(*), y(*), z(*), w(*)
        end function ffh_tetra_duffy

        integer(c_int) function ffh_nedelec_dof_map(k, ntet, tets, base, &
                global_dofs, edge_orient, face_perm, ndofs) &
                bind(C, name="ffh_nedelec_dof_map")
            import :: c_int
            integer(c_int), value :: k, ntet, base
            integer(c_int), intent(in) :: tets(*)
            integer(c_int), intent(out) :: global_dofs(*), edge_orient(*), face_perm(*)
            integer(c_int), intent(out) :: ndofs
        end function ffh_nedelec_dof_map

        ! points: c_null_ptr to query nq, else c_loc of a (4, nq) array
        integer(c_int) function ffh_nedelec_load_points(k, nq, points) &
                bind(C, name="ffh_nedelec_load_points")
            import :: c_int, c_ptr
            integer(c_int), value :: k
            integer(c_int), intent(out) :: nq
            type(c_ptr), value :: points
        end function ffh_nedelec_load_points

        integer(c_int) function ffh_nedelec_basis(k, npts, points, values, curls) &
                bind(C, name="ffh_nedelec_basis")
            import :: c_int, c_double
            integer(c_int), value :: k, npts
            real(c_double), intent(in) :: points(*)
            real(c_double), intent(out) :: values(*), curls(*)
        end function ffh_nedelec_basis

        integer(c_int) function ffh_problem_create(ctx, dim, nvert, vertices, &
                ncell, cells, prob) bind(C, name="ffh_problem_create")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: dim, nvert, ncell
            real(c_double), intent(in) :: vertices(*)
            integer(c_int), intent(in) :: cells(*)
            type(c_ptr), intent(out) :: prob
        end function ffh_problem_create

        integer(c_int) function ffh_problem_free(prob) &
                bind(C, name="ffh_problem_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: prob
        end function ffh_problem_free

        integer(c_int) function ffh_pattern_build(prob, family, order, &
                n_global, nnz) bind(C, name="ffh_pattern_build")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: family, order
            integer(c_int), intent(out) :: n_global
            integer(c_int64_t), intent(out) :: nnz
        end function ffh_pattern_build

        ! stretch: complex (3, T) (kind 0) or (3, 3, T) (kind 1); values:
        ! complex (nnz) on the pattern's FULL CSC order
        integer(c_int) function ffh_nedelec_pml_values(prob, kind, stretch, &
                wave_number, values) bind(C, name="ffh_nedelec_pml_values")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: prob
            integer(c_int), value :: kind
            complex(c_double_complex), intent(in) :: stretch(*)
            real(c_double), value :: wave_number
            complex(c_double_complex), intent(out) :: values(*)
        end function ffh_nedelec_pml_values

        integer(c_int) function ffh_nedelec_pml_values_jvp(prob, kind, stretch, &
                stretch_dot, wave_number, wave_number_dot, vertices_dot, values_dot) &
                bind(C, name="ffh_nedelec_pml_values_jvp")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: prob
            integer(c_int), value :: kind
            complex(c_double_complex), intent(in) :: stretch(*), stretch_dot(*)
            real(c_double), value :: wave_number, wave_number_dot
            real(c_double), intent(in) :: vertices_dot(*)
            complex(c_double_complex), intent(out) :: values_dot(*)
        end function ffh_nedelec_pml_values_jvp

        integer(c_int) function ffh_nedelec_pml_values_vjp(prob, kind, stretch, &
                wave_number, values_bar, vertices_bar, stretch_bar, wave_number_bar) &
                bind(C, name="ffh_nedelec_pml_values_vjp")
            import :: c_int, c_ptr, c_double, c_double_complex
            type(c_ptr), value :: prob
            integer(c_int), value :: kind
            complex(c_double_complex), intent(in) :: stretch(*), values_bar(*)
            real(c_double), value :: wave_number
            real(c_double), intent(out) :: vertices_bar(*)
            complex(c_double_complex), intent(out) :: stretch_bar(*)
            real(c_double), intent(out) :: wave_number_bar
        end function ffh_nedelec_pml_values_vjp

        integer(c_int) function ffh_problem_sizes(prob, n_global, nnz, &
                n_free, nnz_free, n_local, n_colors) &
                bind(C, name="ffh_problem_sizes")
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), intent(out) :: n_global, n_free, n_local, n_colors
            integer(c_int64_t), intent(out) :: nnz, nnz_free
        end function ffh_problem_sizes

        integer(c_int) function ffh_global_dofs_download(prob, base, &
                global_dofs) bind(C, name="ffh_global_dofs_download")
            import :: c_int, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: base
            integer(c_int), intent(out) :: global_dofs(*)
        end function ffh_global_dofs_download

        integer(c_int) function ffh_assemble(prob, quadrature_degree, &
                coeff_kind, coeff, mass_coefficient, precision) &
                bind(C, name="ffh_assemble")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: quadrature_degree, coeff_kind, precision
            real(c_double), intent(in) :: coeff(*)
            real(c_double), value :: mass_coefficient
        end function ffh_assemble

        integer(c_int) function ffh_assemble_jvp(prob, quadrature_degree, &
                coeff_kind, coeff, mass_coefficient, vertices_dot, coeff_dot, &
                mass_dot, values_dot) bind(C, name="ffh_assemble_jvp")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: quadrature_degree, coeff_kind
            real(c_double), intent(in) :: coeff(*), vertices_dot(*), coeff_dot(*)
            real(c_double), value :: mass_coefficient, mass_dot
            real(c_double), intent(out) :: values_dot(*)
        end function ffh_assemble_jvp

        integer(c_int) function ffh_assemble_vjp(prob, quadrature_degree, &
                coeff_kind, coeff, mass_coefficient, values_bar, vertices_bar, &
                coeff_bar, mass_bar) bind(C, name="ffh_assemble_vjp")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: quadrature_degree, coeff_kind
            real(c_double), intent(in) :: coeff(*), values_bar(*)
            real(c_double), value :: mass_coefficient
            real(c_double), intent(out) :: vertices_bar(*), coeff_bar(*), mass_bar
        end function ffh_assemble_vjp

        integer(c_int) function ffh_load(prob, quadrature_degree, source_kind, &
                params, nparams) bind(C, name="ffh_load")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: quadrature_degree, source_kind
            real(c_double), intent(in) :: params(*)
            integer(c_int64_t), value :: nparams
        end function ffh_load

        integer(c_int) function ffh_dirichlet(prob, mode, values) &
                bind(C, name="ffh_dirichlet")
            import :: c_int, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: mode
            type(c_ptr), value :: values
        end function ffh_dirichlet

        integer(c_int) function ffh_constrained_download(prob, constrained) &
                bind(C, name="ffh_constrained_download")
            import :: c_int, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), intent(out) :: constrained(*)
        end function ffh_constrained_download

        integer(c_int) function ffh_set_rhs(prob, which, rhs) &
                bind(C, name="ffh_set_rhs")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: which
            real(c_double), intent(in) :: rhs(*)
        end function ffh_set_rhs

        integer(c_int) function ffh_rhs_download(prob, which, rhs) &
                bind(C, name="ffh_rhs_download")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: which
            real(c_double), intent(out) :: rhs(*)
        end function ffh_rhs_download

        integer(c_int) function ffh_problem_matrix(prob, which, mat) &
                bind(C, name="ffh_problem_matrix")
            import :: c_int, c_ptr
            type(c_ptr), value :: prob
            integer(c_int), value :: which
            type(c_ptr), intent(out) :: mat
        end function ffh_problem_matrix

        integer(c_int) function ffh_problem_solve(prob, opts, solution, stats) &
                bind(C, name="ffh_problem_solve")
            import :: c_int, c_double, c_ptr, ffh_solver_options, ffh_solver_stats
            type(c_ptr), value :: prob
            type(ffh_solver_options), intent(in) :: opts
            real(c_double), intent(out) :: solution(*)
            type(ffh_solver_stats), intent(out) :: stats
        end function ffh_problem_solve

        integer(c_int) function ffh_matrix_upload(ctx, n, nnz, base, col_ptr, &
                row_idx, val, mat) bind(C, name="ffh_matrix_upload")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n, base
            integer(c_int64_t), value :: nnz
            integer(c_int), intent(in) :: col_ptr(*), row_idx(*)
            real(c_double), intent(in) :: val(*)
            type(c_ptr), intent(out) :: mat
        end function ffh_matrix_upload

        integer(c_int) function ffh_matrix_free(mat) bind(C, name="ffh_matrix_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: mat
        end function ffh_matrix_free

        integer(c_int) function ffh_matrix_download(mat, base, capacity, &
                col_ptr, row_idx, val, n, nnz) bind(C, name="ffh_matrix_download")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: mat
            integer(c_int), value :: base
            integer(c_int64_t), value :: capacity
            integer(c_int), intent(out) :: col_ptr(*), row_idx(*)
            real(c_double), intent(out) :: val(*)
            integer(c_int), intent(out) :: n
            integer(c_int64_t), intent(out) :: nnz
        end function ffh_matrix_download

        integer(c_int) function ffh_spmv(mat, x, y) bind(C, name="ffh_spmv")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: x(*)
            real(c_double), intent(out) :: y(*)
        end function ffh_spmv

        integer(c_int) function ffh_solve_sparse(mat, b, x, opts, stats) &
                bind(C, name="ffh_solve_sparse")
            import :: c_int, c_double, c_ptr, ffh_solver_options, ffh_solver_stats
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: b(*)
            real(c_double), intent(inout) :: x(*)
            type(ffh_solver_options), intent(in) :: opts
            type(ffh_solver_stats), intent(out) :: stats
        end function ffh_solve_sparse

        integer(c_int) function ffh_spmv_jvp(mat, x, values_dot, x_dot, y_dot) &
                bind(C, name="ffh_spmv_jvp")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: x(*), values_dot(*), x_dot(*)
            real(c_double), intent(out) :: y_dot(*)
        end function ffh_spmv_jvp

        integer(c_int) function ffh_spmv_vjp(mat, x, y_bar, values_bar, x_bar) &
                bind(C, name="ffh_spmv_vjp")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: x(*), y_bar(*)
            real(c_double), intent(out) :: values_bar(*), x_bar(*)
        end function ffh_spmv_vjp

        integer(c_int) function ffh_solve_sparse_jvp(mat, x, values_dot, b_dot, &
                opts, x_dot, stats) bind(C, name="ffh_solve_sparse_jvp")
            import :: c_int, c_double, c_ptr, ffh_solver_options, ffh_solver_stats
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: x(*), values_dot(*), b_dot(*)
            type(ffh_solver_options), intent(in) :: opts
            real(c_double), intent(out) :: x_dot(*)
            type(ffh_solver_stats), intent(out) :: stats
        end function ffh_solve_sparse_jvp

        integer(c_int) function ffh_solve_sparse_vjp(mat, x, x_bar, opts, &
                values_bar, b_bar, stats) bind(C, name="ffh_solve_sparse_vjp")
            import :: c_int, c_double, c_ptr, ffh_solver_options, ffh_solver_stats
            type(c_ptr), value :: mat
            real(c_double), intent(in) :: x(*), x_bar(*)
            type(ffh_solver_options), intent(in) :: opts
            real(c_double), intent(out) :: values_bar(*), b_bar(*)
            type(ffh_solver_stats), intent(out) :: stats
        end function ffh_solve_sparse_vjp

        integer(c_int) function ffh_bench_spmv(mat, precision, reps, avg_ms, &
                bytes) bind(C, name="ffh_bench_spmv")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: mat
            integer(c_int), value :: precision, reps
            real(c_double), intent(out) :: avg_ms, bytes
        end function ffh_bench_spmv

        integer(c_int) function ffh_set_profiling(ctx, enable) &
                bind(C, name="ffh_set_profiling")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: enable
        end function ffh_set_profiling

        integer(c_int) function ffh_last_solve_profile(ctx, which, count, &
                avg_ms) bind(C, name="ffh_last_solve_profile")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: which
            integer(c_int), intent(out) :: count
            real(c_double), intent(out) :: avg_ms
        end function ffh_last_solve_profile
    end interface

end module fortfem_hip_capi
