! fortfem_cpu_standin.f90 -- for building the drop-in library OUTSIDE
! fortfem only.  Inside fortfem, the GPU build renames each replaced CPU
! module M to M_cpu (the sources stay as they are apart from the module name)
! and each drop-in module M forwards the procedures it does not move to the
! GPU to M_cpu, so that a caller of M finds every public name of the
! reference module.  Outside fortfem there are no CPU modules; this file
! provides modules of those names with the same public interfaces (argument
! lists restated from the reference, file:line at each module) so that the
! drop-in modules and programs written against the reference's import lists
! compile and link.  The procedures stop with a message when called: the CPU
! solvers, element-level routines and evaluators are fortfem's, not this
! library's.  The solver option types, solver_options() and the
! preconditioner-name functions are the reference's small data definitions
! and work.

! src/solvers/advanced_solvers.f90:21-68, 87-1000, 1430-1489
module fortfem_advanced_solvers_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use fortsparse, only: sparse_matrix_t => csc_t
    implicit none
    private

    public :: solver_options_t, solver_stats_t
    public :: solve, solve_sparse, solver_options
    public :: cg_solve, pcg_solve, bicgstab_solve, gmres_solve
    public :: cg_solve_jvp, cg_solve_vjp
    public :: pcg_solve_jvp, pcg_solve_vjp
    public :: bicgstab_solve_jvp, bicgstab_solve_vjp
    public :: gmres_solve_jvp, gmres_solve_vjp
    public :: jacobi_preconditioner, ilu_preconditioner, ichol_preconditioner
    public :: ichol_controlled_preconditioner

    ! :32-44
    type :: solver_options_t
        character(len=32) :: method = "auto"
        real(dp) :: tolerance = 1.0e-6_dp
        integer :: max_iterations = 1000
        character(len=32) :: preconditioner = "none"
        character(len=32) :: tolerance_type = "relative"
        integer :: restart = 50
        logical :: parallel = .false.
        integer :: num_threads = 1
        integer :: verbosity = 0
        real(dp) :: drop_tolerance = 1.0e-4_dp
        integer :: fill_level = 0
    end type solver_options_t

    ! :47-57
    type :: solver_stats_t
        logical :: converged = .false.
        integer :: iterations = 0
        real(dp) :: final_residual = 0.0_dp
        real(dp) :: solve_time = 0.0_dp
        integer :: memory_usage = 0
        character(len=32) :: method_used = ""
        integer :: restarts = 0
        real(dp) :: parallel_efficiency = 0.0_dp
        real(dp) :: condition_estimate = 0.0_dp
    end type solver_stats_t

contains

    ! :1430-1465: the defaults of the type, then the present arguments
    function solver_options(method, tolerance, max_iterations, preconditioner, &
            tolerance_type, restart, parallel, num_threads, &
            verbosity, drop_tolerance, fill_level) result(opts)
        character(len=*), intent(in), optional :: method, preconditioner, tolerance_type
        real(dp), intent(in), optional :: tolerance, drop_tolerance
        integer, intent(in), optional :: max_iterations, restart, num_threads, &
            verbosity, fill_level
        logical, intent(in), optional :: parallel
        type(solver_options_t) :: opts

        if (present(method)) opts%method = method
        if (present(tolerance)) opts%tolerance = tolerance
        if (present(max_iterations)) opts%max_iterations = max_iterations
        if (present(preconditioner)) opts%preconditioner = preconditioner
        if (present(tolerance_type)) opts%tolerance_type = tolerance_type
        if (present(restart)) opts%restart = restart
        if (present(parallel)) opts%parallel = parallel
        if (present(num_threads)) opts%num_threads = num_threads
        if (present(verbosity)) opts%verbosity = verbosity
        if (present(drop_tolerance)) opts%drop_tolerance = drop_tolerance
        if (present(fill_level)) opts%fill_level = fill_level
    end function solver_options

    ! :1469-1489
    function jacobi_preconditioner() result(precond_name)
        character(len=32) :: precond_name
        precond_name = "jacobi"
    end function jacobi_preconditioner

    function ilu_preconditioner(drop_tol, fill_level) result(precond_name)
        real(dp), intent(in), optional :: drop_tol
        integer, intent(in), optional :: fill_level
        character(len=32) :: precond_name
        precond_name = "ilu"
    end function ilu_preconditioner

    function ichol_preconditioner() result(precond_name)
        character(len=32) :: precond_name
        precond_name = "ichol"
    end function ichol_preconditioner

    function ichol_controlled_preconditioner() result(precond_name)
        character(len=32) :: precond_name
        precond_name = "ichol_controlled"
    end function ichol_controlled_preconditioner

    subroutine solve(A, b, x, opts, stats)
        real(dp), intent(in) :: A(:, :), b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "solve: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine solve

    subroutine solve_sparse(A_sparse, b, x, opts, stats)
        type(sparse_matrix_t), intent(in) :: A_sparse
        real(dp), intent(in) :: b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "solve_sparse: method '" // trim(opts%method) // "' with preconditioner '" // &
            trim(opts%preconditioner) // "' runs on fortfem's CPU solvers, which this build of " // &
            "the drop-in library is not linked with"
    end subroutine solve_sparse

    subroutine cg_solve(A, b, x, opts, stats)
        real(dp), intent(in) :: A(:, :), b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "cg_solve: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine cg_solve

    subroutine pcg_solve(A, b, x, opts, stats)
        real(dp), intent(in) :: A(:, :), b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "pcg_solve: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine pcg_solve

    subroutine bicgstab_solve(A, b, x, opts, stats)
        real(dp), intent(in) :: A(:, :), b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "bicgstab_solve: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine bicgstab_solve

    subroutine gmres_solve(A, b, x, opts, stats)
        real(dp), intent(in) :: A(:, :), b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats
        error stop "gmres_solve: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine gmres_solve

    subroutine cg_solve_jvp(  A, b, x, opts, A_dot, b_dot, x_dot, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), A_dot(:, :), b_dot(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: x_dot(:)
        integer, intent(out) :: status
        error stop "cg_solve_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine cg_solve_jvp

    subroutine cg_solve_vjp(  A, b, x, opts, x_bar, A_bar, b_bar, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), x_bar(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: A_bar(:, :), b_bar(:)
        integer, intent(out) :: status
        error stop "cg_solve_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine cg_solve_vjp

    subroutine pcg_solve_jvp(  A, b, x, opts, A_dot, b_dot, x_dot, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), A_dot(:, :), b_dot(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: x_dot(:)
        integer, intent(out) :: status
        error stop "pcg_solve_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine pcg_solve_jvp

    subroutine pcg_solve_vjp(  A, b, x, opts, x_bar, A_bar, b_bar, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), x_bar(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: A_bar(:, :), b_bar(:)
        integer, intent(out) :: status
        error stop "pcg_solve_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine pcg_solve_vjp

    subroutine bicgstab_solve_jvp(  A, b, x, opts, A_dot, b_dot, x_dot, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), A_dot(:, :), b_dot(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: x_dot(:)
        integer, intent(out) :: status
        error stop "bicgstab_solve_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine bicgstab_solve_jvp

    subroutine bicgstab_solve_vjp(  A, b, x, opts, x_bar, A_bar, b_bar, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), x_bar(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: A_bar(:, :), b_bar(:)
        integer, intent(out) :: status
        error stop "bicgstab_solve_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine bicgstab_solve_vjp

    subroutine gmres_solve_jvp(  A, b, x, opts, A_dot, b_dot, x_dot, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), A_dot(:, :), b_dot(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: x_dot(:)
        integer, intent(out) :: status
        error stop "gmres_solve_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine gmres_solve_jvp

    subroutine gmres_solve_vjp(  A, b, x, opts, x_bar, A_bar, b_bar, status)
        real(dp), intent(in) :: A(:, :), b(:), x(:), x_bar(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: A_bar(:, :), b_bar(:)
        integer, intent(out) :: status
        error stop "gmres_solve_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine gmres_solve_vjp
end module fortfem_advanced_solvers_cpu

! src/solvers/fortfem_sparse_matrix.f90:16-44
module fortfem_sparse_matrix_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use fortsparse, only: sparse_matrix_t => csc_t
    implicit none
    private

    public :: sparse_from_dense

contains

    ! entries with |a| > 1e-14 in column-major order (the reference's
    ! triplets, which csc_from_triplet turns into this CSC: one entry per
    ! (row, column), rows ascending)
    subroutine sparse_from_dense(dense_matrix, sparse_matrix)
        real(dp), intent(in) :: dense_matrix(:, :)
        type(sparse_matrix_t), intent(out) :: sparse_matrix

        integer :: column, row, k

        sparse_matrix%nrow = size(dense_matrix, 1)
        sparse_matrix%ncol = size(dense_matrix, 2)
        sparse_matrix%nnz = count(abs(dense_matrix) > 1.0e-14_dp)
        allocate(sparse_matrix%col_ptr(sparse_matrix%ncol + 1))
        allocate(sparse_matrix%row_idx(sparse_matrix%nnz), sparse_matrix%val(sparse_matrix%nnz))
        k = 0
        sparse_matrix%col_ptr(1) = 1
        do column = 1, sparse_matrix%ncol
            do row = 1, sparse_matrix%nrow
                if (abs(dense_matrix(row, column)) <= 1.0e-14_dp) cycle
                k = k + 1
                sparse_matrix%row_idx(k) = row
                sparse_matrix%val(k) = dense_matrix(row, column)
            end do
            sparse_matrix%col_ptr(column + 1) = k + 1
        end do
    end subroutine sparse_from_dense

end module fortfem_sparse_matrix_cpu

! src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:21-27, 40-275, 530-703
module fortfem_assembly_tetra_lagrange_arbitrary_order_3d_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use fortsparse, only: fortsparse_status_t
    implicit none
    private

    public :: assemble_tetra_lagrange_stiffness_element
    public :: assemble_tetra_lagrange_stiffness_element_jvp
    public :: assemble_tetra_lagrange_stiffness_element_vjp
    public :: assemble_tetra_lagrange_scalar_load_samples_jvp
    public :: assemble_tetra_lagrange_scalar_load_samples_vjp

contains

    subroutine assemble_tetra_lagrange_stiffness_element(  vertices, degree, &
        quadrature_degree, matrix, status,  stiffness_coefficient, mass_coefficient)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: degree, quadrature_degree
        real(dp), allocatable, intent(out) :: matrix(:, :)
        integer, intent(out) :: status
        real(dp), intent(in), optional :: stiffness_coefficient
        real(dp), intent(in), optional :: mass_coefficient
        error stop "assemble_tetra_lagrange_stiffness_element: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_lagrange_stiffness_element

    subroutine assemble_tetra_lagrange_stiffness_element_jvp(  vertices, degree, &
        quadrature_degree, stiffness_coefficient,  mass_coefficient, vertices_dot, &
        stiffness_coefficient_dot,  mass_coefficient_dot, matrix_dot, status)
        real(dp), intent(in) :: vertices(3, 4), vertices_dot(3, 4)
        integer, intent(in) :: degree, quadrature_degree
        real(dp), intent(in) :: stiffness_coefficient, mass_coefficient
        real(dp), intent(in) :: stiffness_coefficient_dot
        real(dp), intent(in) :: mass_coefficient_dot
        real(dp), allocatable, intent(out) :: matrix_dot(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_lagrange_stiffness_element_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_lagrange_stiffness_element_jvp

    subroutine assemble_tetra_lagrange_stiffness_element_vjp(  vertices, degree, &
        quadrature_degree, stiffness_coefficient,  mass_coefficient, matrix_bar, vertices_bar, &
        stiffness_coefficient_bar, mass_coefficient_bar, status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: degree, quadrature_degree
        real(dp), intent(in) :: stiffness_coefficient, mass_coefficient
        real(dp), intent(in) :: matrix_bar(:, :)
        real(dp), intent(out) :: vertices_bar(3, 4)
        real(dp), intent(out) :: stiffness_coefficient_bar
        real(dp), intent(out) :: mass_coefficient_bar
        integer, intent(out) :: status
        error stop "assemble_tetra_lagrange_stiffness_element_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_lagrange_stiffness_element_vjp

    subroutine assemble_tetra_lagrange_scalar_load_samples_jvp(  mesh_vertices, tetrahedra, &
        degree, quadrature_degree,  source_values, source_gradients, mesh_vertices_dot, &
        source_parameter_dot, right_hand_side_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        real(dp), intent(in) :: source_values(:, :)
        real(dp), intent(in) :: source_gradients(:, :, :)
        real(dp), intent(in) :: source_parameter_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree, quadrature_degree
        real(dp), allocatable, intent(out) :: right_hand_side_dot(:)
        type(fortsparse_status_t), intent(out) :: status
        error stop "assemble_tetra_lagrange_scalar_load_samples_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_lagrange_scalar_load_samples_jvp

    subroutine assemble_tetra_lagrange_scalar_load_samples_vjp(  mesh_vertices, tetrahedra, &
        degree, quadrature_degree,  source_values, source_gradients, right_hand_side_bar, &
        mesh_vertices_bar, source_values_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :), source_values(:, :)
        real(dp), intent(in) :: source_gradients(:, :, :)
        real(dp), intent(in) :: right_hand_side_bar(:)
        integer, intent(in) :: tetrahedra(:, :), degree, quadrature_degree
        real(dp), intent(out) :: mesh_vertices_bar(:, :)
        real(dp), intent(out) :: source_values_bar(:, :)
        type(fortsparse_status_t), intent(out) :: status
        error stop "assemble_tetra_lagrange_scalar_load_samples_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_lagrange_scalar_load_samples_vjp
end module fortfem_assembly_tetra_lagrange_arbitrary_order_3d_cpu

! src/assembly/assembly_tetra_nedelec_3d.f90:33-41, 51-53 (element-level and
! sampled-load routines)
module fortfem_assembly_tetra_nedelec_3d_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use fortsparse, only: fortsparse_status_t
    implicit none
    private

    public :: assemble_tetra_nedelec_curl_mass_element
    public :: assemble_tetra_nedelec_curl_mass_element_jvp
    public :: assemble_tetra_nedelec_curl_mass_element_vjp
    public :: assemble_tetra_nedelec_pml_element
    public :: assemble_tetra_nedelec_pml_element_jvp
    public :: assemble_tetra_nedelec_pml_element_vjp
    public :: assemble_tetra_nedelec_curvilinear_pml_element
    public :: assemble_tetra_nedelec_curvilinear_pml_element_jvp
    public :: assemble_tetra_nedelec_curvilinear_pml_element_vjp
    public :: assemble_tetra_nedelec_vector_load_samples
    public :: assemble_tetra_nedelec_vector_load_samples_jvp
    public :: assemble_tetra_nedelec_vector_load_samples_vjp

contains

    subroutine assemble_tetra_nedelec_curl_mass_element(  vertices, order, &
        quadrature_degree, matrix, status,  curl_coefficient, mass_coefficient)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        real(dp), allocatable, intent(out) :: matrix(:, :)
        integer, intent(out) :: status
        real(dp), intent(in), optional :: curl_coefficient, mass_coefficient
        error stop "assemble_tetra_nedelec_curl_mass_element: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curl_mass_element

    subroutine assemble_tetra_nedelec_curl_mass_element_jvp(  vertices, order, &
        quadrature_degree, curl_coefficient,  mass_coefficient, vertices_dot, &
        curl_coefficient_dot,  mass_coefficient_dot, matrix_dot, status)
        real(dp), intent(in) :: vertices(3, 4), vertices_dot(3, 4)
        integer, intent(in) :: order, quadrature_degree
        real(dp), intent(in) :: curl_coefficient, mass_coefficient
        real(dp), intent(in) :: curl_coefficient_dot, mass_coefficient_dot
        real(dp), allocatable, intent(out) :: matrix_dot(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_curl_mass_element_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curl_mass_element_jvp

    subroutine assemble_tetra_nedelec_curl_mass_element_vjp(  vertices, order, &
        quadrature_degree, curl_coefficient,  mass_coefficient, matrix_bar, vertices_bar, &
        curl_coefficient_bar,  mass_coefficient_bar, status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        real(dp), intent(in) :: curl_coefficient, mass_coefficient
        real(dp), intent(in) :: matrix_bar(:, :)
        real(dp), intent(out) :: vertices_bar(3, 4)
        real(dp), intent(out) :: curl_coefficient_bar, mass_coefficient_bar
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_curl_mass_element_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curl_mass_element_vjp

    subroutine assemble_tetra_nedelec_pml_element(  vertices, order, quadrature_degree, &
        stretch, wave_number, matrix,  status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3)
        real(dp), intent(in) :: wave_number
        complex(dp), allocatable, intent(out) :: matrix(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_pml_element: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_pml_element

    subroutine assemble_tetra_nedelec_pml_element_jvp(  vertices, order, quadrature_degree, &
        stretch, wave_number,  vertices_dot, stretch_dot, wave_number_dot, matrix_dot, status)
        real(dp), intent(in) :: vertices(3, 4), vertices_dot(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3), stretch_dot(3)
        real(dp), intent(in) :: wave_number, wave_number_dot
        complex(dp), allocatable, intent(out) :: matrix_dot(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_pml_element_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_pml_element_jvp

    subroutine assemble_tetra_nedelec_pml_element_vjp(  vertices, order, quadrature_degree, &
        stretch, wave_number,  matrix_bar, vertices_bar, stretch_bar, wave_number_bar, status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3), matrix_bar(:, :)
        real(dp), intent(in) :: wave_number
        real(dp), intent(out) :: vertices_bar(3, 4), wave_number_bar
        complex(dp), intent(out) :: stretch_bar(3)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_pml_element_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_pml_element_vjp

    subroutine assemble_tetra_nedelec_curvilinear_pml_element(  vertices, order, &
        quadrature_degree, stretch, wave_number, matrix,  status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3, 3)
        real(dp), intent(in) :: wave_number
        complex(dp), allocatable, intent(out) :: matrix(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_curvilinear_pml_element: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curvilinear_pml_element

    subroutine assemble_tetra_nedelec_curvilinear_pml_element_jvp(  vertices, order, &
        quadrature_degree, stretch, wave_number,  vertices_dot, stretch_dot, wave_number_dot, &
        matrix_dot, status)
        real(dp), intent(in) :: vertices(3, 4), vertices_dot(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3, 3), stretch_dot(3, 3)
        real(dp), intent(in) :: wave_number, wave_number_dot
        complex(dp), allocatable, intent(out) :: matrix_dot(:, :)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_curvilinear_pml_element_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curvilinear_pml_element_jvp

    subroutine assemble_tetra_nedelec_curvilinear_pml_element_vjp(  vertices, order, &
        quadrature_degree, stretch, wave_number,  matrix_bar, vertices_bar, stretch_bar, &
        wave_number_bar, status)
        real(dp), intent(in) :: vertices(3, 4)
        integer, intent(in) :: order, quadrature_degree
        complex(dp), intent(in) :: stretch(3, 3), matrix_bar(:, :)
        real(dp), intent(in) :: wave_number
        real(dp), intent(out) :: vertices_bar(3, 4), wave_number_bar
        complex(dp), intent(out) :: stretch_bar(3, 3)
        integer, intent(out) :: status
        error stop "assemble_tetra_nedelec_curvilinear_pml_element_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_curvilinear_pml_element_vjp

    subroutine assemble_tetra_nedelec_vector_load_samples(  mesh_vertices, tetrahedra, &
        order, quadrature_degree,  source_values, vector, status)
        real(dp), intent(in) :: mesh_vertices(:, :), source_values(:, :, :)
        integer, intent(in) :: tetrahedra(:, :), order, quadrature_degree
        real(dp), allocatable, intent(out) :: vector(:)
        type(fortsparse_status_t), intent(out) :: status
        error stop "assemble_tetra_nedelec_vector_load_samples: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_vector_load_samples

    subroutine assemble_tetra_nedelec_vector_load_samples_jvp(  mesh_vertices, tetrahedra, &
        order, quadrature_degree,  source_values, source_gradients, mesh_vertices_dot, &
        source_parameter_dot, vector_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), order, quadrature_degree
        real(dp), intent(in) :: source_values(:, :, :)
        real(dp), intent(in) :: source_gradients(:, :, :, :)
        real(dp), intent(in) :: source_parameter_dot(:, :, :)
        real(dp), allocatable, intent(out) :: vector_dot(:)
        type(fortsparse_status_t), intent(out) :: status
        error stop "assemble_tetra_nedelec_vector_load_samples_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_vector_load_samples_jvp

    subroutine assemble_tetra_nedelec_vector_load_samples_vjp(  mesh_vertices, tetrahedra, &
        order, quadrature_degree,  source_values, source_gradients, vector_bar, &
        mesh_vertices_bar,  source_values_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :), source_values(:, :, :)
        integer, intent(in) :: tetrahedra(:, :), order, quadrature_degree
        real(dp), intent(in) :: source_gradients(:, :, :, :), vector_bar(:)
        real(dp), intent(out) :: mesh_vertices_bar(:, :)
        real(dp), intent(out) :: source_values_bar(:, :, :)
        type(fortsparse_status_t), intent(out) :: status
        error stop "assemble_tetra_nedelec_vector_load_samples_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine assemble_tetra_nedelec_vector_load_samples_vjp
end module fortfem_assembly_tetra_nedelec_3d_cpu

! src/solvers/fortfem_tetra_lagrange_solver_3d.f90:23-46 (solution evaluators)
module fortfem_tetra_lagrange_solver_3d_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    implicit none
    private

    public :: assignment(=)
    public :: evaluate_tetra_lagrange_solution
    public :: evaluate_tetra_lagrange_solution_jvp
    public :: evaluate_tetra_lagrange_solution_vjp
    public :: evaluate_tetra_lagrange_solution_at_point
    public :: evaluate_tetra_lagrange_solution_at_point_jvp
    public :: evaluate_tetra_lagrange_solution_at_point_vjp
    public :: evaluate_tetra_lagrange_solution_prepared
    public :: initialize_tetra_lagrange_solution_evaluator
    public :: tetra_lagrange_solution_evaluator_t

    ! the prepared evaluator (its contents are the CPU module's)
    type :: tetra_lagrange_solution_evaluator_t
        integer, allocatable :: global_dofs(:, :)
        real(dp), allocatable :: basis_gradients(:, :), basis_values(:)
        real(dp), allocatable :: inverse_jacobians(:, :, :)
        integer :: global_dof_count = 0
    end type tetra_lagrange_solution_evaluator_t

    interface assignment(=)
        module procedure assign_evaluator
    end interface

contains

    subroutine assign_evaluator(left, right)
        type(tetra_lagrange_solution_evaluator_t), intent(out) :: left
        type(tetra_lagrange_solution_evaluator_t), intent(in) :: right

        if (allocated(right%global_dofs)) left%global_dofs = right%global_dofs
        if (allocated(right%basis_gradients)) left%basis_gradients = right%basis_gradients
        if (allocated(right%basis_values)) left%basis_values = right%basis_values
        if (allocated(right%inverse_jacobians)) left%inverse_jacobians = right%inverse_jacobians
        left%global_dof_count = right%global_dof_count
    end subroutine assign_evaluator

    subroutine evaluate_tetra_lagrange_solution(  vertices, tetrahedra, degree, solution, &
        tetrahedron,  reference_point, value, gradient, status)
        real(dp), intent(in) :: vertices(:, :), solution(:)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(in) :: reference_point(3)
        real(dp), intent(out) :: value, gradient(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution

    subroutine evaluate_tetra_lagrange_solution_jvp(  vertices, tetrahedra, degree, &
        solution, tetrahedron,  reference_point, vertices_dot, solution_dot, value_dot, &
        gradient_dot, status, reference_point_dot)
        real(dp), intent(in) :: vertices(:, :), solution(:)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(in) :: reference_point(3)
        real(dp), intent(in), optional :: reference_point_dot(3)
        real(dp), intent(in) :: vertices_dot(:, :), solution_dot(:)
        real(dp), intent(out) :: value_dot, gradient_dot(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_jvp

    subroutine evaluate_tetra_lagrange_solution_vjp(  vertices, tetrahedra, degree, &
        solution, tetrahedron,  reference_point, value_bar, gradient_bar, vertices_bar, &
        solution_bar, status, reference_point_bar)
        real(dp), intent(in) :: vertices(:, :), solution(:)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(in) :: reference_point(3), value_bar, gradient_bar(3)
        real(dp), intent(out) :: vertices_bar(:, :), solution_bar(:)
        real(dp), intent(out), optional :: reference_point_bar(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_vjp

    subroutine evaluate_tetra_lagrange_solution_at_point(  vertices, tetrahedra, degree, &
        solution, tetrahedron, point,  value, gradient, status)
        real(dp), intent(in) :: vertices(:, :), solution(:), point(3)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(out) :: value, gradient(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_at_point: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_at_point

    subroutine evaluate_tetra_lagrange_solution_at_point_jvp(  vertices, tetrahedra, degree, &
        solution, tetrahedron, point,  vertices_dot, solution_dot, point_dot, value_dot, &
        gradient_dot,  status)
        real(dp), intent(in) :: vertices(:, :), solution(:), point(3)
        real(dp), intent(in) :: vertices_dot(:, :), solution_dot(:), point_dot(3)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(out) :: value_dot, gradient_dot(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_at_point_jvp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_at_point_jvp

    subroutine evaluate_tetra_lagrange_solution_at_point_vjp(  vertices, tetrahedra, degree, &
        solution, tetrahedron, point,  value_bar, gradient_bar, vertices_bar, solution_bar, &
        point_bar,  status)
        real(dp), intent(in) :: vertices(:, :), solution(:), point(3)
        integer, intent(in) :: tetrahedra(:, :), degree, tetrahedron
        real(dp), intent(in) :: value_bar, gradient_bar(3)
        real(dp), intent(out) :: vertices_bar(:, :), solution_bar(:)
        real(dp), intent(out) :: point_bar(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_at_point_vjp: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_at_point_vjp

    subroutine evaluate_tetra_lagrange_solution_prepared(  evaluator, solution, tetrahedron, &
        reference_point, value,  gradient, status)
        type(tetra_lagrange_solution_evaluator_t), intent(inout) :: evaluator
        real(dp), intent(in) :: solution(:), reference_point(3)
        integer, intent(in) :: tetrahedron
        real(dp), intent(out) :: value, gradient(3)
        integer, intent(out) :: status
        error stop "evaluate_tetra_lagrange_solution_prepared: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine evaluate_tetra_lagrange_solution_prepared

    subroutine initialize_tetra_lagrange_solution_evaluator(  vertices, tetrahedra, degree, &
        evaluator, status)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree
        type(tetra_lagrange_solution_evaluator_t), intent(out) :: evaluator
        integer, intent(out) :: status
        error stop "initialize_tetra_lagrange_solution_evaluator: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine initialize_tetra_lagrange_solution_evaluator
end module fortfem_tetra_lagrange_solver_3d_cpu

! src/solvers/fortfem_tetra_nedelec_solver_3d.f90:21-22 (the PML solves)
module fortfem_tetra_nedelec_solver_3d_cpu
    use, intrinsic :: iso_fortran_env, only: dp => real64
    implicit none
    private

    public :: solve_tetra_nedelec_pml, solve_tetra_nedelec_curvilinear_pml

contains

    subroutine solve_tetra_nedelec_pml(  vertices, tetrahedra, order, stretch, wave_number, &
        volume_load,  dirichlet_dofs, dirichlet_values, solution, status, &
        boundary_operator_dofs, boundary_operator)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :)
        real(dp), intent(in) :: wave_number
        complex(dp), intent(in) :: volume_load(:)
        integer, intent(in) :: dirichlet_dofs(:)
        complex(dp), intent(in) :: dirichlet_values(:)
        complex(dp), allocatable, intent(out) :: solution(:)
        integer, intent(out) :: status
        integer, intent(in), optional :: boundary_operator_dofs(:)
        complex(dp), intent(in), optional :: boundary_operator(:, :)
        error stop "solve_tetra_nedelec_pml: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine solve_tetra_nedelec_pml

    subroutine solve_tetra_nedelec_curvilinear_pml(  vertices, tetrahedra, order, stretch, &
        wave_number, volume_load,  dirichlet_dofs, dirichlet_values, solution, status, &
        boundary_operator_dofs, boundary_operator)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :, :)
        real(dp), intent(in) :: wave_number
        complex(dp), intent(in) :: volume_load(:)
        integer, intent(in) :: dirichlet_dofs(:)
        complex(dp), intent(in) :: dirichlet_values(:)
        complex(dp), allocatable, intent(out) :: solution(:)
        integer, intent(out) :: status
        integer, intent(in), optional :: boundary_operator_dofs(:)
        complex(dp), intent(in), optional :: boundary_operator(:, :)
        error stop "solve_tetra_nedelec_curvilinear_pml: not available: this build of the drop-in library is not linked with fortfem's CPU modules"
    end subroutine solve_tetra_nedelec_curvilinear_pml
end module fortfem_tetra_nedelec_solver_3d_cpu
