! Drop-in Fortran modules for fortfem's sparse-assembly + Jacobi-PCG path.
!
! Each module keeps the name, procedure names, argument lists, argument
! meaning and status behaviour of the reference module it replaces
! (lazy-fortran/fortfem); the bodies call the HIP library through
! fortfem_hip_capi.  Callers that `use` these modules compile unchanged.
!
!   fortfem_hip_types            csc_t / fortsparse_status_t layout of the
!                                fortsparse dependency (col_ptr, row_idx 1-based)
!   fortfem_hip_runtime          the process-wide device context
!   fortfem_structured_tetra_box_mesh    src/mesh/structured_tetra_box_mesh.f90
!   fortfem_tetra_duffy_quadrature       src/quadrature/tetra_duffy_quadrature.f90
!   fortfem_assembly_tetra_lagrange_arbitrary_order_3d
!                                src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90
!   fortfem_sparse_matrix        src/solvers/fortfem_sparse_matrix.f90
!   fortfem_advanced_solvers     src/solvers/advanced_solvers.f90 (CG / PCG part)
!   fortfem_tetra_lagrange_solver_3d     src/solvers/fortfem_tetra_lagrange_solver_3d.f90

module fortfem_hip_types
    ! csc_t, csc_z_t, fortsparse_status_t, status_set and the status codes are
    ! fortsparse's (the CSC layout of the path is unchanged); dp is real64
    use, intrinsic :: iso_fortran_env, only: dp => real64
    use fortsparse, only: csc_t, csc_z_t, fortsparse_status_t, status_set, &
        FORTSPARSE_OK, FORTSPARSE_INVALID_MATRIX, FORTSPARSE_INTERNAL_ERROR
    implicit none
    private

    public :: csc_t, csc_z_t, fortsparse_status_t, status_set, dp
    public :: FORTSPARSE_OK, FORTSPARSE_INVALID_MATRIX, FORTSPARSE_INTERNAL_ERROR
end module fortfem_hip_types

module fortfem_hip_runtime
    use, intrinsic :: iso_c_binding, only: c_ptr, c_null_ptr, c_associated, &
        c_int, c_char, c_f_pointer
    use fortfem_hip_capi, only: ffh_init, ffh_finalize, ffh_last_error, FFH_OK
    implicit none
    private

    type(c_ptr), save :: context = c_null_ptr

    public :: hip_context, fortfem_hip_finalize, hip_error_message

contains

    ! The process-wide context on device 0, created on first use.  A failed
    ! initialisation stops the program: there is no CPU fallback.
    function hip_context() result(ctx)
        type(c_ptr) :: ctx
        integer(c_int) :: code

        if (.not. c_associated(context)) then
            code = ffh_init(0_c_int, context)
            if (code /= FFH_OK) then
                write (*, '(a,i0,2a)') "fortfem_hip: device initialisation failed (", &
                    code, "): ", trim(hip_error_message(c_null_ptr))
                error stop "fortfem_hip: no usable HIP device"
            end if
        end if
        ctx = context
    end function hip_context

    subroutine fortfem_hip_finalize()
        integer(c_int) :: code

        if (c_associated(context)) code = ffh_finalize(context)
        context = c_null_ptr
    end subroutine fortfem_hip_finalize

    function hip_error_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: i

        msg = ""
        p = ffh_last_error(ctx)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, chars, [256])
        do i = 1, 256
            if (chars(i) == char(0)) exit
            msg(i:i) = chars(i)
        end do
    end function hip_error_message

end module fortfem_hip_runtime

module fortfem_structured_tetra_box_mesh
    use, intrinsic :: iso_c_binding, only: c_int, c_double
    use fortfem_hip_types, only: dp
    use fortfem_hip_capi, only: ffh_box_mesh_size, ffh_box_mesh, FFH_OK
    implicit none
    private

    public :: generate_structured_tetra_box_mesh

contains

    ! src/mesh/structured_tetra_box_mesh.f90:12-70
    subroutine generate_structured_tetra_box_mesh(bounds, counts, vertices, &
            tetrahedra, status)
        real(dp), intent(in) :: bounds(3, 2)
        integer, intent(in) :: counts(3)
        real(dp), allocatable, intent(out) :: vertices(:, :)
        integer, allocatable, intent(out) :: tetrahedra(:, :)
        integer, intent(out) :: status

        integer(c_int) :: nvert, ntet, ccounts(3)
        real(c_double) :: cbounds(3, 2)

        status = 1
        if (any(counts < 1)) return
        if (any(bounds(:, 2) <= bounds(:, 1))) return
        ccounts = int(counts, c_int)
        cbounds = real(bounds, c_double)
        if (ffh_box_mesh_size(ccounts, nvert, ntet) /= FFH_OK) return
        allocate(vertices(3, nvert), tetrahedra(4, ntet))
        if (ffh_box_mesh(cbounds, ccounts, vertices, tetrahedra) /= FFH_OK) return
        status = 0
    end subroutine generate_structured_tetra_box_mesh

end module fortfem_structured_tetra_box_mesh

module fortfem_tetra_duffy_quadrature
    use, intrinsic :: iso_c_binding, only: c_int
    use fortfem_hip_types, only: dp
    use fortfem_hip_capi, only: ffh_tetra_duffy, FFH_OK
    implicit none
    private

    public :: tetra_duffy_quadrature

contains

    ! src/quadrature/tetra_duffy_quadrature.f90:12-51
    subroutine tetra_duffy_quadrature(polynomial_degree, x, y, z, weights, status)
        integer, intent(in) :: polynomial_degree
        real(dp), allocatable, intent(out) :: x(:), y(:), z(:), weights(:)
        integer, intent(out) :: status

        integer(c_int) :: nq
        integer :: m

        status = 1
        if (polynomial_degree < 0) return
        ! (p + 4)/2 points per direction (tetra_duffy_quadrature.f90:25); the
        ! C routine fills all nq^3 entries, so the arrays are sized first
        m = (polynomial_degree + 4)/2
        allocate(x(m**3), y(m**3), z(m**3), weights(m**3))
        if (ffh_tetra_duffy(int(polynomial_degree, c_int), nq, x, y, z, &
            weights) /= FFH_OK) return
        if (nq /= m**3) return
        status = 0
    end subroutine tetra_duffy_quadrature

end module fortfem_tetra_duffy_quadrature

module fortfem_assembly_tetra_lagrange_arbitrary_order_3d
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, &
        c_null_ptr
    use fortfem_hip_types, only: dp, csc_t, fortsparse_status_t, status_set, &
        FORTSPARSE_INVALID_MATRIX
    use fortfem_hip_capi
    use fortfem_hip_runtime, only: hip_context
    use fortfem_tetra_duffy_quadrature, only: tetra_duffy_quadrature
    ! element-level routines and the sampled-load tangents stay on the CPU
    use fortfem_assembly_tetra_lagrange_arbitrary_order_3d_cpu, only: &
        assemble_tetra_lagrange_stiffness_element, &
        assemble_tetra_lagrange_stiffness_element_jvp, &
        assemble_tetra_lagrange_stiffness_element_vjp, &
        assemble_tetra_lagrange_scalar_load_samples_jvp, &
        assemble_tetra_lagrange_scalar_load_samples_vjp
    implicit none
    private

    public :: assemble_tetra_lagrange_stiffness_element
    public :: assemble_tetra_lagrange_stiffness_element_jvp
    public :: assemble_tetra_lagrange_stiffness_element_vjp
    public :: assemble_tetra_lagrange_scalar_load_samples_jvp
    public :: assemble_tetra_lagrange_scalar_load_samples_vjp
    public :: assemble_tetra_lagrange_stiffness_csc
    public :: assemble_tetra_lagrange_stiffness_csc_jvp
    public :: assemble_tetra_lagrange_stiffness_csc_vjp
    public :: assemble_tetra_lagrange_scalar_load
    public :: assemble_tetra_lagrange_scalar_load_samples
    public :: scalar_source_3d
    public :: download_csc

    abstract interface
        pure subroutine scalar_source_3d(x, y, z, value)
            import :: dp
            real(dp), intent(in) :: x, y, z
            real(dp), intent(out) :: value
        end subroutine scalar_source_3d
    end interface

contains

    ! Full CSC of a resident matrix into a csc_t (1-based, the reference layout)
    subroutine download_csc(mat, matrix, ok)
        type(c_ptr), intent(in) :: mat
        type(csc_t), intent(out) :: matrix
        logical, intent(out) :: ok

        integer(c_int) :: n, code
        integer(c_int64_t) :: nnz
        integer(c_int) :: idummy(1)
        real(c_double) :: ddummy(1)

        ok = .false.
        code = ffh_matrix_download(mat, 1_c_int, 0_c_int64_t, idummy, idummy, &
            ddummy, n, nnz)
        if (code /= FFH_ERR_CAPACITY .and. code /= FFH_OK) return
        allocate(matrix%col_ptr(n + 1), matrix%row_idx(nnz), matrix%val(nnz))
        code = ffh_matrix_download(mat, 1_c_int, nnz, matrix%col_ptr, &
            matrix%row_idx, matrix%val, n, nnz)
        if (code /= FFH_OK) return
        matrix%nrow = n
        matrix%ncol = n
        matrix%nnz = int(nnz)
        ok = .true.
    end subroutine download_csc

    ! src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:276-334.
    ! cell_coefficient (optional, new): per-cell stiffness coefficient.
    subroutine assemble_tetra_lagrange_stiffness_csc(mesh_vertices, tetrahedra, &
            degree, quadrature_degree, matrix, status, stiffness_coefficient, &
            mass_coefficient, cell_coefficient)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        integer, intent(in) :: degree, quadrature_degree
        type(csc_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        real(dp), intent(in), optional :: stiffness_coefficient
        real(dp), intent(in), optional :: mass_coefficient
        real(dp), intent(in), optional :: cell_coefficient(:)

        type(c_ptr) :: prob, mat
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        real(c_double) :: kappa(1), mu
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral H1 sparse assembly failed")
        if (size(mesh_vertices, 1) /= 3) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        kappa = 1.0_dp
        mu = 0.0_dp
        if (present(stiffness_coefficient)) kappa = stiffness_coefficient
        if (present(mass_coefficient)) mu = mass_coefficient
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_H1, int(degree, c_int), &
            n_global, nnz)
        if (code == FFH_OK) then
            if (present(cell_coefficient)) then
                code = ffh_assemble(prob, int(quadrature_degree, c_int), &
                    FFH_COEFF_CELL, cell_coefficient, mu, FFH_FP64)
            else
                code = ffh_assemble(prob, int(quadrature_degree, c_int), &
                    FFH_COEFF_SCALAR, kappa, mu, FFH_FP64)
            end if
        end if
        if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
        ok = .false.
        if (code == FFH_OK) call download_csc(mat, matrix, ok)
        code = ffh_problem_free(prob)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_lagrange_stiffness_csc

    ! :336-386.  matrix_dot has the primal matrix's pattern.
    subroutine assemble_tetra_lagrange_stiffness_csc_jvp(mesh_vertices, &
            tetrahedra, degree, quadrature_degree, stiffness_coefficient, &
            mass_coefficient, mesh_vertices_dot, stiffness_coefficient_dot, &
            mass_coefficient_dot, matrix_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree, quadrature_degree
        real(dp), intent(in) :: stiffness_coefficient, mass_coefficient
        real(dp), intent(in) :: stiffness_coefficient_dot
        real(dp), intent(in) :: mass_coefficient_dot
        type(csc_t), intent(out) :: matrix_dot
        type(fortsparse_status_t), intent(out) :: status

        type(c_ptr) :: prob, mat
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral H1 JVP assembly failed")
        if (size(mesh_vertices, 1) /= 3) return
        if (any(shape(mesh_vertices_dot) /= shape(mesh_vertices))) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (degree < 1 .or. degree > 4) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_H1, int(degree, c_int), &
            n_global, nnz)
        ok = .false.
        if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
        if (code == FFH_OK) call download_csc(mat, matrix_dot, ok)
        if (ok) then
            code = ffh_assemble_jvp(prob, int(quadrature_degree, c_int), &
                FFH_COEFF_SCALAR, [stiffness_coefficient], mass_coefficient, &
                mesh_vertices_dot, [stiffness_coefficient_dot], &
                mass_coefficient_dot, matrix_dot%val)
            ok = code == FFH_OK
        end if
        code = ffh_problem_free(prob)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_lagrange_stiffness_csc_jvp

    ! :388-471.
    subroutine assemble_tetra_lagrange_stiffness_csc_vjp(mesh_vertices, &
            tetrahedra, degree, quadrature_degree, stiffness_coefficient, &
            mass_coefficient, matrix_values_bar, mesh_vertices_bar, &
            stiffness_coefficient_bar, mass_coefficient_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree, quadrature_degree
        real(dp), intent(in) :: stiffness_coefficient, mass_coefficient
        real(dp), intent(in) :: matrix_values_bar(:)
        real(dp), intent(out) :: mesh_vertices_bar(:, :)
        real(dp), intent(out) :: stiffness_coefficient_bar
        real(dp), intent(out) :: mass_coefficient_bar
        type(fortsparse_status_t), intent(out) :: status

        type(c_ptr) :: prob
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        real(c_double) :: kbar(1)
        logical :: ok

        mesh_vertices_bar = 0.0_dp
        stiffness_coefficient_bar = 0.0_dp
        mass_coefficient_bar = 0.0_dp
        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral H1 VJP shapes differ")
        if (size(mesh_vertices, 1) /= 3) return
        if (any(shape(mesh_vertices_bar) /= shape(mesh_vertices))) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (degree < 1 .or. degree > 4) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_H1, int(degree, c_int), &
            n_global, nnz)
        ok = code == FFH_OK
        if (ok) ok = size(matrix_values_bar) == nnz
        if (ok) then
            code = ffh_assemble_vjp(prob, int(quadrature_degree, c_int), &
                FFH_COEFF_SCALAR, [stiffness_coefficient], mass_coefficient, &
                matrix_values_bar, mesh_vertices_bar, kbar, mass_coefficient_bar)
            ok = code == FFH_OK
        end if
        code = ffh_problem_free(prob)
        if (ok) then
            stiffness_coefficient_bar = kbar(1)
            call status_set(status, 0, "")
        else
            mesh_vertices_bar = 0.0_dp
            mass_coefficient_bar = 0.0_dp
        end if
    end subroutine assemble_tetra_lagrange_stiffness_csc_vjp

    ! :473-528.  source_values(nq, ncell) at the Duffy points of each cell.
    subroutine assemble_tetra_lagrange_scalar_load_samples(mesh_vertices, &
            tetrahedra, degree, quadrature_degree, source_values, &
            right_hand_side, status)
        real(dp), intent(in) :: mesh_vertices(:, :), source_values(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree, quadrature_degree
        real(dp), allocatable, intent(out) :: right_hand_side(:)
        type(fortsparse_status_t), intent(out) :: status

        type(c_ptr) :: prob
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Sampled tetrahedral H1 scalar load failed")
        if (size(mesh_vertices, 1) /= 3 .or. size(tetrahedra, 1) /= 4) return
        if (size(source_values, 2) /= size(tetrahedra, 2)) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_H1, int(degree, c_int), &
            n_global, nnz)
        if (code == FFH_OK) code = ffh_load(prob, int(quadrature_degree, c_int), &
            FFH_SOURCE_SAMPLES, source_values, int(size(source_values), c_int64_t))
        if (code == FFH_OK) then
            allocate(right_hand_side(n_global))
            code = ffh_rhs_download(prob, FFH_FULL, right_hand_side)
        end if
        if (code == FFH_OK) call status_set(status, 0, "")
        code = ffh_problem_free(prob)
    end subroutine assemble_tetra_lagrange_scalar_load_samples

    ! :705-774.  The callback runs on the host at the mapped Duffy points
    ! x_q = v_1 + J xi_q (:761-763); the samples go to the device kernel.
    subroutine assemble_tetra_lagrange_scalar_load(mesh_vertices, tetrahedra, &
            degree, quadrature_degree, source, right_hand_side, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        integer, intent(in) :: degree, quadrature_degree
        procedure(scalar_source_3d) :: source
        real(dp), allocatable, intent(out) :: right_hand_side(:)
        type(fortsparse_status_t), intent(out) :: status

        real(dp), allocatable :: samples(:, :)

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral H1 scalar load assembly failed")
        if (.not. sample_source(mesh_vertices, tetrahedra, quadrature_degree, &
            source, samples)) return
        call assemble_tetra_lagrange_scalar_load_samples(mesh_vertices, &
            tetrahedra, degree, quadrature_degree, samples, right_hand_side, status)
    end subroutine assemble_tetra_lagrange_scalar_load

    logical function sample_source(mesh_vertices, tetrahedra, quadrature_degree, &
            source, samples) result(ok)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), quadrature_degree
        procedure(scalar_source_3d) :: source
        real(dp), allocatable, intent(out) :: samples(:, :)

        real(dp), allocatable :: x(:), y(:), z(:), w(:)
        real(dp) :: jac(3, 3), point(3)
        integer :: cell, q, st

        ok = .false.
        if (size(mesh_vertices, 1) /= 3 .or. size(tetrahedra, 1) /= 4) return
        if (any(tetrahedra < 1) .or. any(tetrahedra > size(mesh_vertices, 2))) return
        call tetra_duffy_quadrature(quadrature_degree, x, y, z, w, st)
        if (st /= 0) return
        allocate(samples(size(w), size(tetrahedra, 2)))
        do cell = 1, size(tetrahedra, 2)
            jac(:, 1) = mesh_vertices(:, tetrahedra(2, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 2) = mesh_vertices(:, tetrahedra(3, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 3) = mesh_vertices(:, tetrahedra(4, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            do q = 1, size(w)
                point = mesh_vertices(:, tetrahedra(1, cell)) + &
                    matmul(jac, [x(q), y(q), z(q)])
                call source(point(1), point(2), point(3), samples(q, cell))
            end do
        end do
        ok = .true.
    end function sample_source

end module fortfem_assembly_tetra_lagrange_arbitrary_order_3d

! src/assembly/assembly_tetra_nedelec_3d.f90: curl-mass and tensor-weighted
! first-kind Nedelec CSC assembly, orders 1..6, on the GPU.
module fortfem_assembly_tetra_nedelec_3d
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, c_null_ptr, c_loc
    use fortfem_hip_types, only: dp, csc_t, csc_z_t, fortsparse_status_t, status_set, &
        FORTSPARSE_INVALID_MATRIX
    use fortfem_hip_capi
    use fortfem_hip_runtime, only: hip_context
    use fortfem_assembly_tetra_lagrange_arbitrary_order_3d, only: download_csc
    ! element-level routines and the sampled vector loads stay on the CPU
    use fortfem_assembly_tetra_nedelec_3d_cpu, only: &
        assemble_tetra_nedelec_curl_mass_element, &
        assemble_tetra_nedelec_curl_mass_element_jvp, &
        assemble_tetra_nedelec_curl_mass_element_vjp, &
        assemble_tetra_nedelec_pml_element, assemble_tetra_nedelec_pml_element_jvp, &
        assemble_tetra_nedelec_pml_element_vjp, &
        assemble_tetra_nedelec_curvilinear_pml_element, &
        assemble_tetra_nedelec_curvilinear_pml_element_jvp, &
        assemble_tetra_nedelec_curvilinear_pml_element_vjp, &
        assemble_tetra_nedelec_vector_load_samples, &
        assemble_tetra_nedelec_vector_load_samples_jvp, &
        assemble_tetra_nedelec_vector_load_samples_vjp
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
    public :: assemble_tetra_nedelec_pml_csc
    public :: assemble_tetra_nedelec_curvilinear_pml_csc
    public :: assemble_tetra_nedelec_pml_csc_jvp, assemble_tetra_nedelec_pml_csc_vjp
    public :: assemble_tetra_nedelec_curvilinear_pml_csc_jvp
    public :: assemble_tetra_nedelec_curvilinear_pml_csc_vjp

    public :: assemble_tetra_nedelec_curl_mass_csc
    public :: assemble_tetra_nedelec_curl_mass_csc_jvp
    public :: assemble_tetra_nedelec_curl_mass_csc_vjp
    public :: assemble_tetra_nedelec_weighted_csc
    public :: assemble_tetra_nedelec_weighted_csc_cells
    public :: assemble_tetra_nedelec_vector_load
    public :: assemble_tetra_nedelec_vector_load_order
    public :: tensor_coefficient_3d
    public :: vector_source_3d
    public :: sample_vector_source, sample_rule_tensors

    abstract interface
        pure subroutine tensor_coefficient_3d(x, y, z, value)
            import :: dp
            real(dp), intent(in) :: x, y, z
            real(dp), intent(out) :: value(3, 3)
        end subroutine tensor_coefficient_3d
        pure subroutine vector_source_3d(x, y, z, value)
            import :: dp
            real(dp), intent(in) :: x, y, z
            real(dp), intent(out) :: value(3)
        end subroutine vector_source_3d
    end interface

contains

    ! :2002-2040 (order 1).  The callback runs on the host at the mapped load
    ! points x_q = v_1 + J xi_q (:2539-2541); the samples go to the device.
    subroutine assemble_tetra_nedelec_vector_load(mesh_vertices, tetrahedra, &
            source, vector, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        procedure(vector_source_3d) :: source
        real(dp), allocatable, intent(out) :: vector(:)
        type(fortsparse_status_t), intent(out) :: status

        call assemble_tetra_nedelec_vector_load_order(mesh_vertices, tetrahedra, &
            1, source, vector, status)
    end subroutine assemble_tetra_nedelec_vector_load

    ! :71-127 (orders 2..6; order 1 as above).
    subroutine assemble_tetra_nedelec_vector_load_order(mesh_vertices, &
            tetrahedra, order, source, vector, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(vector_source_3d) :: source
        real(dp), allocatable, intent(out) :: vector(:)
        type(fortsparse_status_t), intent(out) :: status

        real(c_double), allocatable :: samples(:, :, :)
        type(c_ptr) :: prob
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral Nedelec vector load assembly failed")
        if (.not. sample_vector_source(mesh_vertices, tetrahedra, order, source, &
            samples)) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(order, c_int), &
            n_global, nnz)
        if (code == FFH_OK) code = ffh_load(prob, -1_c_int, FFH_SOURCE_SAMPLES, &
            samples, int(size(samples), c_int64_t))
        if (code == FFH_OK) then
            allocate(vector(n_global))
            code = ffh_rhs_download(prob, FFH_FULL, vector)
        end if
        if (code == FFH_OK) call status_set(status, 0, "")
        code = ffh_problem_free(prob)
    end subroutine assemble_tetra_nedelec_vector_load_order

    ! The source at the mapped load points of every cell, samples(3, nq, T)
    ! (the device load's sample layout); false on an invalid mesh or order.
    logical function sample_vector_source(mesh_vertices, tetrahedra, order, &
            source, samples) result(ok)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(vector_source_3d) :: source
        real(c_double), allocatable, intent(out) :: samples(:, :, :)

        real(c_double), allocatable, target :: pts(:, :)
        real(dp) :: jac(3, 3), point(3)
        integer(c_int) :: nq
        integer :: cell, q

        ok = .false.
        if (size(mesh_vertices, 1) /= 3 .or. size(tetrahedra, 1) /= 4) return
        if (size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (order < 1 .or. order > 6) return
        if (ffh_nedelec_load_points(int(order, c_int), nq, c_null_ptr) /= FFH_OK) return
        allocate(pts(4, nq))
        if (ffh_nedelec_load_points(int(order, c_int), nq, c_loc(pts)) /= FFH_OK) return
        allocate(samples(3, nq, size(tetrahedra, 2)))
        do cell = 1, size(tetrahedra, 2)
            jac(:, 1) = mesh_vertices(:, tetrahedra(2, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 2) = mesh_vertices(:, tetrahedra(3, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 3) = mesh_vertices(:, tetrahedra(4, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            do q = 1, nq
                point = mesh_vertices(:, tetrahedra(1, cell)) + &
                    matmul(jac, pts(1:3, q))
                call source(point(1), point(2), point(3), samples(:, q, cell))
            end do
        end do
        ok = .true.
    end function sample_vector_source

    ! The curl tensor of the weighted element where the reference evaluates
    ! it (:2435, :2499): at every point of the rule (ffh_nedelec_load_points)
    ! mapped into each cell.  values holds 9 per point (column-major),
    ! cell-major, and kind is FFH_COEFF_POINT_TENSOR; when the tensor is the
    ! same at all points of every cell, one tensor per cell and
    ! FFH_COEFF_CELL_TENSOR (the class-table element).  False on bad input.
    logical function sample_rule_tensors(mesh_vertices, tetrahedra, order, &
            coefficient, values, kind) result(ok)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(tensor_coefficient_3d) :: coefficient
        real(c_double), allocatable, intent(out) :: values(:)
        integer(c_int), intent(out) :: kind

        real(c_double), allocatable, target :: pts(:, :)
        real(dp), allocatable :: kp(:, :, :, :)
        real(dp) :: jac(3, 3), point(3)
        integer(c_int) :: nq
        integer :: cell, q, ncell
        logical :: constant

        ok = .false.
        kind = FFH_COEFF_CELL_TENSOR
        if (size(mesh_vertices, 1) /= 3 .or. size(tetrahedra, 1) /= 4) return
        if (size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (order < 1 .or. order > 6) return
        if (ffh_nedelec_load_points(int(order, c_int), nq, c_null_ptr) /= FFH_OK) return
        allocate(pts(4, nq))
        if (ffh_nedelec_load_points(int(order, c_int), nq, c_loc(pts)) /= FFH_OK) return
        ncell = size(tetrahedra, 2)
        allocate(kp(3, 3, nq, ncell))
        constant = .true.
        do cell = 1, ncell
            jac(:, 1) = mesh_vertices(:, tetrahedra(2, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 2) = mesh_vertices(:, tetrahedra(3, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            jac(:, 3) = mesh_vertices(:, tetrahedra(4, cell)) - &
                mesh_vertices(:, tetrahedra(1, cell))
            do q = 1, nq
                point = mesh_vertices(:, tetrahedra(1, cell)) + &
                    matmul(jac, pts(1:3, q))
                call coefficient(point(1), point(2), point(3), kp(:, :, q, cell))
                if (any(kp(:, :, q, cell) /= kp(:, :, 1, cell))) constant = .false.
            end do
        end do
        if (constant) then
            values = reshape(kp(:, :, 1, :), [9*ncell])
        else
            values = reshape(kp, [9*int(nq)*ncell])
            kind = FFH_COEFF_POINT_TENSOR
        end if
        ok = .true.
    end function sample_rule_tensors

    ! Shared driver: pattern + assembly of one Nedelec problem, downloaded.
    subroutine nedelec_csc(mesh_vertices, tetrahedra, k, kind, coeff, mass, &
            matrix, ok)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        integer, intent(in) :: k
        integer(c_int), intent(in) :: kind
        real(c_double), intent(in) :: coeff(*)
        real(dp), intent(in) :: mass
        type(csc_t), intent(out) :: matrix
        logical, intent(out) :: ok
        type(c_ptr) :: prob, mat
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz

        ok = .false.
        if (size(mesh_vertices, 1) /= 3) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(k, c_int), &
            n_global, nnz)
        if (code == FFH_OK) code = ffh_assemble(prob, -1_c_int, kind, coeff, &
            real(mass, c_double), FFH_FP64)
        if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
        if (code == FFH_OK) call download_csc(mat, matrix, ok)
        code = ffh_problem_free(prob)
    end subroutine nedelec_csc

    ! Complex PML operator on the Nedelec pattern: the pattern comes from the
    ! device matrix, the complex values from ffh_nedelec_pml_values.
    subroutine nedelec_pml(mesh_vertices, tetrahedra, k, kind, stretch, &
            wave_number, matrix, ok)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        integer, intent(in) :: k
        integer(c_int), intent(in) :: kind
        complex(dp), intent(in) :: stretch(*)
        real(dp), intent(in) :: wave_number
        type(csc_z_t), intent(out) :: matrix
        logical, intent(out) :: ok
        type(c_ptr) :: prob, mat
        type(csc_t) :: pattern
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz

        ok = .false.
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(k, c_int), &
            n_global, nnz)
        if (code == FFH_OK) then
            allocate(matrix%val(max(nnz, 1_c_int64_t)))
            code = ffh_nedelec_pml_values(prob, kind, stretch, &
                real(wave_number, c_double), matrix%val)
        end if
        if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
        if (code == FFH_OK) call download_csc(mat, pattern, ok)
        code = ffh_problem_free(prob)
        if (.not. ok) return
        matrix%nrow = pattern%nrow
        matrix%ncol = pattern%ncol
        matrix%nnz = pattern%nnz
        call move_alloc(pattern%col_ptr, matrix%col_ptr)
        call move_alloc(pattern%row_idx, matrix%row_idx)
        if (size(matrix%val) /= matrix%nnz) matrix%val = matrix%val(1:matrix%nnz)
    end subroutine nedelec_pml

    logical function pml_input_ok(mesh_vertices, tetrahedra, order, &
            wave_number)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        real(dp), intent(in) :: wave_number

        pml_input_ok = .false.
        if (size(mesh_vertices, 1) /= 3) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (order < 1 .or. order > 6) return
        if (.not. (wave_number > 0.0_dp)) return
        pml_input_ok = .true.
    end function pml_input_ok

    ! Tangent or cotangent of the PML operator on a fresh pattern (mode 0:
    ! JVP into matrix_dot; mode 1: VJP of values_bar).
    subroutine nedelec_pml_ad(mesh_vertices, tetrahedra, k, kind, mode, stretch, &
            wave_number, vertices_dot, stretch_dot, wave_number_dot, matrix_dot, &
            values_bar, vertices_bar, stretch_bar, wave_number_bar, ok)
        real(dp), intent(in) :: mesh_vertices(:, :), vertices_dot(*)
        integer, intent(in) :: tetrahedra(:, :), k, mode
        integer(c_int), intent(in) :: kind
        complex(dp), intent(in) :: stretch(*), stretch_dot(*), values_bar(*)
        real(dp), intent(in) :: wave_number, wave_number_dot
        type(csc_z_t), intent(out) :: matrix_dot
        real(dp), intent(out) :: vertices_bar(*), wave_number_bar
        complex(dp), intent(out) :: stretch_bar(*)
        logical, intent(out) :: ok
        type(c_ptr) :: prob, mat
        type(csc_t) :: pattern
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        real(c_double) :: kb

        ok = .false.
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(k, c_int), n_global, nnz)
        if (code == FFH_OK .and. mode == 0) then
            allocate(matrix_dot%val(max(nnz, 1_c_int64_t)))
            code = ffh_nedelec_pml_values_jvp(prob, kind, stretch, stretch_dot, &
                real(wave_number, c_double), real(wave_number_dot, c_double), &
                vertices_dot, matrix_dot%val)
            if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
            if (code == FFH_OK) call download_csc(mat, pattern, ok)
            if (ok) then
                matrix_dot%nrow = pattern%nrow
                matrix_dot%ncol = pattern%ncol
                matrix_dot%nnz = pattern%nnz
                call move_alloc(pattern%col_ptr, matrix_dot%col_ptr)
                call move_alloc(pattern%row_idx, matrix_dot%row_idx)
                if (size(matrix_dot%val) /= matrix_dot%nnz) &
                    matrix_dot%val = matrix_dot%val(1:matrix_dot%nnz)
            end if
        else if (code == FFH_OK) then
            code = ffh_nedelec_pml_values_vjp(prob, kind, stretch, &
                real(wave_number, c_double), values_bar, vertices_bar, stretch_bar, kb)
            wave_number_bar = kb
            ok = code == FFH_OK
        end if
        code = ffh_problem_free(prob)
    end subroutine nedelec_pml_ad

    ! :1687-1768.  matrix_dot has the primal pattern.
    subroutine assemble_tetra_nedelec_pml_csc_jvp(mesh_vertices, tetrahedra, order, &
            stretch, wave_number, mesh_vertices_dot, stretch_dot, wave_number_dot, &
            matrix_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :), stretch_dot(:, :)
        real(dp), intent(in) :: wave_number, wave_number_dot
        type(csc_z_t), intent(out) :: matrix_dot
        type(fortsparse_status_t), intent(out) :: status
        complex(dp) :: cz(1)
        real(dp) :: rz(1), wb
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Complex tetrahedral Nedelec PML JVP failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3 .or. size(stretch, 2) /= size(tetrahedra, 2)) return
        if (any(shape(stretch_dot) /= shape(stretch))) return
        if (any(shape(mesh_vertices_dot) /= shape(mesh_vertices))) return
        call nedelec_pml_ad(mesh_vertices, tetrahedra, order, 0_c_int, 0, stretch, &
            wave_number, mesh_vertices_dot, stretch_dot, wave_number_dot, matrix_dot, &
            cz, rz, cz, wb, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_pml_csc_jvp

    ! :1770-1868: the reference's convention (cotangent = sum of seed times
    ! the conjugate derivative; real part for the vertices and wave number).
    subroutine assemble_tetra_nedelec_pml_csc_vjp(mesh_vertices, tetrahedra, order, &
            stretch, wave_number, matrix_values_bar, mesh_vertices_bar, stretch_bar, &
            wave_number_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :), matrix_values_bar(:)
        real(dp), intent(in) :: wave_number
        real(dp), intent(out) :: mesh_vertices_bar(:, :), wave_number_bar
        complex(dp), intent(out) :: stretch_bar(:, :)
        type(fortsparse_status_t), intent(out) :: status
        type(csc_z_t) :: unused
        complex(dp) :: cz(1)
        real(dp) :: rz(1)
        logical :: ok

        mesh_vertices_bar = 0.0_dp
        stretch_bar = (0.0_dp, 0.0_dp)
        wave_number_bar = 0.0_dp
        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Complex tetrahedral Nedelec PML VJP failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3 .or. size(stretch, 2) /= size(tetrahedra, 2)) return
        if (any(shape(stretch_bar) /= shape(stretch))) return
        if (any(shape(mesh_vertices_bar) /= shape(mesh_vertices))) return
        call nedelec_pml_ad(mesh_vertices, tetrahedra, order, 0_c_int, 1, stretch, &
            wave_number, rz, cz, 0.0_dp, unused, matrix_values_bar, mesh_vertices_bar, &
            stretch_bar, wave_number_bar, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_pml_csc_vjp

    ! :1446-1521
    subroutine assemble_tetra_nedelec_curvilinear_pml_csc_jvp(mesh_vertices, &
            tetrahedra, order, stretch, wave_number, mesh_vertices_dot, stretch_dot, &
            wave_number_dot, matrix_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :, :), stretch_dot(:, :, :)
        real(dp), intent(in) :: wave_number, wave_number_dot
        type(csc_z_t), intent(out) :: matrix_dot
        type(fortsparse_status_t), intent(out) :: status
        complex(dp) :: cz(1)
        real(dp) :: rz(1), wb
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Curvilinear tetrahedral Nedelec PML JVP failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3 .or. size(stretch, 2) /= 3) return
        if (size(stretch, 3) /= size(tetrahedra, 2)) return
        if (any(shape(stretch_dot) /= shape(stretch))) return
        if (any(shape(mesh_vertices_dot) /= shape(mesh_vertices))) return
        call nedelec_pml_ad(mesh_vertices, tetrahedra, order, 1_c_int, 0, stretch, &
            wave_number, mesh_vertices_dot, stretch_dot, wave_number_dot, matrix_dot, &
            cz, rz, cz, wb, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_curvilinear_pml_csc_jvp

    ! :1523-1611
    subroutine assemble_tetra_nedelec_curvilinear_pml_csc_vjp(mesh_vertices, &
            tetrahedra, order, stretch, wave_number, matrix_values_bar, &
            mesh_vertices_bar, stretch_bar, wave_number_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :, :), matrix_values_bar(:)
        real(dp), intent(in) :: wave_number
        real(dp), intent(out) :: mesh_vertices_bar(:, :), wave_number_bar
        complex(dp), intent(out) :: stretch_bar(:, :, :)
        type(fortsparse_status_t), intent(out) :: status
        type(csc_z_t) :: unused
        complex(dp) :: cz(1)
        real(dp) :: rz(1)
        logical :: ok

        mesh_vertices_bar = 0.0_dp
        stretch_bar = (0.0_dp, 0.0_dp)
        wave_number_bar = 0.0_dp
        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Curvilinear tetrahedral Nedelec PML VJP failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3 .or. size(stretch, 2) /= 3) return
        if (size(stretch, 3) /= size(tetrahedra, 2)) return
        if (any(shape(stretch_bar) /= shape(stretch))) return
        if (any(shape(mesh_vertices_bar) /= shape(mesh_vertices))) return
        call nedelec_pml_ad(mesh_vertices, tetrahedra, order, 1_c_int, 1, stretch, &
            wave_number, rz, cz, 0.0_dp, unused, matrix_values_bar, mesh_vertices_bar, &
            stretch_bar, wave_number_bar, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_curvilinear_pml_csc_vjp

    ! :1613-1685.  stretch(3, n_tetrahedra): the cell's complex stretch
    ! factors along x, y, z.
    subroutine assemble_tetra_nedelec_pml_csc(mesh_vertices, tetrahedra, &
            order, stretch, wave_number, matrix, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :)
        real(dp), intent(in) :: wave_number
        type(csc_z_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        complex(dp), allocatable :: packed(:, :)
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Complex tetrahedral Nedelec PML assembly failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3) return
        if (size(stretch, 2) /= size(tetrahedra, 2)) return
        packed = stretch
        call nedelec_pml(mesh_vertices, tetrahedra, order, 0_c_int, packed, &
            wave_number, matrix, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_pml_csc

    ! :1372-1444.  stretch(3, 3, n_tetrahedra): the cell's complex stretch
    ! matrix.
    subroutine assemble_tetra_nedelec_curvilinear_pml_csc(mesh_vertices, &
            tetrahedra, order, stretch, wave_number, matrix, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        complex(dp), intent(in) :: stretch(:, :, :)
        real(dp), intent(in) :: wave_number
        type(csc_z_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        complex(dp), allocatable :: packed(:, :, :)
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Curvilinear tetrahedral Nedelec PML assembly failed")
        if (.not. pml_input_ok(mesh_vertices, tetrahedra, order, wave_number)) return
        if (size(stretch, 1) /= 3 .or. size(stretch, 2) /= 3) return
        if (size(stretch, 3) /= size(tetrahedra, 2)) return
        packed = stretch
        call nedelec_pml(mesh_vertices, tetrahedra, order, 1_c_int, packed, &
            wave_number, matrix, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_curvilinear_pml_csc

    ! :2042-2112.  curl_coefficient and mass_coefficient default to 1, order
    ! to 1.  cell_coefficient (optional, new): per-cell curl coefficient.
    subroutine assemble_tetra_nedelec_curl_mass_csc(mesh_vertices, tetrahedra, &
            matrix, status, curl_coefficient, mass_coefficient, order, &
            cell_coefficient)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        type(csc_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        real(dp), intent(in), optional :: curl_coefficient, mass_coefficient
        integer, intent(in), optional :: order
        real(dp), intent(in), optional :: cell_coefficient(:)
        real(c_double) :: alpha(1)
        real(dp) :: beta
        integer :: k
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral Nedelec sparse assembly failed")
        alpha = 1.0_dp
        beta = 1.0_dp
        k = 1
        if (present(curl_coefficient)) alpha = curl_coefficient
        if (present(mass_coefficient)) beta = mass_coefficient
        if (present(order)) k = order
        if (k < 1 .or. k > 6) return
        if (present(cell_coefficient)) then
            if (size(cell_coefficient) /= size(tetrahedra, 2)) return
            call nedelec_csc(mesh_vertices, tetrahedra, k, FFH_COEFF_CELL, &
                cell_coefficient, beta, matrix, ok)
        else
            call nedelec_csc(mesh_vertices, tetrahedra, k, FFH_COEFF_SCALAR, &
                alpha, beta, matrix, ok)
        end if
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_curl_mass_csc

    ! :2114-2180.  matrix_dot has the primal matrix's pattern.
    subroutine assemble_tetra_nedelec_curl_mass_csc_jvp(mesh_vertices, &
            tetrahedra, order, curl_coefficient, mass_coefficient, &
            mesh_vertices_dot, curl_coefficient_dot, mass_coefficient_dot, &
            matrix_dot, status)
        real(dp), intent(in) :: mesh_vertices(:, :), mesh_vertices_dot(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        real(dp), intent(in) :: curl_coefficient, mass_coefficient
        real(dp), intent(in) :: curl_coefficient_dot, mass_coefficient_dot
        type(csc_t), intent(out) :: matrix_dot
        type(fortsparse_status_t), intent(out) :: status

        type(c_ptr) :: prob, mat
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral Nedelec curl-mass JVP assembly failed")
        if (size(mesh_vertices, 1) /= 3) return
        if (any(shape(mesh_vertices_dot) /= shape(mesh_vertices))) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (order < 1 .or. order > 6) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(order, c_int), &
            n_global, nnz)
        ok = .false.
        if (code == FFH_OK) code = ffh_problem_matrix(prob, FFH_FULL, mat)
        if (code == FFH_OK) call download_csc(mat, matrix_dot, ok)
        if (ok) then
            code = ffh_assemble_jvp(prob, -1_c_int, FFH_COEFF_SCALAR, &
                [curl_coefficient], mass_coefficient, mesh_vertices_dot, &
                [curl_coefficient_dot], mass_coefficient_dot, matrix_dot%val)
            ok = code == FFH_OK
        end if
        code = ffh_problem_free(prob)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_curl_mass_csc_jvp

    ! :2182-2287.
    subroutine assemble_tetra_nedelec_curl_mass_csc_vjp(mesh_vertices, &
            tetrahedra, order, curl_coefficient, mass_coefficient, &
            matrix_values_bar, mesh_vertices_bar, curl_coefficient_bar, &
            mass_coefficient_bar, status)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        real(dp), intent(in) :: curl_coefficient, mass_coefficient
        real(dp), intent(in) :: matrix_values_bar(:)
        real(dp), intent(out) :: mesh_vertices_bar(:, :)
        real(dp), intent(out) :: curl_coefficient_bar, mass_coefficient_bar
        type(fortsparse_status_t), intent(out) :: status

        type(c_ptr) :: prob
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz
        real(c_double) :: abar(1)
        logical :: ok

        mesh_vertices_bar = 0.0_dp
        curl_coefficient_bar = 0.0_dp
        mass_coefficient_bar = 0.0_dp
        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral Nedelec curl-mass VJP shapes differ")
        if (size(mesh_vertices, 1) /= 3) return
        if (any(shape(mesh_vertices_bar) /= shape(mesh_vertices))) return
        if (size(tetrahedra, 1) /= 4 .or. size(tetrahedra, 2) < 1) return
        if (any(tetrahedra < 1) .or. &
            any(tetrahedra > size(mesh_vertices, 2))) return
        if (order < 1 .or. order > 6) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(mesh_vertices, 2), c_int), mesh_vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(order, c_int), &
            n_global, nnz)
        ok = code == FFH_OK
        if (ok) ok = size(matrix_values_bar) == nnz
        if (ok) then
            code = ffh_assemble_vjp(prob, -1_c_int, FFH_COEFF_SCALAR, &
                [curl_coefficient], mass_coefficient, matrix_values_bar, &
                mesh_vertices_bar, abar, mass_coefficient_bar)
            ok = code == FFH_OK
        end if
        code = ffh_problem_free(prob)
        if (ok) then
            curl_coefficient_bar = abar(1)
            call status_set(status, 0, "")
        else
            mesh_vertices_bar = 0.0_dp
            mass_coefficient_bar = 0.0_dp
        end if
    end subroutine assemble_tetra_nedelec_curl_mass_csc_vjp

    ! :1870-2000.  The callback is evaluated at every point of the rule in
    ! every cell, as the reference does; tensors constant on each cell take
    ! the class-table element, others the per-point element (dense curl
    ! contraction on MFMA).
    subroutine assemble_tetra_nedelec_weighted_csc(mesh_vertices, tetrahedra, &
            coefficient, mass_coefficient, matrix, status, order)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        procedure(tensor_coefficient_3d) :: coefficient
        real(dp), intent(in) :: mass_coefficient
        type(csc_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        integer, intent(in), optional :: order
        real(c_double), allocatable :: values(:)
        integer(c_int) :: kind
        integer :: k
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral weighted Nedelec assembly failed")
        k = 1
        if (present(order)) k = order
        if (.not. sample_rule_tensors(mesh_vertices, tetrahedra, k, coefficient, &
            values, kind)) return
        call nedelec_csc(mesh_vertices, tetrahedra, k, kind, values, &
            mass_coefficient, matrix, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_weighted_csc

    ! New: one symmetric tensor K(3,3,cell) per cell.
    subroutine assemble_tetra_nedelec_weighted_csc_cells(mesh_vertices, &
            tetrahedra, cell_tensor, mass_coefficient, matrix, status, order)
        real(dp), intent(in) :: mesh_vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :)
        real(dp), intent(in) :: cell_tensor(:, :, :)
        real(dp), intent(in) :: mass_coefficient
        type(csc_t), intent(out) :: matrix
        type(fortsparse_status_t), intent(out) :: status
        integer, intent(in), optional :: order
        integer :: k
        logical :: ok

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral weighted Nedelec assembly failed")
        k = 1
        if (present(order)) k = order
        if (k < 1 .or. k > 6) return
        if (size(cell_tensor, 1) /= 3 .or. size(cell_tensor, 2) /= 3) return
        if (size(cell_tensor, 3) /= size(tetrahedra, 2)) return
        call nedelec_csc(mesh_vertices, tetrahedra, k, FFH_COEFF_CELL_TENSOR, &
            cell_tensor, mass_coefficient, matrix, ok)
        if (ok) call status_set(status, 0, "")
    end subroutine assemble_tetra_nedelec_weighted_csc_cells

end module fortfem_assembly_tetra_nedelec_3d

module fortfem_sparse_matrix
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_ptr
    use fortfem_hip_types, only: dp, sparse_matrix_t => csc_t
    use fortfem_hip_capi, only: ffh_matrix_upload, ffh_matrix_free, ffh_spmv, &
        ffh_spmv_jvp, ffh_spmv_vjp, FFH_OK
    use fortfem_hip_runtime, only: hip_context
    use fortfem_sparse_matrix_cpu, only: sparse_from_dense
    implicit none
    private

    public :: sparse_matrix_t, sparse_from_dense, spmv, spmv_jvp, spmv_vjp

contains

    ! src/solvers/fortfem_sparse_matrix.f90:46-58
    subroutine spmv(matrix, vector, product)
        type(sparse_matrix_t), intent(in) :: matrix
        real(dp), intent(in) :: vector(:)
        real(dp), intent(out) :: product(:)

        type(c_ptr) :: mat
        integer(c_int) :: code

        if (size(vector) /= matrix%ncol) then
            error stop "spmv: vector has incompatible size"
        end if
        if (size(product) /= matrix%nrow) then
            error stop "spmv: product has incompatible size"
        end if
        code = ffh_matrix_upload(hip_context(), int(matrix%ncol, c_int), &
            int(matrix%nnz, c_int64_t), 1_c_int, matrix%col_ptr, matrix%row_idx, &
            matrix%val, mat)
        if (code /= FFH_OK) error stop "spmv: device upload failed"
        code = ffh_spmv(mat, vector, product)
        if (code /= FFH_OK) error stop "spmv: device product failed"
        code = ffh_matrix_free(mat)
    end subroutine spmv

    ! :60-91.  product_dot = A vector_dot + A_dot vector, A_dot given by its
    ! CSC values on the matrix's pattern.
    subroutine spmv_jvp(matrix, vector, matrix_values_dot, vector_dot, product_dot)
        type(sparse_matrix_t), intent(in) :: matrix
        real(dp), intent(in) :: vector(:), matrix_values_dot(:)
        real(dp), intent(in) :: vector_dot(:)
        real(dp), intent(out) :: product_dot(:)

        type(c_ptr) :: mat
        integer(c_int) :: code

        if (size(vector) /= matrix%ncol) then
            error stop "spmv_jvp: vector has incompatible size"
        end if
        if (size(vector_dot) /= matrix%ncol) then
            error stop "spmv_jvp: vector_dot has incompatible size"
        end if
        if (size(matrix_values_dot) /= matrix%nnz) then
            error stop "spmv_jvp: matrix value tangent has incompatible size"
        end if
        if (size(product_dot) /= matrix%nrow) then
            error stop "spmv_jvp: product_dot has incompatible size"
        end if
        code = ffh_matrix_upload(hip_context(), int(matrix%ncol, c_int), &
            int(matrix%nnz, c_int64_t), 1_c_int, matrix%col_ptr, matrix%row_idx, &
            matrix%val, mat)
        if (code /= FFH_OK) error stop "spmv_jvp: device upload failed"
        code = ffh_spmv_jvp(mat, vector, matrix_values_dot, vector_dot, product_dot)
        if (code /= FFH_OK) error stop "spmv_jvp: device product failed"
        code = ffh_matrix_free(mat)
    end subroutine spmv_jvp

    ! :93-126.  matrix_values_bar(k) = product_bar(row_k) vector(col_k),
    ! vector_bar = A^T product_bar.
    subroutine spmv_vjp(matrix, vector, product_bar, matrix_values_bar, vector_bar)
        type(sparse_matrix_t), intent(in) :: matrix
        real(dp), intent(in) :: vector(:), product_bar(:)
        real(dp), intent(out) :: matrix_values_bar(:), vector_bar(:)

        type(c_ptr) :: mat
        integer(c_int) :: code

        if (size(vector) /= matrix%ncol) then
            error stop "spmv_vjp: vector has incompatible size"
        end if
        if (size(product_bar) /= matrix%nrow) then
            error stop "spmv_vjp: product_bar has incompatible size"
        end if
        if (size(matrix_values_bar) /= matrix%nnz) then
            error stop "spmv_vjp: matrix value cotangent has incompatible size"
        end if
        if (size(vector_bar) /= matrix%ncol) then
            error stop "spmv_vjp: vector_bar has incompatible size"
        end if
        code = ffh_matrix_upload(hip_context(), int(matrix%ncol, c_int), &
            int(matrix%nnz, c_int64_t), 1_c_int, matrix%col_ptr, matrix%row_idx, &
            matrix%val, mat)
        if (code /= FFH_OK) error stop "spmv_vjp: device upload failed"
        code = ffh_spmv_vjp(mat, vector, product_bar, matrix_values_bar, vector_bar)
        if (code /= FFH_OK) error stop "spmv_vjp: device product failed"
        code = ffh_matrix_free(mat)
    end subroutine spmv_vjp

end module fortfem_sparse_matrix

module fortfem_advanced_solvers
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr
    use fortfem_hip_types, only: dp, sparse_matrix_t => csc_t
    use fortfem_hip_capi
    use fortfem_hip_runtime, only: hip_context
    ! the reference's option and statistics types, the dense solvers and
    ! their tangents, and the preconditioner names stay those of the CPU
    ! module (src/solvers/advanced_solvers.f90:21-29); solve_sparse moves to
    ! the GPU for the methods and preconditioners it has there
    use fortfem_advanced_solvers_cpu, only: solver_options_t, solver_stats_t, solve, &
        solver_options, cg_solve, pcg_solve, bicgstab_solve, gmres_solve, cg_solve_jvp, &
        cg_solve_vjp, pcg_solve_jvp, pcg_solve_vjp, bicgstab_solve_jvp, bicgstab_solve_vjp, &
        gmres_solve_jvp, gmres_solve_vjp, jacobi_preconditioner, ilu_preconditioner, &
        ichol_preconditioner, ichol_controlled_preconditioner, &
        cpu_solve_sparse => solve_sparse
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
    ! additions of the GPU path
    public :: solve_sparse_jvp, solve_sparse_vjp
    public :: set_hip_solver_mode, to_c_options, solve_sparse_on_gpu

    ! GPU solve settings outside solver_options_t (which stays the
    ! reference's): precision "fp64" (the reference's recurrence) or "mixed"
    ! (FP32 iterations, FP64 x, reductions and residual refreshes); breakdown
    ! "absolute" (the reference's |p.Ap| < 1e-14, :897) or "relative"
    ! (|p.Ap| <= 1e-14 |r.z|)
    character(len=16), save :: hip_precision = "fp64"
    character(len=16), save :: hip_breakdown = "absolute"

contains

    subroutine set_hip_solver_mode(precision, breakdown)
        character(len=*), intent(in), optional :: precision, breakdown

        if (present(precision)) then
            if (trim(precision) /= "fp64" .and. trim(precision) /= "mixed") &
                error stop "set_hip_solver_mode: precision must be fp64 or mixed"
            hip_precision = precision
        end if
        if (present(breakdown)) then
            if (trim(breakdown) /= "absolute" .and. trim(breakdown) /= "relative") &
                error stop "set_hip_solver_mode: breakdown must be absolute or relative"
            hip_breakdown = breakdown
        end if
    end subroutine set_hip_solver_mode

    ! The methods and preconditioners the GPU path runs: solve_sparse's
    ! 'auto' -> 'pcg' and its default branch (any other method name runs
    ! pcg, :159-161, :183-186), with no or the Jacobi preconditioner.
    ! gmres, bicgstab and the ILU / IC preconditioners stay on the CPU.
    logical function solve_sparse_on_gpu(opts)
        type(solver_options_t), intent(in) :: opts

        select case (trim(opts%method))
        case ("gmres", "bicgstab")
            solve_sparse_on_gpu = .false.
        case default
            solve_sparse_on_gpu = trim(opts%preconditioner) == "none" .or. &
                trim(opts%preconditioner) == "jacobi"
        end select
    end function solve_sparse_on_gpu

    function to_c_options(opts) result(c)
        type(solver_options_t), intent(in) :: opts
        type(ffh_solver_options) :: c
        integer(c_int) :: code

        code = ffh_solver_options_default(c)
        c%method = 1
        if (trim(opts%method) == "cg") c%method = 0
        c%preconditioner = 0
        if (trim(opts%preconditioner) == "jacobi") c%preconditioner = 1
        c%tolerance_type = 0
        if (trim(opts%tolerance_type) == "absolute") c%tolerance_type = 1
        c%max_iterations = int(opts%max_iterations, c_int)
        c%tolerance = opts%tolerance
        c%precision = FFH_FP64
        if (trim(hip_precision) == "mixed") c%precision = FFH_MIXED
        c%breakdown_mode = 0
        if (trim(hip_breakdown) == "relative") c%breakdown_mode = 1
    end function to_c_options

    ! :135-195.  Size errors stop the program as in the reference.
    subroutine solve_sparse(A_sparse, b, x, opts, stats)
        type(sparse_matrix_t), intent(in) :: A_sparse
        real(dp), intent(in) :: b(:)
        real(dp), intent(inout) :: x(:)
        type(solver_options_t), intent(in) :: opts
        type(solver_stats_t), intent(out) :: stats

        type(c_ptr) :: mat
        type(ffh_solver_stats) :: cst
        integer(c_int) :: code

        if (A_sparse%ncol /= A_sparse%nrow) then
            error stop "solve_sparse: only square matrices are supported"
        end if
        if (size(b) /= A_sparse%nrow .or. size(x) /= A_sparse%nrow) then
            error stop "solve_sparse: incompatible vector size"
        end if
        if (.not. solve_sparse_on_gpu(opts)) then
            call cpu_solve_sparse(A_sparse, b, x, opts, stats)
            return
        end if
        code = ffh_matrix_upload(hip_context(), int(A_sparse%ncol, c_int), &
            int(A_sparse%nnz, c_int64_t), 1_c_int, A_sparse%col_ptr, &
            A_sparse%row_idx, A_sparse%val, mat)
        if (code /= FFH_OK) error stop "solve_sparse: device upload failed"
        code = ffh_solve_sparse(mat, b, x, to_c_options(opts), cst)
        if (code /= FFH_OK) error stop "solve_sparse: device solve failed"
        code = ffh_matrix_free(mat)
        stats%converged = cst%converged /= 0
        stats%iterations = cst%iterations
        stats%final_residual = cst%final_residual
        stats%solve_time = cst%solve_time
        stats%memory_usage = int(min(huge(1_c_int64_t), &
            int(A_sparse%nnz, c_int64_t)*8 + int(A_sparse%nrow, c_int64_t)*80))
        stats%method_used = "pcg"
        if (trim(opts%method) == "cg") stats%method_used = "cg"
    end subroutine solve_sparse

    ! Sparse counterpart of pcg_solve_jvp / cg_solve_jvp (:387-419, 458-489)
    ! for a converged x = A^-1 b: x_dot solves A x_dot = b_dot - A_dot x with
    ! A_dot given by its CSC values on A's pattern.  status 0, 1 (invalid
    ! input), 2 (tangent solve not converged).  b is the primal right-hand
    ! side, kept for the converged-state contract; it is not read.
    subroutine solve_sparse_jvp(A_sparse, b, x, opts, A_values_dot, b_dot, x_dot, status)
        type(sparse_matrix_t), intent(in) :: A_sparse
        real(dp), intent(in) :: b(:), x(:), A_values_dot(:), b_dot(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: x_dot(:)
        integer, intent(out) :: status

        type(c_ptr) :: mat
        type(ffh_solver_stats) :: cst
        integer(c_int) :: code, freed
        integer :: n

        x_dot = 0.0_dp
        status = 1
        n = A_sparse%nrow
        if (A_sparse%ncol /= n .or. size(b) /= n .or. size(x) /= n .or. &
            size(b_dot) /= n .or. size(x_dot) /= n .or. &
            size(A_values_dot) /= A_sparse%nnz) return
        if (opts%max_iterations < 1 .or. opts%tolerance <= 0.0_dp) return
        if (.not. solve_sparse_on_gpu(opts)) return ! CG / PCG with none or Jacobi only
        code = ffh_matrix_upload(hip_context(), int(n, c_int), &
            int(A_sparse%nnz, c_int64_t), 1_c_int, A_sparse%col_ptr, &
            A_sparse%row_idx, A_sparse%val, mat)
        if (code /= FFH_OK) return
        code = ffh_solve_sparse_jvp(mat, x, A_values_dot, b_dot, to_c_options(opts), &
            x_dot, cst)
        freed = ffh_matrix_free(mat)
        if (code /= FFH_OK) then
            x_dot = 0.0_dp
            return
        end if
        status = 0
        if (cst%converged == 0) status = 2
    end subroutine solve_sparse_jvp

    ! Sparse counterpart of pcg_solve_vjp / cg_solve_vjp (:421-456, 491-522):
    ! lambda solves A^T lambda = x_bar; b_bar = lambda and A_values_bar(k) =
    ! -lambda(row_k) x(col_k), the reference's A_bar = -lambda x^T on A's
    ! stored entries.  Zeros and status 2 when the adjoint solve does not
    ! converge.
    subroutine solve_sparse_vjp(A_sparse, b, x, opts, x_bar, A_values_bar, b_bar, status)
        type(sparse_matrix_t), intent(in) :: A_sparse
        real(dp), intent(in) :: b(:), x(:), x_bar(:)
        type(solver_options_t), intent(in) :: opts
        real(dp), intent(out) :: A_values_bar(:), b_bar(:)
        integer, intent(out) :: status

        type(c_ptr) :: mat
        type(ffh_solver_stats) :: cst
        integer(c_int) :: code, freed
        integer :: n

        A_values_bar = 0.0_dp
        b_bar = 0.0_dp
        status = 1
        n = A_sparse%nrow
        if (A_sparse%ncol /= n .or. size(b) /= n .or. size(x) /= n .or. &
            size(x_bar) /= n .or. size(b_bar) /= n .or. &
            size(A_values_bar) /= A_sparse%nnz) return
        if (opts%max_iterations < 1 .or. opts%tolerance <= 0.0_dp) return
        if (.not. solve_sparse_on_gpu(opts)) return ! CG / PCG with none or Jacobi only
        code = ffh_matrix_upload(hip_context(), int(n, c_int), &
            int(A_sparse%nnz, c_int64_t), 1_c_int, A_sparse%col_ptr, &
            A_sparse%row_idx, A_sparse%val, mat)
        if (code /= FFH_OK) return
        code = ffh_solve_sparse_vjp(mat, x, x_bar, to_c_options(opts), A_values_bar, &
            b_bar, cst)
        freed = ffh_matrix_free(mat)
        if (code /= FFH_OK) then
            A_values_bar = 0.0_dp
            b_bar = 0.0_dp
            return
        end if
        status = 0
        if (cst%converged == 0) then
            A_values_bar = 0.0_dp
            b_bar = 0.0_dp
            status = 2
        end if
    end subroutine solve_sparse_vjp

end module fortfem_advanced_solvers

module fortfem_tetra_lagrange_solver_3d
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, &
        c_null_ptr, c_loc
    use fortfem_hip_types, only: dp, fortsparse_status_t, status_set, &
        FORTSPARSE_INVALID_MATRIX, FORTSPARSE_INTERNAL_ERROR
    use fortfem_hip_capi
    use fortfem_hip_runtime, only: hip_context
    use fortfem_assembly_tetra_lagrange_arbitrary_order_3d, only: scalar_source_3d
    use fortfem_tetra_duffy_quadrature, only: tetra_duffy_quadrature
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, to_c_options
    ! the solution evaluators stay on the CPU
    use fortfem_tetra_lagrange_solver_3d_cpu, only: assignment(=), &
        evaluate_tetra_lagrange_solution, evaluate_tetra_lagrange_solution_jvp, &
        evaluate_tetra_lagrange_solution_vjp, evaluate_tetra_lagrange_solution_at_point, &
        evaluate_tetra_lagrange_solution_at_point_jvp, &
        evaluate_tetra_lagrange_solution_at_point_vjp, &
        evaluate_tetra_lagrange_solution_prepared, &
        initialize_tetra_lagrange_solution_evaluator, tetra_lagrange_solution_evaluator_t
    implicit none
    private

    public :: solve_tetra_lagrange_poisson, solve_tetra_lagrange_diffusion_reaction
    public :: assignment(=)
    public :: evaluate_tetra_lagrange_solution, evaluate_tetra_lagrange_solution_jvp
    public :: evaluate_tetra_lagrange_solution_vjp, evaluate_tetra_lagrange_solution_at_point
    public :: evaluate_tetra_lagrange_solution_at_point_jvp
    public :: evaluate_tetra_lagrange_solution_at_point_vjp
    public :: evaluate_tetra_lagrange_solution_prepared
    public :: initialize_tetra_lagrange_solution_evaluator
    public :: tetra_lagrange_solution_evaluator_t

contains

    ! src/solvers/fortfem_tetra_lagrange_solver_3d.f90:50-62
    subroutine solve_tetra_lagrange_poisson(vertices, tetrahedra, degree, &
            source, boundary_value, solution, status, options, stats, &
            cell_coefficient)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree
        procedure(scalar_source_3d) :: source, boundary_value
        real(dp), allocatable, intent(out) :: solution(:)
        type(fortsparse_status_t), intent(out) :: status
        type(solver_options_t), intent(in), optional :: options
        type(solver_stats_t), intent(out), optional :: stats
        real(dp), intent(in), optional :: cell_coefficient(:)

        call solve_tetra_lagrange_diffusion_reaction(vertices, tetrahedra, &
            degree, source, 1.0_dp, 0.0_dp, solution, status, boundary_value, &
            options, stats, cell_coefficient)
    end subroutine solve_tetra_lagrange_poisson

    ! :64-129 with the sparse direct solve replaced by Jacobi-PCG on the
    ! reduced system (default: relative 1e-10, scale-invariant breakdown).
    subroutine solve_tetra_lagrange_diffusion_reaction(vertices, tetrahedra, &
            degree, source, stiffness_coefficient, mass_coefficient, solution, &
            status, boundary_value, options, stats, cell_coefficient)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree
        procedure(scalar_source_3d) :: source
        real(dp), intent(in) :: stiffness_coefficient, mass_coefficient
        real(dp), allocatable, intent(out) :: solution(:)
        type(fortsparse_status_t), intent(out) :: status
        procedure(scalar_source_3d), optional :: boundary_value
        type(solver_options_t), intent(in), optional :: options
        type(solver_stats_t), intent(out), optional :: stats
        real(dp), intent(in), optional :: cell_coefficient(:)

        type(c_ptr) :: prob
        type(solver_options_t) :: opts
        type(ffh_solver_options) :: copts
        type(ffh_solver_stats) :: cst
        integer(c_int) :: n_global, code, q
        integer(c_int64_t) :: nnz
        real(c_double) :: kappa(1)
        real(dp), allocatable, target :: samples(:, :), values(:)

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral H1 diffusion-reaction solve failed")
        if (degree < 1 .or. degree > 4) return
        if (stiffness_coefficient < 0.0_dp .or. mass_coefficient < 0.0_dp) return
        if (stiffness_coefficient == 0.0_dp .and. mass_coefficient == 0.0_dp) return
        if (.not. present(boundary_value) .and. mass_coefficient == 0.0_dp) return
        q = int(2*degree + 4, c_int)
        if (.not. sample_points(vertices, tetrahedra, int(q), source, samples)) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(vertices, 2), c_int), vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_H1, int(degree, c_int), &
            n_global, nnz)
        kappa = stiffness_coefficient
        if (code == FFH_OK) then
            if (present(cell_coefficient)) then
                code = ffh_assemble(prob, q, FFH_COEFF_CELL, &
                    cell_coefficient*stiffness_coefficient, mass_coefficient, FFH_FP64)
            else
                code = ffh_assemble(prob, q, FFH_COEFF_SCALAR, kappa, &
                    mass_coefficient, FFH_FP64)
            end if
        end if
        if (code == FFH_OK) code = ffh_load(prob, q, FFH_SOURCE_SAMPLES, samples, &
            int(size(samples), c_int64_t))
        if (code == FFH_OK .and. present(boundary_value)) then
            allocate(values(n_global))
            call node_values(vertices, tetrahedra, degree, prob, boundary_value, values)
            code = ffh_dirichlet(prob, 0_c_int, c_loc(values))
        end if
        opts%method = "pcg"
        opts%preconditioner = "jacobi"
        opts%tolerance = 1.0e-10_dp
        opts%max_iterations = 100000
        if (present(options)) opts = options
        if (code == FFH_OK) then
            allocate(solution(n_global))
            ! the scale-invariant breakdown test: these drivers replace a
            ! direct solve, and the absolute one fires early on fine meshes
            copts = to_c_options(opts)
            copts%breakdown_mode = 1
            code = ffh_problem_solve(prob, copts, solution, cst)
        end if
        if (present(stats)) then
            stats%converged = cst%converged /= 0
            stats%iterations = cst%iterations
            stats%final_residual = cst%final_residual
            stats%solve_time = cst%solve_time
            stats%method_used = "pcg"
        end if
        if (code == FFH_OK) then
            call status_set(status, 0, "")
        else
            call status_set(status, FORTSPARSE_INTERNAL_ERROR, &
                "Tetrahedral H1 device solve failed")
        end if
        code = ffh_problem_free(prob)
    end subroutine solve_tetra_lagrange_diffusion_reaction

    logical function sample_points(vertices, tetrahedra, qdeg, source, samples) &
            result(ok)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), qdeg
        procedure(scalar_source_3d) :: source
        real(dp), allocatable, intent(out) :: samples(:, :)

        real(dp), allocatable :: x(:), y(:), z(:), w(:)
        real(dp) :: jac(3, 3), point(3)
        integer :: cell, q, st

        ok = .false.
        if (size(vertices, 1) /= 3 .or. size(tetrahedra, 1) /= 4) return
        if (any(tetrahedra < 1) .or. any(tetrahedra > size(vertices, 2))) return
        call tetra_duffy_quadrature(qdeg, x, y, z, w, st)
        if (st /= 0) return
        allocate(samples(size(w), size(tetrahedra, 2)))
        do cell = 1, size(tetrahedra, 2)
            jac(:, 1) = vertices(:, tetrahedra(2, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 2) = vertices(:, tetrahedra(3, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 3) = vertices(:, tetrahedra(4, cell)) - vertices(:, tetrahedra(1, cell))
            do q = 1, size(w)
                point = vertices(:, tetrahedra(1, cell)) + matmul(jac, [x(q), y(q), z(q)])
                call source(point(1), point(2), point(3), samples(q, cell))
            end do
        end do
        ok = .true.
    end function sample_points

    ! g at the Lagrange node v_1 + J node of every dof (build_dirichlet_data,
    ! :598-606); only constrained entries are read by the library.
    subroutine node_values(vertices, tetrahedra, degree, prob, g, values)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), degree
        type(c_ptr), intent(in) :: prob
        procedure(scalar_source_3d) :: g
        real(dp), intent(out) :: values(:)

        integer, allocatable :: gd(:, :)
        integer :: n_local, cell, a, b, c, l, code
        real(dp) :: node(3), jac(3, 3), point(3)

        n_local = (degree + 1)*(degree + 2)*(degree + 3)/6
        allocate(gd(n_local, size(tetrahedra, 2)))
        code = ffh_global_dofs_download(prob, 1_c_int, gd)
        values = 0.0_dp
        do cell = 1, size(tetrahedra, 2)
            jac(:, 1) = vertices(:, tetrahedra(2, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 2) = vertices(:, tetrahedra(3, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 3) = vertices(:, tetrahedra(4, cell)) - vertices(:, tetrahedra(1, cell))
            l = 0
            do a = 0, degree
                do b = 0, degree - a
                    do c = 0, degree - a - b
                        l = l + 1
                        node = real([b, c, degree - a - b - c], dp)/real(degree, dp)
                        point = vertices(:, tetrahedra(1, cell)) + matmul(jac, node)
                        call g(point(1), point(2), point(3), values(gd(l, cell)))
                    end do
                end do
            end do
        end do
    end subroutine node_values

end module fortfem_tetra_lagrange_solver_3d

! src/solvers/fortfem_tetra_nedelec_solver_3d.f90: the curl-mass drivers with
! the sparse direct solve replaced by Jacobi-PCG (relative 1e-12 unless
! options says otherwise); zero_tangential_boundary eliminates the boundary
! edge and face dofs (build_zero_tangential_mask, :416-470) with zero values.
module fortfem_tetra_nedelec_solver_3d
    use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, &
        c_null_ptr
    use fortfem_hip_types, only: dp, fortsparse_status_t, status_set, &
        FORTSPARSE_INVALID_MATRIX, FORTSPARSE_INTERNAL_ERROR
    use fortfem_hip_capi
    use fortfem_hip_runtime, only: hip_context
    use fortfem_assembly_tetra_nedelec_3d, only: tensor_coefficient_3d, &
        vector_source_3d, sample_vector_source, sample_rule_tensors
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, to_c_options
    ! the PML solves stay on the CPU
    use fortfem_tetra_nedelec_solver_3d_cpu, only: solve_tetra_nedelec_pml, &
        solve_tetra_nedelec_curvilinear_pml
    implicit none
    private

    public :: solve_tetra_nedelec_curl_mass, solve_tetra_nedelec_weighted_curl_mass
    public :: solve_tetra_nedelec_pml, solve_tetra_nedelec_curvilinear_pml

contains

    ! :361-414
    subroutine solve_tetra_nedelec_curl_mass(vertices, tetrahedra, order, &
            source, curl_coefficient, mass_coefficient, solution, status, &
            zero_tangential_boundary, options, stats)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(vector_source_3d) :: source
        real(dp), intent(in) :: curl_coefficient, mass_coefficient
        real(dp), allocatable, intent(out) :: solution(:)
        type(fortsparse_status_t), intent(out) :: status
        logical, intent(in), optional :: zero_tangential_boundary
        type(solver_options_t), intent(in), optional :: options
        type(solver_stats_t), intent(out), optional :: stats

        real(c_double) :: alpha(1)
        type(solver_stats_t) :: st

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Tetrahedral Nedelec curl-mass solve failed")
        if (order < 1 .or. mass_coefficient <= 0.0_dp) return
        alpha = curl_coefficient
        call nedelec_solve(vertices, tetrahedra, order, source, FFH_COEFF_SCALAR, &
            alpha, mass_coefficient, solution, status, zero_tangential_boundary, &
            options, st)
        if (present(stats)) stats = st
    end subroutine solve_tetra_nedelec_curl_mass

    ! :41-95.  The tensor is evaluated at the rule's points as in
    ! assemble_tetra_nedelec_weighted_csc.
    subroutine solve_tetra_nedelec_weighted_curl_mass(vertices, tetrahedra, &
            order, curl_coefficient, source, mass_coefficient, solution, status, &
            zero_tangential_boundary, options, stats)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(tensor_coefficient_3d) :: curl_coefficient
        procedure(vector_source_3d) :: source
        real(dp), intent(in) :: mass_coefficient
        real(dp), allocatable, intent(out) :: solution(:)
        type(fortsparse_status_t), intent(out) :: status
        logical, intent(in), optional :: zero_tangential_boundary
        type(solver_options_t), intent(in), optional :: options
        type(solver_stats_t), intent(out), optional :: stats

        real(c_double), allocatable :: values(:)
        integer(c_int) :: kind
        type(solver_stats_t) :: st

        call status_set(status, FORTSPARSE_INVALID_MATRIX, &
            "Weighted tetrahedral Nedelec curl-mass solve failed")
        if (order < 1 .or. mass_coefficient <= 0.0_dp) return
        if (.not. sample_rule_tensors(vertices, tetrahedra, order, &
            curl_coefficient, values, kind)) return
        call nedelec_solve(vertices, tetrahedra, order, source, &
            kind, values, mass_coefficient, solution, status, &
            zero_tangential_boundary, options, st)
        if (present(stats)) stats = st
    end subroutine solve_tetra_nedelec_weighted_curl_mass

    subroutine nedelec_solve(vertices, tetrahedra, order, source, kind, coeff, &
            mass, solution, status, zero_tangential_boundary, options, stats)
        real(dp), intent(in) :: vertices(:, :)
        integer, intent(in) :: tetrahedra(:, :), order
        procedure(vector_source_3d) :: source
        integer(c_int), intent(in) :: kind
        real(c_double), intent(in) :: coeff(*)
        real(dp), intent(in) :: mass
        real(dp), allocatable, intent(out) :: solution(:)
        type(fortsparse_status_t), intent(inout) :: status
        logical, intent(in), optional :: zero_tangential_boundary
        type(solver_options_t), intent(in), optional :: options
        type(solver_stats_t), intent(out) :: stats

        real(c_double), allocatable :: samples(:, :, :)
        type(solver_options_t) :: opts
        type(ffh_solver_options) :: copts
        type(ffh_solver_stats) :: cst
        type(c_ptr) :: prob
        integer(c_int) :: n_global, code
        integer(c_int64_t) :: nnz

        if (.not. sample_vector_source(vertices, tetrahedra, order, source, &
            samples)) return
        code = ffh_problem_create(hip_context(), 3_c_int, &
            int(size(vertices, 2), c_int), vertices, &
            int(size(tetrahedra, 2), c_int), tetrahedra, prob)
        if (code /= FFH_OK) return
        code = ffh_pattern_build(prob, FFH_FAMILY_NEDELEC, int(order, c_int), &
            n_global, nnz)
        if (code == FFH_OK) code = ffh_assemble(prob, -1_c_int, kind, coeff, &
            real(mass, c_double), FFH_FP64)
        if (code == FFH_OK) code = ffh_load(prob, -1_c_int, FFH_SOURCE_SAMPLES, &
            samples, int(size(samples), c_int64_t))
        if (code == FFH_OK .and. present(zero_tangential_boundary)) then
            if (zero_tangential_boundary) code = ffh_dirichlet(prob, 0_c_int, c_null_ptr)
        end if
        opts%method = "pcg"
        opts%preconditioner = "jacobi"
        opts%tolerance = 1.0e-12_dp
        opts%max_iterations = 100000
        if (present(options)) opts = options
        if (code == FFH_OK) then
            allocate(solution(n_global))
            ! the scale-invariant breakdown test (a direct solve is replaced)
            copts = to_c_options(opts)
            copts%breakdown_mode = 1
            code = ffh_problem_solve(prob, copts, solution, cst)
        end if
        stats%converged = cst%converged /= 0
        stats%iterations = cst%iterations
        stats%final_residual = cst%final_residual
        stats%solve_time = cst%solve_time
        stats%method_used = "pcg"
        if (code == FFH_OK) then
            call status_set(status, 0, "")
        else
            call status_set(status, FORTSPARSE_INTERNAL_ERROR, &
                "Tetrahedral Nedelec device solve failed")
        end if
        code = ffh_problem_free(prob)
    end subroutine nedelec_solve

end module fortfem_tetra_nedelec_solver_3d
