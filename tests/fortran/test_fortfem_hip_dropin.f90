! Known-answer tests of the drop-in Fortran modules, written like the
! reference's own test programs (check + error stop 1 on failure).
! Restated checks: test/test_structured_tetra_box_mesh.f90:19-51,
! test/test_tetra_lagrange_arbitrary_order_assembly.f90:37-66,
! test/test_advanced_solvers.f90:39-84,342-348,
! test/test_tetra_lagrange_poisson_p_convergence.f90:26-40,
! test/test_tetra_nedelec_sparse_assembly_slow.f90:31-59,
! test/test_sparse_matrix_ad.f90:19-65, test/test_pcg_solver_ad.f90:18-60,
! test/test_tetra_lagrange_stiffness_csc_ad.f90:10-93,
! test/test_tetra_nedelec_curl_mass_csc_ad.f90:10-96,
! test/test_public_tetra_nedelec_arbitrary_order_solve_slow.f90:14-46.
module dropin_test_sources
    use fortfem_hip_types, only: dp
    implicit none
contains

    pure subroutine bubble_source(x, y, z, value)
        real(dp), intent(in) :: x, y, z
        real(dp), intent(out) :: value
        value = 2.0_dp*(x*y + x*z + y*z)
    end subroutine bubble_source

    pure subroutine zero_boundary(x, y, z, value)
        real(dp), intent(in) :: x, y, z
        real(dp), intent(out) :: value
        value = 0.0_dp*(x + y + z)
    end subroutine zero_boundary

    pure subroutine constant_field(x, y, z, value)
        real(dp), intent(in) :: x, y, z
        real(dp), intent(out) :: value(3)
        value = [0.7_dp, -1.3_dp, 0.4_dp] + 0.0_dp*(x + y + z)
    end subroutine constant_field

    pure subroutine sin_source(x, y, z, value)
        real(dp), intent(in) :: x, y, z
        real(dp), intent(out) :: value
        real(dp), parameter :: pi = 3.14159265358979323846_dp
        value = 3.0_dp*pi*pi*sin(pi*x)*sin(pi*y)*sin(pi*z)
    end subroutine sin_source

    ! a symmetric tensor field linear in x: K0 + x K1
    pure subroutine linear_tensor(x, y, z, value)
        real(dp), intent(in) :: x, y, z
        real(dp), intent(out) :: value(3, 3)
        value = reshape([2.0_dp, 0.3_dp, 0.1_dp, 0.3_dp, 1.5_dp, -0.2_dp, &
            0.1_dp, -0.2_dp, 1.0_dp], [3, 3]) + x*reshape([0.8_dp, 0.2_dp, 0.0_dp, &
            0.2_dp, -0.4_dp, 0.1_dp, 0.0_dp, 0.1_dp, 0.6_dp], [3, 3]) + 0.0_dp*(y + z)
    end subroutine linear_tensor

end module dropin_test_sources

program test_fortfem_hip_dropin
    use fortfem_hip_types, only: dp, csc_t, csc_z_t, fortsparse_status_t
    use fortfem_hip_runtime, only: fortfem_hip_finalize
    use fortfem_structured_tetra_box_mesh, only: generate_structured_tetra_box_mesh
    use fortfem_assembly_tetra_lagrange_arbitrary_order_3d, only: &
        assemble_tetra_lagrange_stiffness_csc, &
        assemble_tetra_lagrange_stiffness_csc_jvp, &
        assemble_tetra_lagrange_stiffness_csc_vjp
    use fortfem_assembly_tetra_nedelec_3d, only: assemble_tetra_nedelec_curl_mass_csc, &
        assemble_tetra_nedelec_vector_load, assemble_tetra_nedelec_curl_mass_csc_jvp, &
        assemble_tetra_nedelec_curl_mass_csc_vjp, assemble_tetra_nedelec_pml_csc, &
        assemble_tetra_nedelec_curvilinear_pml_csc, assemble_tetra_nedelec_pml_csc_jvp, &
        assemble_tetra_nedelec_pml_csc_vjp, assemble_tetra_nedelec_weighted_csc
    use fortfem_sparse_matrix, only: spmv, spmv_jvp, spmv_vjp
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, &
        solver_options, solve_sparse, solve_sparse_jvp, solve_sparse_vjp
    use fortfem_tetra_lagrange_solver_3d, only: solve_tetra_lagrange_poisson
    use fortfem_tetra_nedelec_solver_3d, only: solve_tetra_nedelec_curl_mass
    use fortfem_hip_capi, only: ffh_nedelec_dof_map
    use dropin_test_sources, only: bubble_source, zero_boundary, sin_source, constant_field, &
        linear_tensor
    implicit none

    logical :: all_passed

    all_passed = .true.
    call check_box_mesh()
    call check_linear_energy()
    call check_cg_tridiagonal()
    call check_quartic_p4()
    call check_p2_box_poisson()
    call check_nedelec_rotational_energy()
    call check_nedelec_point_tensor_energy()
    call check_nedelec_constant_load()
    call check_spmv_ad()
    call check_pcg_ad()
    call check_h1_csc_ad()
    call check_nedelec_csc_ad()
    call check_nedelec_constant_solve()
    call check_nedelec_pml()
    call check_nedelec_pml_ad()
    call fortfem_hip_finalize()
    if (.not. all_passed) error stop 1
    write (*, '(a)') "drop-in Fortran checks passed"

contains

    subroutine record(condition, description)
        use, intrinsic :: iso_fortran_env, only: output_unit
        logical, intent(in) :: condition
        character(len=*), intent(in) :: description

        if (condition) then
            write (*, '(2a)') "PASS ", description
        else
            write (*, '(2a)') "FAIL ", description
            all_passed = .false.
        end if
        flush (output_unit)
    end subroutine record

    real(dp) function det3(a)
        real(dp), intent(in) :: a(3, 3)
        det3 = a(1, 1)*(a(2, 2)*a(3, 3) - a(2, 3)*a(3, 2)) - &
            a(1, 2)*(a(2, 1)*a(3, 3) - a(2, 3)*a(3, 1)) + &
            a(1, 3)*(a(2, 1)*a(3, 2) - a(2, 2)*a(3, 1))
    end function det3

    subroutine check_box_mesh()
        real(dp), parameter :: bounds(3, 2) = reshape([-1.0_dp, 2.0_dp, 4.0_dp, &
            1.0_dp, 5.0_dp, 10.0_dp], [3, 2])
        integer, parameter :: counts(3) = [2, 1, 3]
        real(dp), allocatable :: vertices(:, :)
        integer, allocatable :: tetrahedra(:, :)
        real(dp) :: jac(3, 3), volume
        integer :: cell, status
        logical :: positive

        call generate_structured_tetra_box_mesh(bounds, counts, vertices, &
            tetrahedra, status)
        call record(status == 0, "box mesh generation succeeds")
        call record(size(vertices, 2) == product(counts + 1), "tensor-product vertex count")
        call record(size(tetrahedra, 2) == 6*product(counts), "six tetrahedra per brick")
        call record(all(abs(minval(vertices, dim=2) - bounds(:, 1)) < 1.0e-14_dp) .and. &
            all(abs(maxval(vertices, dim=2) - bounds(:, 2)) < 1.0e-14_dp), "bounds reached")
        volume = 0.0_dp
        positive = .true.
        do cell = 1, size(tetrahedra, 2)
            jac(:, 1) = vertices(:, tetrahedra(2, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 2) = vertices(:, tetrahedra(3, cell)) - vertices(:, tetrahedra(1, cell))
            jac(:, 3) = vertices(:, tetrahedra(4, cell)) - vertices(:, tetrahedra(1, cell))
            positive = positive .and. det3(jac) > 0.0_dp
            volume = volume + det3(jac)/6.0_dp
        end do
        call record(positive, "every tetrahedron positively oriented")
        call record(abs(volume - product(bounds(:, 2) - bounds(:, 1))) < 1.0e-12_dp, &
            "exact volume partition")
    end subroutine check_box_mesh

    ! test_tetra_nedelec_sparse_assembly_slow.f90: the Whitney interpolant of
    ! the rotational field 0.5 c x x (curl = c) on the two-tet mesh has
    ! alpha |c|^2 |Omega| = 2 * 14 / 3 = 28/3 curl energy.
    subroutine check_nedelec_rotational_energy()
        real(dp) :: mesh_vertices(3, 5), c(3), mid(3), tangent(3), field(3)
        real(dp), allocatable :: dofs(:), product(:)
        integer :: tetrahedra(4, 2), edges(2, 12), cell, l, a, b, e, nedges
        integer, parameter :: local(2, 6) = reshape([1, 2, 1, 3, 1, 4, 2, 3, 2, 4, 3, 4], [2, 6])
        logical :: found
        type(csc_t) :: matrix
        type(fortsparse_status_t) :: status

        mesh_vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, -1.0_dp], [3, 5])
        tetrahedra(:, 1) = [1, 2, 3, 4]
        tetrahedra(:, 2) = [1, 3, 2, 5]
        call assemble_tetra_nedelec_curl_mass_csc(mesh_vertices, tetrahedra, matrix, &
            status, curl_coefficient=2.0_dp, mass_coefficient=0.0_dp, order=1)
        call record(status%code == 0 .and. matrix%nrow == 9, "Nedelec k=1 CSC assembly succeeds")
        ! edges by first appearance, (min, max)
        nedges = 0
        do cell = 1, 2
            do l = 1, 6
                a = minval(tetrahedra(local(:, l), cell))
                b = maxval(tetrahedra(local(:, l), cell))
                found = .false.
                do e = 1, nedges
                    if (edges(1, e) == a .and. edges(2, e) == b) found = .true.
                end do
                if (.not. found) then
                    nedges = nedges + 1
                    edges(:, nedges) = [a, b]
                end if
            end do
        end do
        allocate(dofs(nedges), product(nedges))
        c = [1.0_dp, -2.0_dp, 3.0_dp]
        do e = 1, nedges
            mid = 0.5_dp*(mesh_vertices(:, edges(1, e)) + mesh_vertices(:, edges(2, e)))
            tangent = mesh_vertices(:, edges(2, e)) - mesh_vertices(:, edges(1, e))
            field = 0.5_dp*[c(2)*mid(3) - c(3)*mid(2), c(3)*mid(1) - c(1)*mid(3), &
                c(1)*mid(2) - c(2)*mid(1)]
            dofs(e) = dot_product(field, tangent)
        end do
        call spmv(matrix, dofs, product)
        call record(abs(dot_product(dofs, product) - 28.0_dp/3.0_dp) < 3.0e-12_dp, &
            "Nedelec rotational-field energy is 28/3")
    end subroutine check_nedelec_rotational_energy

    ! assemble_tetra_nedelec_weighted_csc (:1870-2000) with a tensor varying
    ! inside the cells, K(x) = K0 + x K1: the callback is evaluated at the
    ! rule's points (per-point element on MFMA).  The Whitney interpolant of
    ! 0.5 c x x has curl c on both cells of the two-tet mesh, and the 4-point
    ! rule integrates the linear field exactly, so the curl energy is
    ! c^T (|T1| K(centroid 1) + |T2| K(centroid 2)) c = c^T (K0 + K1 / 4) c / 3.
    subroutine check_nedelec_point_tensor_energy()
        real(dp) :: mesh_vertices(3, 5), c(3), mid(3), tangent(3), field(3)
        real(dp) :: k0(3, 3), k1(3, 3), exact
        real(dp), allocatable :: dofs(:), product(:)
        integer :: tetrahedra(4, 2), edges(2, 12), cell, l, a, b, e, nedges
        integer, parameter :: local(2, 6) = reshape([1, 2, 1, 3, 1, 4, 2, 3, 2, 4, 3, 4], [2, 6])
        logical :: found
        type(csc_t) :: matrix
        type(fortsparse_status_t) :: status

        mesh_vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, -1.0_dp], [3, 5])
        tetrahedra(:, 1) = [1, 2, 3, 4]
        tetrahedra(:, 2) = [1, 3, 2, 5]
        call assemble_tetra_nedelec_weighted_csc(mesh_vertices, tetrahedra, &
            linear_tensor, 0.0_dp, matrix, status, order=1)
        call record(status%code == 0 .and. matrix%nrow == 9, &
            "weighted Nedelec CSC with a per-point tensor succeeds")
        nedges = 0
        do cell = 1, 2
            do l = 1, 6
                a = minval(tetrahedra(local(:, l), cell))
                b = maxval(tetrahedra(local(:, l), cell))
                found = .false.
                do e = 1, nedges
                    if (edges(1, e) == a .and. edges(2, e) == b) found = .true.
                end do
                if (.not. found) then
                    nedges = nedges + 1
                    edges(:, nedges) = [a, b]
                end if
            end do
        end do
        allocate(dofs(nedges), product(nedges))
        c = [1.0_dp, -2.0_dp, 3.0_dp]
        do e = 1, nedges
            mid = 0.5_dp*(mesh_vertices(:, edges(1, e)) + mesh_vertices(:, edges(2, e)))
            tangent = mesh_vertices(:, edges(2, e)) - mesh_vertices(:, edges(1, e))
            field = 0.5_dp*[c(2)*mid(3) - c(3)*mid(2), c(3)*mid(1) - c(1)*mid(3), &
                c(1)*mid(2) - c(2)*mid(1)]
            dofs(e) = dot_product(field, tangent)
        end do
        call spmv(matrix, dofs, product)
        call linear_tensor(0.0_dp, 0.0_dp, 0.0_dp, k0)
        call linear_tensor(1.0_dp, 0.0_dp, 0.0_dp, k1)
        k1 = k1 - k0
        exact = dot_product(c, matmul(k0 + 0.25_dp*k1, c))/3.0_dp
        call record(abs(dot_product(dofs, product) - exact) < 1.0e-12_dp*abs(exact), &
            "per-point tensor curl energy is c^T (K0 + K1/4) c / 3")
    end subroutine check_nedelec_point_tensor_energy

    ! assemble_tetra_nedelec_vector_load (:2002-2040): a constant field f lies
    ! in the Whitney space with dofs d_e = f . (v_b - v_a) on the edges
    ! (a, b) = (min, max) by first appearance, so its load is M d with M the
    ! mass matrix (curl coefficient 0).
    subroutine check_nedelec_constant_load()
        real(dp) :: mesh_vertices(3, 5)
        real(dp), allocatable :: dofs(:), product(:), load(:)
        integer :: tetrahedra(4, 2), edges(2, 12), cell, l, a, b, e, nedges
        integer, parameter :: local(2, 6) = reshape([1, 2, 1, 3, 1, 4, 2, 3, 2, 4, 3, 4], [2, 6])
        logical :: found
        type(csc_t) :: mass
        type(fortsparse_status_t) :: status

        mesh_vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.1_dp, 0.0_dp, &
            0.2_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.3_dp, 1.0_dp, 0.1_dp, 0.0_dp, -1.0_dp], [3, 5])
        tetrahedra(:, 1) = [1, 2, 3, 4]
        tetrahedra(:, 2) = [1, 3, 2, 5]
        call assemble_tetra_nedelec_curl_mass_csc(mesh_vertices, tetrahedra, mass, &
            status, curl_coefficient=0.0_dp, mass_coefficient=1.0_dp, order=1)
        call record(status%code == 0, "Nedelec k=1 mass assembly succeeds")
        call assemble_tetra_nedelec_vector_load(mesh_vertices, tetrahedra, &
            constant_field, load, status)
        call record(status%code == 0 .and. size(load) == 9, "Nedelec k=1 vector load succeeds")
        nedges = 0
        do cell = 1, 2
            do l = 1, 6
                a = minval(tetrahedra(local(:, l), cell))
                b = maxval(tetrahedra(local(:, l), cell))
                found = .false.
                do e = 1, nedges
                    if (edges(1, e) == a .and. edges(2, e) == b) found = .true.
                end do
                if (.not. found) then
                    nedges = nedges + 1
                    edges(:, nedges) = [a, b]
                end if
            end do
        end do
        allocate(dofs(nedges), product(nedges))
        do e = 1, nedges
            dofs(e) = dot_product([0.7_dp, -1.3_dp, 0.4_dp], &
                mesh_vertices(:, edges(2, e)) - mesh_vertices(:, edges(1, e)))
        end do
        call spmv(mass, dofs, product)
        call record(maxval(abs(load - product)) < 1.0e-13_dp*maxval(abs(product)), &
            "Nedelec constant-field load equals the mass times the interpolant")
    end subroutine check_nedelec_constant_load

    ! a dense 3x3 matrix stored with all nine entries (column-major CSC)
    subroutine full_csc(dense, a)
        real(dp), intent(in) :: dense(3, 3)
        type(csc_t), intent(out) :: a
        integer :: c, r

        a%nrow = 3
        a%ncol = 3
        a%nnz = 9
        allocate(a%col_ptr(4), a%row_idx(9), a%val(9))
        do c = 1, 3
            a%col_ptr(c) = 3*(c - 1) + 1
            do r = 1, 3
                a%row_idx(3*(c - 1) + r) = r
                a%val(3*(c - 1) + r) = dense(r, c)
            end do
        end do
        a%col_ptr(4) = 10
    end subroutine full_csc

    subroutine check_spmv_ad()
        real(dp), parameter :: step = 2.0e-7_dp
        real(dp) :: matrix(3, 3), matrix_dot(3, 3)
        real(dp) :: vector(3), vector_dot(3), product_dot(3), plus(3), minus(3)
        real(dp) :: product_bar(3), vector_bar(3), values_bar(9), values_dot(9)
        real(dp) :: lhs, rhs, relative_error
        type(csc_t) :: a, ap, am

        matrix = reshape([4.0_dp, -1.0_dp, 0.5_dp, 0.7_dp, 3.0_dp, -0.4_dp, &
            -0.2_dp, 0.8_dp, 2.5_dp], shape(matrix))
        matrix_dot = reshape([0.12_dp, -0.03_dp, 0.05_dp, -0.07_dp, 0.09_dp, &
            0.02_dp, 0.04_dp, -0.06_dp, 0.08_dp], shape(matrix_dot))
        vector = [0.6_dp, -0.8_dp, 0.35_dp]
        vector_dot = [-0.11_dp, 0.07_dp, 0.09_dp]
        product_bar = [0.31_dp, -0.27_dp, 0.44_dp]
        call full_csc(matrix, a)
        values_dot = reshape(matrix_dot, [9])
        call spmv_jvp(a, vector, values_dot, vector_dot, product_dot)
        call full_csc(matrix + step*matrix_dot, ap)
        call full_csc(matrix - step*matrix_dot, am)
        call spmv(ap, vector + step*vector_dot, plus)
        call spmv(am, vector - step*vector_dot, minus)
        relative_error = maxval(abs(product_dot - (plus - minus)/(2.0_dp*step)))/ &
            max(1.0_dp, maxval(abs(product_dot)))
        call record(relative_error < 2.0e-8_dp, &
            "Sparse matrix JVP matches an independent re-evaluation")
        call spmv_vjp(a, vector, product_bar, values_bar, vector_bar)
        lhs = dot_product(product_bar, product_dot)
        rhs = dot_product(vector_bar, vector_dot) + dot_product(values_bar, values_dot)
        call record(abs(lhs - rhs) < 2.0e-10_dp*max(1.0_dp, abs(lhs), abs(rhs)), &
            "Sparse matrix VJP satisfies the adjoint identity")
    end subroutine check_spmv_ad

    subroutine check_pcg_ad()
        real(dp), parameter :: step = 2.0e-7_dp
        real(dp) :: matrix(3, 3), matrix_dot(3, 3), rhs(3), rhs_dot(3)
        real(dp) :: solution(3), solution_dot(3), plus(3), minus(3), solution_bar(3)
        real(dp) :: values_dot(9), values_bar(9), rhs_bar(3), lhs, identity
        real(dp) :: relative_error
        type(csc_t) :: a, ap, am
        type(solver_options_t) :: opts
        type(solver_stats_t) :: stats
        integer :: status

        matrix = reshape([5.0_dp, 1.0_dp, 0.0_dp, 1.0_dp, 4.0_dp, 0.3_dp, &
            0.0_dp, 0.3_dp, 2.5_dp], shape(matrix))
        matrix_dot = reshape([0.09_dp, -0.02_dp, 0.01_dp, -0.02_dp, 0.05_dp, &
            0.03_dp, 0.01_dp, 0.03_dp, -0.06_dp], shape(matrix_dot))
        rhs = [0.8_dp, -0.25_dp, 0.55_dp]
        rhs_dot = [-0.06_dp, 0.11_dp, 0.04_dp]
        solution_bar = [0.22_dp, -0.35_dp, 0.41_dp]
        opts = solver_options(method="pcg", tolerance=1.0e-12_dp, max_iterations=40, &
            preconditioner="jacobi", tolerance_type="absolute")
        call full_csc(matrix, a)
        values_dot = reshape(matrix_dot, [9])
        solution = 0.0_dp
        call solve_sparse(a, rhs, solution, opts, stats)
        call record(stats%converged, "PCG primal solve converges")
        call solve_sparse_jvp(a, rhs, solution, opts, values_dot, rhs_dot, &
            solution_dot, status)
        call full_csc(matrix + step*matrix_dot, ap)
        call full_csc(matrix - step*matrix_dot, am)
        plus = 0.0_dp
        call solve_sparse(ap, rhs + step*rhs_dot, plus, opts, stats)
        minus = 0.0_dp
        call solve_sparse(am, rhs - step*rhs_dot, minus, opts, stats)
        relative_error = maxval(abs(solution_dot - (plus - minus)/(2.0_dp*step)))/ &
            max(1.0_dp, maxval(abs(solution_dot)))
        call record(status == 0 .and. relative_error < 2.0e-7_dp, &
            "PCG JVP matches re-evaluation of the converged state")
        call solve_sparse_vjp(a, rhs, solution, opts, solution_bar, values_bar, &
            rhs_bar, status)
        lhs = dot_product(solution_bar, solution_dot)
        identity = dot_product(values_bar, values_dot) + dot_product(rhs_bar, rhs_dot)
        call record(status == 0 .and. abs(lhs - identity) < &
            3.0e-8_dp*max(1.0_dp, abs(lhs), abs(identity)), &
            "PCG VJP satisfies the implicit-solve adjoint identity")
    end subroutine check_pcg_ad

    subroutine check_h1_csc_ad()
        integer, parameter :: degree = 2, quadrature_degree = 6
        integer, parameter :: tets(4, 2) = reshape([1, 2, 3, 4, 2, 3, 4, 5], [4, 2])
        real(dp), parameter :: step = 2.0e-7_dp
        real(dp) :: vertices(3, 5), vertices_dot(3, 5), vertices_bar(3, 5)
        real(dp), allocatable :: values_bar(:)
        type(csc_t) :: matrix, matrix_dot, minus, plus
        type(fortsparse_status_t) :: st
        real(dp) :: kappa, mu, kappa_dot, mu_dot, kappa_bar, mu_bar
        real(dp) :: lhs, rhs, relative_error
        integer :: entry

        vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 1.0_dp, 1.0_dp, &
            1.0_dp], [3, 5])
        vertices_dot = reshape([0.01_dp, -0.02_dp, 0.015_dp, -0.03_dp, 0.01_dp, &
            0.02_dp, 0.025_dp, -0.015_dp, 0.005_dp, -0.01_dp, 0.03_dp, -0.02_dp, &
            0.02_dp, -0.01_dp, 0.025_dp], [3, 5])
        kappa = 1.7_dp
        mu = 0.8_dp
        kappa_dot = 0.13_dp
        mu_dot = -0.09_dp
        call assemble_tetra_lagrange_stiffness_csc(vertices, tets, degree, &
            quadrature_degree, matrix, st, kappa, mu)
        call record(st%code == 0, "global H1 primal assembly succeeds")
        call assemble_tetra_lagrange_stiffness_csc_jvp(vertices, tets, degree, &
            quadrature_degree, kappa, mu, vertices_dot, kappa_dot, mu_dot, &
            matrix_dot, st)
        call record(st%code == 0, "global H1 JVP succeeds")
        call assemble_tetra_lagrange_stiffness_csc(vertices + step*vertices_dot, &
            tets, degree, quadrature_degree, plus, st, kappa + step*kappa_dot, &
            mu + step*mu_dot)
        call assemble_tetra_lagrange_stiffness_csc(vertices - step*vertices_dot, &
            tets, degree, quadrature_degree, minus, st, kappa - step*kappa_dot, &
            mu - step*mu_dot)
        call record(matrix_dot%nnz == matrix%nnz .and. &
            all(matrix_dot%col_ptr == matrix%col_ptr) .and. &
            all(matrix_dot%row_idx == matrix%row_idx), &
            "global H1 JVP preserves the merged CSC pattern")
        relative_error = maxval(abs(matrix_dot%val - &
            (plus%val - minus%val)/(2.0_dp*step)))/ &
            max(1.0_dp, maxval(abs(matrix_dot%val)))
        call record(relative_error < 5.0e-8_dp, &
            "global H1 JVP matches independent complete mesh reassembly")
        allocate(values_bar(matrix%nnz))
        do entry = 1, matrix%nnz
            values_bar(entry) = 0.003_dp*entry - 0.00001_dp*entry**2
        end do
        call assemble_tetra_lagrange_stiffness_csc_vjp(vertices, tets, degree, &
            quadrature_degree, kappa, mu, values_bar, vertices_bar, kappa_bar, &
            mu_bar, st)
        call record(st%code == 0, "global H1 VJP succeeds")
        lhs = dot_product(values_bar, matrix_dot%val)
        rhs = sum(vertices_bar*vertices_dot) + kappa_bar*kappa_dot + mu_bar*mu_dot
        call record(abs(lhs - rhs) < 3.0e-10_dp*max(1.0_dp, abs(lhs), abs(rhs)), &
            "global H1 products accumulate the shared-vertex adjoint")
    end subroutine check_h1_csc_ad

    subroutine check_nedelec_csc_ad()
        integer, parameter :: order = 2
        integer, parameter :: tets(4, 2) = reshape([1, 2, 3, 4, 2, 3, 4, 5], [4, 2])
        real(dp), parameter :: step = 2.0e-7_dp
        real(dp) :: vertices(3, 5), vertices_dot(3, 5), vertices_bar(3, 5)
        real(dp), allocatable :: values_bar(:)
        type(csc_t) :: matrix, matrix_dot, minus, plus
        type(fortsparse_status_t) :: st
        real(dp) :: alpha, beta, alpha_dot, beta_dot, alpha_bar, beta_bar
        real(dp) :: lhs, rhs, relative_error
        integer :: entry

        vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 1.0_dp, 1.0_dp, &
            1.0_dp], [3, 5])
        vertices_dot = reshape([0.01_dp, -0.02_dp, 0.015_dp, -0.03_dp, 0.01_dp, &
            0.02_dp, 0.025_dp, -0.015_dp, 0.005_dp, -0.01_dp, 0.03_dp, -0.02_dp, &
            0.02_dp, -0.01_dp, 0.025_dp], [3, 5])
        alpha = 1.7_dp
        beta = -0.8_dp
        alpha_dot = 0.13_dp
        beta_dot = -0.09_dp
        call assemble_tetra_nedelec_curl_mass_csc(vertices, tets, matrix, st, &
            alpha, beta, order)
        call assemble_tetra_nedelec_curl_mass_csc_jvp(vertices, tets, order, &
            alpha, beta, vertices_dot, alpha_dot, beta_dot, matrix_dot, st)
        call record(st%code == 0, "Nedelec curl-mass JVP succeeds")
        call assemble_tetra_nedelec_curl_mass_csc(vertices + step*vertices_dot, &
            tets, plus, st, alpha + step*alpha_dot, beta + step*beta_dot, order)
        call assemble_tetra_nedelec_curl_mass_csc(vertices - step*vertices_dot, &
            tets, minus, st, alpha - step*alpha_dot, beta - step*beta_dot, order)
        call record(matrix_dot%nnz == matrix%nnz .and. &
            all(matrix_dot%col_ptr == matrix%col_ptr) .and. &
            all(matrix_dot%row_idx == matrix%row_idx), &
            "Global curl-mass JVP preserves the merged CSC pattern")
        relative_error = maxval(abs(matrix_dot%val - &
            (plus%val - minus%val)/(2.0_dp*step)))/ &
            max(1.0_dp, maxval(abs(matrix_dot%val)))
        call record(relative_error < 5.0e-8_dp, &
            "Global curl-mass JVP matches complete merged-mesh reassembly")
        allocate(values_bar(matrix%nnz))
        do entry = 1, matrix%nnz
            values_bar(entry) = 0.003_dp*entry - 0.00001_dp*entry**2
        end do
        call assemble_tetra_nedelec_curl_mass_csc_vjp(vertices, tets, order, &
            alpha, beta, values_bar, vertices_bar, alpha_bar, beta_bar, st)
        call record(st%code == 0, "Global curl-mass CSC VJP succeeds")
        lhs = sum(values_bar*matrix_dot%val)
        rhs = sum(vertices_bar*vertices_dot) + alpha_bar*alpha_dot + beta_bar*beta_dot
        call record(abs(lhs - rhs) < 3.0e-10_dp*max(1.0_dp, abs(lhs), abs(rhs)), &
            "Global curl-mass products accumulate the shared-vertex adjoint")
    end subroutine check_nedelec_csc_ad

    ! The curl-mass solve with a constant source reproduces the constant
    ! field (it lies in the space); for k = 1 the global dofs are then the
    ! Whitney moments o_e f . (x_j - x_i) of every local edge (i, j).
    subroutine check_nedelec_constant_solve()
        use, intrinsic :: iso_c_binding, only: c_int
        integer, parameter :: tets(4, 2) = reshape([1, 2, 3, 4, 2, 3, 4, 5], [4, 2])
        integer, parameter :: ev(2, 6) = reshape([1, 2, 1, 3, 1, 4, 2, 3, 2, 4, 3, 4], [2, 6])
        real(dp) :: vertices(3, 5), err
        real(dp), allocatable :: solution(:)
        type(fortsparse_status_t) :: st
        integer(c_int) :: gd(6, 2), eo(6, 2), fp(3, 4, 2), nd
        integer :: cell, e, order
        logical :: ok

        vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 1.0_dp, 1.0_dp, &
            1.0_dp], [3, 5])
        ok = .true.
        do order = 1, 3
            call solve_tetra_nedelec_curl_mass(vertices, tets, order, constant_field, &
                1.0_dp, 1.0_dp, solution, st)
            ok = ok .and. st%code == 0
            if (order == 1 .and. st%code == 0) then
                if (ffh_nedelec_dof_map(1_c_int, 2_c_int, tets, 1_c_int, gd, eo, fp, &
                    nd) /= 0) ok = .false.
                err = 0.0_dp
                do cell = 1, 2
                    do e = 1, 6
                        err = max(err, abs(solution(gd(e, cell)) - eo(e, cell)* &
                            dot_product([0.7_dp, -1.3_dp, 0.4_dp], &
                            vertices(:, tets(ev(2, e), cell)) - &
                            vertices(:, tets(ev(1, e), cell)))))
                    end do
                end do
                call record(err < 1.0e-9_dp, &
                    "Nedelec k=1 solve reproduces the constant field's edge moments")
            end if
            if (allocated(solution)) deallocate(solution)
        end do
        call record(ok, "Nedelec curl-mass solves succeed for k = 1..3")
    end subroutine check_nedelec_constant_solve

    ! test/test_tetra_nedelec_pml_sparse.f90:20-51: on the two-tet mesh the
    ! constant field (1, 2, -1) has PML energy -k^2/3 sum(m_i f_i^2) with
    ! m = (s2 s3/s1, s1 s3/s2, s1 s2/s3); a diagonal curvilinear stretch gives
    ! the cartesian matrix.
    subroutine check_nedelec_pml()
        use, intrinsic :: iso_c_binding, only: c_int
        integer, parameter :: ev(2, 6) = reshape([1, 2, 1, 3, 1, 4, 2, 3, 2, 4, 3, 4], [2, 6])
        real(dp) :: vertices(3, 5), f(3)
        integer :: tets(4, 2), cell, e, c, j, order
        integer(c_int) :: gd(6, 2), eo(6, 2), fp(3, 4, 2), nd
        complex(dp) :: stretch(3, 2), full(3, 3, 2), m(3), energy, expected
        complex(dp), allocatable :: d(:), ad(:)
        type(csc_z_t) :: a, b
        type(fortsparse_status_t) :: st
        logical :: ok

        vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, -1.0_dp], [3, 5])
        tets(:, 1) = [1, 2, 3, 4]
        tets(:, 2) = [1, 3, 2, 5]
        stretch(:, 1) = [cmplx(1.0_dp, 0.6_dp, dp), cmplx(1.0_dp, 0.0_dp, dp), &
            cmplx(1.0_dp, 0.0_dp, dp)]
        stretch(:, 2) = stretch(:, 1)
        f = [1.0_dp, 2.0_dp, -1.0_dp]
        call assemble_tetra_nedelec_pml_csc(vertices, tets, 1, stretch, 1.3_dp, a, st)
        ok = st%code == 0
        if (ok) ok = ffh_nedelec_dof_map(1_c_int, 2_c_int, tets, 1_c_int, gd, eo, fp, nd) == 0
        if (ok) then
            allocate(d(a%ncol), ad(a%nrow))
            do cell = 1, 2
                do e = 1, 6
                    d(gd(e, cell)) = eo(e, cell)*dot_product(f, &
                        vertices(:, tets(ev(2, e), cell)) - vertices(:, tets(ev(1, e), cell)))
                end do
            end do
            ad = (0.0_dp, 0.0_dp)
            do c = 1, a%ncol
                do j = a%col_ptr(c), a%col_ptr(c + 1) - 1
                    ad(a%row_idx(j)) = ad(a%row_idx(j)) + a%val(j)*d(c)
                end do
            end do
            energy = sum(d*ad)
            m = [stretch(2, 1)*stretch(3, 1)/stretch(1, 1), &
                stretch(1, 1)*stretch(3, 1)/stretch(2, 1), &
                stretch(1, 1)*stretch(2, 1)/stretch(3, 1)]
            expected = -1.3_dp**2/3.0_dp*sum(m*f**2)
            ok = abs(energy - expected) < 3.0e-12_dp
        end if
        call record(ok, "Nedelec PML constant-field energy")
        ok = .true.
        do order = 1, 3
            full = (0.0_dp, 0.0_dp)
            do cell = 1, 2
                do e = 1, 3
                    full(e, e, cell) = stretch(e, cell)*cmplx(1.0_dp, 0.1_dp*order, dp)
                end do
            end do
            call assemble_tetra_nedelec_pml_csc(vertices, tets, order, &
                reshape([(full(e, e, 1), e = 1, 3), (full(e, e, 2), e = 1, 3)], [3, 2]), &
                0.7_dp, a, st)
            ok = ok .and. st%code == 0
            call assemble_tetra_nedelec_curvilinear_pml_csc(vertices, tets, order, full, &
                0.7_dp, b, st)
            ok = ok .and. st%code == 0
            if (ok) ok = a%nnz == b%nnz .and. all(a%row_idx == b%row_idx) .and. &
                maxval(abs(a%val - b%val)) <= 1.0e-12_dp*maxval(abs(a%val))
        end do
        call record(ok, "diagonal curvilinear PML equals the cartesian PML")
        stretch(2, 1) = (0.0_dp, 0.0_dp)
        call assemble_tetra_nedelec_pml_csc(vertices, tets, 1, stretch, 1.3_dp, a, st)
        call record(st%code /= 0, "zero PML stretch is rejected")
    end subroutine check_nedelec_pml

    ! The PML cotangent is the adjoint of the tangent in the reference's
    ! convention: Re sum conj(w) A_dot = <v_bar, v_dot> + Re sum conj(s_bar)
    ! s_dot + k_bar k_dot.
    subroutine check_nedelec_pml_ad()
        real(dp) :: vertices(3, 5), vd(3, 5), vb(3, 5), kb, lhs, rhs
        integer :: tets(4, 2), i
        complex(dp) :: stretch(3, 2), sd(3, 2), sb(3, 2)
        complex(dp), allocatable :: w(:)
        type(csc_z_t) :: a
        type(fortsparse_status_t) :: st
        logical :: ok

        vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, -1.0_dp], [3, 5])
        tets(:, 1) = [1, 2, 3, 4]
        tets(:, 2) = [1, 3, 2, 5]
        stretch(:, 1) = [cmplx(1.0_dp, 0.6_dp, dp), cmplx(0.9_dp, 0.1_dp, dp), &
            cmplx(1.2_dp, -0.2_dp, dp)]
        stretch(:, 2) = [cmplx(1.1_dp, 0.3_dp, dp), cmplx(1.0_dp, 0.0_dp, dp), &
            cmplx(0.8_dp, 0.4_dp, dp)]
        do i = 1, 15
            vd(mod(i - 1, 3) + 1, (i - 1) / 3 + 1) = sin(0.7_dp * i)
        end do
        sd = cmplx(cos(1.3_dp * reshape([(real(i, dp), i = 1, 6)], [3, 2])), &
            sin(0.4_dp * reshape([(real(i, dp), i = 1, 6)], [3, 2])), dp)
        call assemble_tetra_nedelec_pml_csc_jvp(vertices, tets, 2, stretch, 1.3_dp, vd, &
            sd, 0.25_dp, a, st)
        ok = st%code == 0
        if (ok) then
            allocate(w(a%nnz))
            do i = 1, a%nnz
                w(i) = cmplx(cos(0.37_dp * i), sin(0.91_dp * i), dp)
            end do
            call assemble_tetra_nedelec_pml_csc_vjp(vertices, tets, 2, stretch, 1.3_dp, w, &
                vb, sb, kb, st)
            ok = st%code == 0
        end if
        if (ok) then
            lhs = real(sum(conjg(w) * a%val), dp)
            rhs = sum(vb * vd) + real(sum(conjg(sb) * sd), dp) + kb * 0.25_dp
            ok = abs(lhs - rhs) <= 1.0e-10_dp * max(abs(lhs), 1.0_dp)
        end if
        call record(ok, "Nedelec PML VJP is the adjoint of the JVP")
    end subroutine check_nedelec_pml_ad

    subroutine check_linear_energy()
        real(dp) :: mesh_vertices(3, 5), coefficients(5), product(5)
        integer :: tetrahedra(4, 2)
        type(csc_t) :: matrix
        type(fortsparse_status_t) :: status

        mesh_vertices = reshape([0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, &
            0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, -1.0_dp], [3, 5])
        tetrahedra(:, 1) = [1, 2, 3, 4]
        tetrahedra(:, 2) = [1, 3, 2, 5]
        call assemble_tetra_lagrange_stiffness_csc(mesh_vertices, tetrahedra, 1, 4, &
            matrix, status)
        call record(status%code == 0 .and. matrix%nrow == 5, "P1 CSC assembly succeeds")
        ! first-appearance numbering on this mesh: 4, 3, 2, 1, 5 -> dofs 1..5
        coefficients = [0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp]
        call spmv(matrix, coefficients, product)
        call record(abs(dot_product(coefficients, product) - 1.0_dp/3.0_dp) < 2.0e-12_dp, &
            "linear-field energy is 1/3")
    end subroutine check_linear_energy

    subroutine check_cg_tridiagonal()
        integer, parameter :: n = 100
        type(csc_t) :: a
        type(solver_options_t) :: opts
        type(solver_stats_t) :: stats
        real(dp) :: b(n), x(n), x_exact(n), ax(n)
        integer :: i, k

        a%nrow = n
        a%ncol = n
        a%nnz = 3*n - 2
        allocate(a%col_ptr(n + 1), a%row_idx(a%nnz), a%val(a%nnz))
        k = 0
        do i = 1, n
            a%col_ptr(i) = k + 1
            if (i > 1) then
                k = k + 1
                a%row_idx(k) = i - 1
                a%val(k) = -1.0_dp
            end if
            k = k + 1
            a%row_idx(k) = i
            a%val(k) = 4.1_dp
            if (i < n) then
                k = k + 1
                a%row_idx(k) = i + 1
                a%val(k) = -1.0_dp
            end if
        end do
        a%col_ptr(n + 1) = k + 1
        x_exact = [(real(i, dp), i=1, n)]
        call spmv(a, x_exact, b)
        x = 0.0_dp
        opts = solver_options(method="cg", tolerance=1.0e-10_dp, max_iterations=n)
        call solve_sparse(a, b, x, opts, stats)
        call spmv(a, x, ax)
        call record(stats%converged, "CG converged")
        call record(stats%iterations < n/2, "CG reasonable iteration count")
        call record(norm2(ax - b) < 1.0e-7_dp, "CG small residual")
        call record(norm2(x - x_exact) < 1.0e-6_dp, "CG accurate solution")
        x = 0.0_dp
        opts = solver_options(method="pcg", tolerance=1.0e-12_dp, max_iterations=5)
        call solve_sparse(a, b, x, opts, stats)
        call record(.not. stats%converged .and. stats%iterations == 5, &
            "iteration limit enforced")
    end subroutine check_cg_tridiagonal



    subroutine check_quartic_p4()
        real(dp), parameter :: vertices(3, 5) = reshape([0.0_dp, 0.0_dp, 0.0_dp, &
            1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, &
            0.25_dp, 0.25_dp, 0.25_dp], [3, 5])
        integer, parameter :: tetrahedra(4, 4) = reshape([1, 2, 3, 5, 1, 4, 2, 5, &
            1, 3, 4, 5, 2, 4, 3, 5], [4, 4])
        real(dp), allocatable :: solution(:)
        type(fortsparse_status_t) :: status
        type(solver_options_t) :: opts
        type(solver_stats_t) :: stats

        opts = solver_options(method="pcg", preconditioner="jacobi", &
            tolerance=1.0e-14_dp, max_iterations=10000)
        call solve_tetra_lagrange_poisson(vertices, tetrahedra, 4, bubble_source, &
            zero_boundary, solution, status, opts, stats)
        call record(status%code == 0 .and. stats%converged, "P4 bubble solve succeeds")
        ! the only interior P4 node of cell 1 with all-equal barycentrics does
        ! not exist; check the maximum of u = xyz(1-x-y-z) instead: u <= 1/256
        call record(maxval(solution) <= 1.0_dp/256.0_dp + 1.0e-10_dp .and. &
            minval(solution) >= -1.0e-12_dp, "P4 solution bounded by the exact maximum")
    end subroutine check_quartic_p4


    subroutine check_p2_box_poisson()
        real(dp), allocatable :: vertices(:, :), solution(:)
        integer, allocatable :: tetrahedra(:, :)
        type(fortsparse_status_t) :: status
        type(solver_stats_t) :: stats
        integer :: mesh_status

        call generate_structured_tetra_box_mesh(reshape([0.0_dp, 0.0_dp, 0.0_dp, &
            1.0_dp, 1.0_dp, 1.0_dp], [3, 2]), [6, 6, 6], vertices, tetrahedra, mesh_status)
        call solve_tetra_lagrange_poisson(vertices, tetrahedra, 2, sin_source, &
            zero_boundary, solution, status, stats=stats)
        call record(status%code == 0 .and. stats%converged, "P2 box Poisson converges")
        ! u = sin(pi x) sin(pi y) sin(pi z) has maximum 1 at the centre node
        call record(abs(maxval(solution) - 1.0_dp) < 1.0e-2_dp, &
            "P2 solution reaches the exact maximum within 1%")
    end subroutine check_p2_box_poisson

end program test_fortfem_hip_dropin
