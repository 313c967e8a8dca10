! bind(C) wrappers around reference modules compiled from their own sources
! (oracle/ref.mk).  TEST INFRASTRUCTURE ONLY: the oracle's restatement is
! checked against these in tests/; nothing in the product links them.
module ffo_ref_bindings
    use, intrinsic :: iso_c_binding, only: c_int, c_double
    use fortfem_tetra_lagrange_arbitrary_order, only: tetra_lagrange_t, &
        initialize_tetra_lagrange, evaluate_tetra_lagrange, tetra_lagrange_dof_count
    use fortfem_tetra_lagrange_global_dof_map, only: build_tetra_lagrange_dof_map
    use fortfem_tetra_edge_dof_map, only: build_tetra_edge_dof_map
    use fortfem_tetra_nedelec_first_order, only: evaluate_tetra_nedelec_first_order
    use fortfem_generated_tetra_nedelec_coefficients, only: load_tetra_nedelec_coefficients
    use fortfem_generated_tetra_nedelec_candidates_order_1, only: evaluate_candidates_order_1
    use fortfem_generated_tetra_nedelec_candidates_order_2, only: evaluate_candidates_order_2
    use fortfem_generated_tetra_nedelec_candidates_order_3, only: evaluate_candidates_order_3
    use fortfem_generated_tetra_nedelec_candidates_order_4, only: evaluate_candidates_order_4
    use fortfem_generated_tetra_face_moment_transforms, only: map_tetra_face_basis_to_local
    use fortfem_generated_tetra_h1_oracle, only: generated_tetra_h1_oracle
    implicit none

contains

    integer(c_int) function ffr_lagrange_eval(degree, point, values, gradients) &
            bind(C, name="ffr_lagrange_eval")
        integer(c_int), value :: degree
        real(c_double), intent(in) :: point(3)
        real(c_double), intent(out) :: values(*), gradients(3, *)
        type(tetra_lagrange_t) :: basis
        integer :: st, n
        real(c_double), allocatable :: v(:), g(:, :)

        call initialize_tetra_lagrange(int(degree), basis, st)
        ffr_lagrange_eval = st
        if (st /= 0) return
        n = tetra_lagrange_dof_count(basis)
        allocate(v(n), g(3, n))
        call evaluate_tetra_lagrange(basis, point, v, g, st)
        values(1:n) = v
        gradients(:, 1:n) = g
        ffr_lagrange_eval = st
    end function ffr_lagrange_eval

    integer(c_int) function ffr_lagrange_dof_map(degree, ntet, tets, gd, count) &
            bind(C, name="ffr_lagrange_dof_map")
        integer(c_int), value :: degree, ntet
        integer(c_int), intent(in) :: tets(4, ntet)
        integer(c_int), intent(out) :: gd(*), count
        integer, allocatable :: g(:, :)
        integer :: st, cnt

        call build_tetra_lagrange_dof_map(int(degree), tets, g, cnt, st)
        ffr_lagrange_dof_map = st
        count = cnt
        if (st /= 0) return
        gd(1:size(g)) = reshape(g, [size(g)])
    end function ffr_lagrange_dof_map

    integer(c_int) function ffr_edge_dof_map(ntet, tets, nedges, edges, gd, orient) &
            bind(C, name="ffr_edge_dof_map")
        integer(c_int), value :: ntet
        integer(c_int), intent(in) :: tets(4, ntet)
        integer(c_int), intent(out) :: nedges, edges(2, *), gd(6, *), orient(6, *)
        integer, allocatable :: e(:, :), g(:, :), o(:, :)
        integer :: st

        call build_tetra_edge_dof_map(tets, e, g, o, st)
        ffr_edge_dof_map = st
        if (st /= 0) return
        nedges = size(e, 2)
        edges(:, 1:size(e, 2)) = e
        gd(:, 1:ntet) = g
        orient(:, 1:ntet) = o
    end function ffr_edge_dof_map

    integer(c_int) function ffr_whitney(point, values, curls) bind(C, name="ffr_whitney")
        real(c_double), intent(in) :: point(3)
        real(c_double), intent(out) :: values(3, 6), curls(3, 6)
        integer :: st

        call evaluate_tetra_nedelec_first_order(point, values, curls, st)
        ffr_whitney = st
    end function ffr_whitney

    ! physical-free reference basis of order k <= 4: candidates x coefficients
    integer(c_int) function ffr_nedelec_basis(order, point, values, curls) &
            bind(C, name="ffr_nedelec_basis")
        integer(c_int), value :: order
        real(c_double), intent(in) :: point(3)
        real(c_double), intent(out) :: values(3, *), curls(3, *)
        real(c_double), allocatable :: coeff(:, :), cv(:, :), cc(:, :)
        integer :: st, n

        ffr_nedelec_basis = 1
        if (order < 1 .or. order > 4) return
        call load_tetra_nedelec_coefficients(int(order), coeff, st)
        if (st /= 0) return
        n = size(coeff, 1)
        allocate(cv(3, n), cc(3, n))
        select case (order)
        case (1)
            call evaluate_candidates_order_1(point(1), point(2), point(3), cv, cc)
        case (2)
            call evaluate_candidates_order_2(point(1), point(2), point(3), cv, cc)
        case (3)
            call evaluate_candidates_order_3(point(1), point(2), point(3), cv, cc)
        case (4)
            call evaluate_candidates_order_4(point(1), point(2), point(3), cv, cc)
        end select
        values(:, 1:n) = matmul(cv, coeff)
        curls(:, 1:n) = matmul(cc, coeff)
        ffr_nedelec_basis = 0
    end function ffr_nedelec_basis

    integer(c_int) function ffr_face_basis_to_local(order, perm, canonical, local) &
            bind(C, name="ffr_face_basis_to_local")
        integer(c_int), value :: order
        integer(c_int), intent(in) :: perm(3)
        real(c_double), intent(in) :: canonical(*)
        real(c_double), intent(out) :: local(*)
        integer :: st, n

        n = order*(order - 1)
        call map_tetra_face_basis_to_local(int(order), int(perm), canonical(1:n), &
            local(1:n), st)
        ffr_face_basis_to_local = st
    end function ffr_face_basis_to_local

    subroutine ffr_h1_oracle(x, y, z, values) bind(C, name="ffr_h1_oracle")
        real(c_double), value :: x, y, z
        real(c_double), intent(out) :: values(5)

        call generated_tetra_h1_oracle(x, y, z, values)
    end subroutine ffr_h1_oracle

end module ffo_ref_bindings
