! Host-only checks of the drop-in Fortran modules (no device needed):
! generate_structured_tetra_box_mesh and tetra_duffy_quadrature through the
! C-ABI.  Restates test/test_tetra_duffy_quadrature.f90:13-33 and
! test/test_structured_tetra_box_mesh.f90:19-51.
program test_fortfem_hip_host
    use, intrinsic :: iso_fortran_env, only: output_unit
    use fortfem_hip_types, only: dp
    use fortfem_structured_tetra_box_mesh, only: generate_structured_tetra_box_mesh
    use fortfem_tetra_duffy_quadrature, only: tetra_duffy_quadrature
    implicit none

    real(dp), allocatable :: x(:), y(:), z(:), w(:), vertices(:, :)
    integer, allocatable :: tetrahedra(:, :)
    real(dp) :: exact, numerical
    integer :: degree, a, b, c, status
    logical :: all_passed

    all_passed = .true.
    do degree = 0, 12
        call tetra_duffy_quadrature(degree, x, y, z, w, status)
        call record(status == 0 .and. size(w) == ((degree + 4)/2)**3, "Duffy rule size")
        call record(abs(sum(w) - 1.0_dp/6.0_dp) < 2.0e-14_dp, "Duffy weights have exact volume")
        do c = 0, min(degree, 7)
            do b = 0, min(degree, 7) - c
                do a = 0, min(degree, 7) - b - c
                    numerical = sum(w*x**a*y**b*z**c)
                    exact = fact(a)*fact(b)*fact(c)/fact(a + b + c + 3)
                    if (abs(numerical - exact) >= 4.0e-14_dp) then
                        call record(.false., "Duffy monomial exactness")
                    end if
                end do
            end do
        end do
    end do
    call tetra_duffy_quadrature(-1, x, y, z, w, status)
    call record(status /= 0, "Duffy rule rejects a negative degree")
    call generate_structured_tetra_box_mesh(reshape([0.0_dp, 0.0_dp, 0.0_dp, &
        1.0_dp, 2.0_dp, 3.0_dp], [3, 2]), [3, 2, 4], vertices, tetrahedra, status)
    call record(status == 0 .and. size(tetrahedra, 2) == 6*24 .and. &
        size(vertices, 2) == 4*3*5, "box mesh sizes")
    if (.not. all_passed) error stop 1
    write (*, '(a)') "host Fortran checks passed"

contains

    real(dp) function fact(n)
        integer, intent(in) :: n
        integer :: i
        fact = 1.0_dp
        do i = 2, n
            fact = fact*i
        end do
    end function fact

    subroutine record(condition, description)
        logical, intent(in) :: condition
        character(len=*), intent(in) :: description

        if (.not. condition) then
            write (*, '(2a)') "FAIL ", description
            all_passed = .false.
        end if
        flush (output_unit)
    end subroutine record

end program test_fortfem_hip_host
