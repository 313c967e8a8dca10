! fortsparse_standin.f90 -- the part of fortsparse's interface that the
! drop-in modules use (csc_t, csc_z_t, fortsparse_status_t, status_set and the
! status codes), for building this library OUTSIDE fortfem.  fortfem's own
! build compiles the real fortsparse (fpm.toml:62, @3c19c978) and does not
! compile this file: the drop-in modules then take these names from it.
! Field names follow the uses in the reference (e.g.
! src/solvers/fortfem_sparse_direct.f90:1080-1090): nrow, ncol, nnz,
! col_ptr(ncol+1), row_idx(nnz), val(nnz), 1-based; status%code, %msg.
module fortsparse
    use, intrinsic :: iso_fortran_env, only: real64
    implicit none
    private

    integer, parameter :: dp = real64
    integer, parameter, public :: FORTSPARSE_OK = 0
    integer, parameter, public :: FORTSPARSE_INVALID_MATRIX = 2
    integer, parameter, public :: FORTSPARSE_INTERNAL_ERROR = 3
    integer, parameter, public :: FORTSPARSE_SINGULAR = 4
    integer, parameter, public :: FORTSPARSE_NOT_FACTORED = 5

    type, public :: csc_t
        integer :: nrow = 0
        integer :: ncol = 0
        integer :: nnz = 0
        integer, allocatable :: col_ptr(:)
        integer, allocatable :: row_idx(:)
        real(dp), allocatable :: val(:)
    end type csc_t

    type, public :: csc_z_t
        integer :: nrow = 0
        integer :: ncol = 0
        integer :: nnz = 0
        integer, allocatable :: col_ptr(:)
        integer, allocatable :: row_idx(:)
        complex(dp), allocatable :: val(:)
    end type csc_z_t

    type, public :: fortsparse_status_t
        integer :: code = 0
        character(len=256) :: msg = ""
    end type fortsparse_status_t

    public :: status_set

contains

    subroutine status_set(status, code, msg)
        type(fortsparse_status_t), intent(inout) :: status
        integer, intent(in) :: code
        character(len=*), intent(in) :: msg

        status%code = code
        status%msg = msg
    end subroutine status_set

end module fortsparse
