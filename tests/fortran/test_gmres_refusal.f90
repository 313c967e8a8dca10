! solve_sparse with a method or preconditioner the GPU path does not run
! (gmres here) goes to fortfem's CPU solve_sparse; in a build without
! fortfem's CPU modules that stops with a message naming the routine, instead
! of silently running PCG.  Expected: a nonzero exit status and the message
! on stderr, before any device call.
program test_gmres_refusal
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, solver_options, &
        solve_sparse, solve_sparse_on_gpu
    use fortfem_sparse_matrix, only: sparse_matrix_t, sparse_from_dense
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    real(dp) :: a(2, 2), b(2), x(2)
    type(sparse_matrix_t) :: s
    type(solver_options_t) :: opts
    type(solver_stats_t) :: stats
    character(len=16) :: arg

    a = reshape([2.0_dp, 1.0_dp, 0.0_dp, 3.0_dp], [2, 2])
    call sparse_from_dense(a, s)
    b = [1.0_dp, 1.0_dp]
    x = 0.0_dp
    arg = "gmres"
    if (command_argument_count() > 0) call get_command_argument(1, arg)
    if (trim(arg) == "ilu") then
        opts = solver_options(method="pcg", preconditioner="ilu")
    else
        opts = solver_options(method=trim(arg))
    end if
    if (solve_sparse_on_gpu(opts)) error stop "the option set was meant for the CPU solver"
    call solve_sparse(s, b, x, opts, stats)
    print "(a)", "solve_sparse returned"
end program test_gmres_refusal
