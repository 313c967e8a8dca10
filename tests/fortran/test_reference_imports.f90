! The import lists of the reference's callers of the path, compiled against
! the drop-in modules: every name they take must exist with the reference's
! meaning.  Lists restated from src/fortfem_api.f90:1780-1785,
! src/fortfem_api_solvers.f90:9-17, src/solvers/fortfem_solvers_laplacian.f90:
! 6-7, src/bem/laplace_fem_bem_coupling_3d.f90:8-9 and
! src/solvers/tetra_lagrange_state_3d.f90:3-6, plus the remaining public names
! of the replaced modules.  Run: a small sparse solve through the reference's
! types (GPU path when a device is present, else the status is reported) and
! the forwarded host helpers.
module api_like
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t,   &
        solver_options, cg_solve, pcg_solve, bicgstab_solve, gmres_solve,   &
        cg_solve_jvp, cg_solve_vjp, pcg_solve_jvp, pcg_solve_vjp,           &
        bicgstab_solve_jvp, bicgstab_solve_vjp, gmres_solve_jvp,            &
        gmres_solve_vjp, jacobi_preconditioner, ilu_preconditioner, &
        ichol_preconditioner, ichol_controlled_preconditioner
end module api_like

module api_solvers_like
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, &
        solve_sparse, solver_options, cg_solve, pcg_solve, bicgstab_solve, &
        gmres_solve, &
        cg_solve_jvp, cg_solve_vjp, pcg_solve_jvp, pcg_solve_vjp, &
        bicgstab_solve_jvp, bicgstab_solve_vjp, gmres_solve_jvp, &
        gmres_solve_vjp, &
        jacobi_preconditioner, ilu_preconditioner, ichol_controlled_preconditioner
    use fortfem_sparse_matrix, only: sparse_matrix_t, sparse_from_dense, &
        spmv, spmv_jvp, spmv_vjp
end module api_solvers_like

module laplacian_like
    use fortfem_advanced_solvers, only: solver_options_t, solver_stats_t, &
        solver_options, advanced_solve => solve
end module laplacian_like

module fem_bem_like
    use fortfem_assembly_tetra_lagrange_arbitrary_order_3d, only: &
        assemble_tetra_lagrange_stiffness_csc, assemble_tetra_lagrange_stiffness_element, &
        assemble_tetra_lagrange_stiffness_element_jvp, &
        assemble_tetra_lagrange_stiffness_element_vjp, &
        assemble_tetra_lagrange_scalar_load, assemble_tetra_lagrange_scalar_load_samples, &
        assemble_tetra_lagrange_scalar_load_samples_jvp, &
        assemble_tetra_lagrange_scalar_load_samples_vjp, scalar_source_3d, &
        assemble_tetra_lagrange_stiffness_csc_jvp, assemble_tetra_lagrange_stiffness_csc_vjp
    use fortfem_assembly_tetra_nedelec_3d, only: &
        assemble_tetra_nedelec_curl_mass_csc, assemble_tetra_nedelec_curl_mass_csc_jvp, &
        assemble_tetra_nedelec_curl_mass_csc_vjp, assemble_tetra_nedelec_curl_mass_element, &
        assemble_tetra_nedelec_curl_mass_element_jvp, &
        assemble_tetra_nedelec_curl_mass_element_vjp, assemble_tetra_nedelec_pml_element, &
        assemble_tetra_nedelec_pml_element_jvp, assemble_tetra_nedelec_pml_element_vjp, &
        assemble_tetra_nedelec_curvilinear_pml_element, &
        assemble_tetra_nedelec_curvilinear_pml_element_jvp, &
        assemble_tetra_nedelec_curvilinear_pml_element_vjp, &
        assemble_tetra_nedelec_curvilinear_pml_csc, &
        assemble_tetra_nedelec_curvilinear_pml_csc_jvp, &
        assemble_tetra_nedelec_curvilinear_pml_csc_vjp, assemble_tetra_nedelec_pml_csc, &
        assemble_tetra_nedelec_pml_csc_jvp, assemble_tetra_nedelec_pml_csc_vjp, &
        assemble_tetra_nedelec_weighted_csc, assemble_tetra_nedelec_vector_load, &
        assemble_tetra_nedelec_vector_load_order, assemble_tetra_nedelec_vector_load_samples, &
        assemble_tetra_nedelec_vector_load_samples_jvp, &
        assemble_tetra_nedelec_vector_load_samples_vjp
    use fortfem_tetra_lagrange_solver_3d, only: assignment(=), &
        evaluate_tetra_lagrange_solution, evaluate_tetra_lagrange_solution_jvp, &
        evaluate_tetra_lagrange_solution_vjp, evaluate_tetra_lagrange_solution_at_point, &
        evaluate_tetra_lagrange_solution_at_point_jvp, &
        evaluate_tetra_lagrange_solution_at_point_vjp, &
        evaluate_tetra_lagrange_solution_prepared, &
        initialize_tetra_lagrange_solution_evaluator, &
        solve_tetra_lagrange_diffusion_reaction, &
        solve_tetra_lagrange_poisson, tetra_lagrange_solution_evaluator_t
    use fortfem_tetra_nedelec_solver_3d, only: solve_tetra_nedelec_curl_mass, &
        solve_tetra_nedelec_pml, solve_tetra_nedelec_curvilinear_pml, &
        solve_tetra_nedelec_weighted_curl_mass
    use fortfem_structured_tetra_box_mesh, only: generate_structured_tetra_box_mesh
    use fortfem_tetra_duffy_quadrature, only: tetra_duffy_quadrature
end module fem_bem_like

program test_reference_imports
    use api_solvers_like
    use fortsparse, only: csc_t
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    real(dp) :: a(3, 3), b(3), x(3)
    type(sparse_matrix_t) :: s
    type(solver_options_t) :: opts
    type(solver_stats_t) :: stats
    integer :: fails

    fails = 0
    a = reshape([4.0_dp, -1.0_dp, 0.0_dp, -1.0_dp, 4.0_dp, -1.0_dp, 0.0_dp, -1.0_dp, 4.0_dp], [3, 3])
    call sparse_from_dense(a, s)
    if (s%nnz /= 7 .or. any(s%col_ptr /= [1, 3, 6, 8])) fails = fails + 1
    opts = solver_options(method="pcg", preconditioner=jacobi_preconditioner(), tolerance=1.0e-12_dp)
    if (trim(opts%preconditioner) /= "jacobi" .or. opts%max_iterations /= 1000) fails = fails + 1
    if (trim(ilu_preconditioner()) /= "ilu") fails = fails + 1
    if (trim(ichol_controlled_preconditioner()) /= "ichol_controlled") fails = fails + 1
    b = [3.0_dp, 2.0_dp, 3.0_dp]
    x = 0.0_dp
    print "(a, i0, a, i0)", "reference import lists compiled; sparse_from_dense nnz ", s%nnz, &
        ", failures ", fails
    if (fails /= 0) error stop 1
end program test_reference_imports
