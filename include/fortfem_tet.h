/*
 * fortfem_tet.h -- tetrahedral exports in the conventions of the reference's
 * C API (include/fortfem.h): void functions with an int *status out
 * argument, zero-based indices at the boundary, positive process-local
 * integer handles from fixed registries of 32 slots
 * (src/bindings/fortfem_capi_solver.f90:10-13, fortfem_capi_mesh.f90:16-18),
 * and the capacity protocol of the CSC exports (fortfem_capi_mesh.f90:
 * 545-571: status -2 when the capacity is too small, sizes still returned).
 *
 * status: 0 OK, -1 invalid input, -2 capacity too small, -3 bad handle or
 * registry full (as in the reference), -4 device (HIP) error or a failed
 * host allocation (no exception leaves these functions), 2 a solve that
 * reached its iteration limit (the solution holds the last iterate).
 *
 * Meshes and operators live on the MI355X: a mesh handle keeps the mesh, an
 * operator handle keeps an assembled (or uploaded) matrix resident in HBM
 * with its solver workspace, in place of the reference's retained sparse
 * factorisation (fortfem_complex_factor_csc / fortfem_complex_solve /
 * fortfem_factor_free); solves run Jacobi-PCG on the device.  All calls use
 * one process-wide device context (device 0); the registries are guarded by
 * a lock.
 */
#ifndef FORTFEM_TET_H
#define FORTFEM_TET_H

#ifdef __cplusplus
extern "C" {
#endif

/* Retain a tetrahedral mesh: vertices [x0, y0, z0, x1, ...], tetrahedra
 * [v0, v1, v2, v3, ...] zero-based (the reference's C convention, cf.
 * fortfem_triangle_mesh_create in include/fortfem.h). */
void fortfem_tetra_mesh_create(int n_vertices, const double *vertices, int n_tetrahedra,
                               const int *tetrahedra, int *handle, int *status);
void fortfem_tetra_mesh_free(int handle, int *status);

/* assemble_tetra_lagrange_stiffness_csc
 * (src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:276-334) on a
 * retained mesh, exported zero-based: degree 1..4. */
void fortfem_tetra_lagrange_stiffness_csc(int mesh, int degree, int quadrature_degree,
                                          double stiffness_coefficient,
                                          double mass_coefficient, int nnz_capacity,
                                          int *n_dofs, int *nnz, int *col_ptr, int *row_ind,
                                          double *values, int *status);

/* assemble_tetra_nedelec_curl_mass_csc
 * (src/assembly/assembly_tetra_nedelec_3d.f90:2042-2112), order 1..6. */
void fortfem_tetra_nedelec_curl_mass_csc(int mesh, int order, double curl_coefficient,
                                         double mass_coefficient, int nnz_capacity,
                                         int *n_dofs, int *nnz, int *col_ptr, int *row_ind,
                                         double *values, int *status);

/* Retained operators.  From a mesh: the H1 operator (boundary = 1 imposes
 * u = 0 on the boundary nodes, build_dirichlet_data,
 * fortfem_tetra_lagrange_solver_3d.f90:551-610) or the Nedelec curl-mass
 * operator (boundary = 1 imposes a zero tangential trace,
 * fortfem_tetra_nedelec_solver_3d.f90:416-470), assembled on the device and
 * kept there; n_dofs is the global dof count (right-hand sides and solutions
 * are full dof vectors, constrained entries 0). */
void fortfem_tetra_lagrange_operator(int mesh, int degree, int quadrature_degree,
                                     double stiffness_coefficient, double mass_coefficient,
                                     int boundary, int *handle, int *n_dofs, int *status);
void fortfem_tetra_nedelec_operator(int mesh, int order, double curl_coefficient,
                                    double mass_coefficient, int boundary, int *handle,
                                    int *n_dofs, int *status);
/* From a zero-based real CSC matrix (symmetric positive definite for the
 * solves): uploaded once, kept resident. */
void fortfem_real_operator_csc(int n, int nnz, const int *col_ptr, const int *row_ind,
                               const double *values, int *handle, int *status);

/* Solve operator x = rhs from x = 0 by Jacobi-PCG to the relative residual
 * tolerance (precision 0 FP64, the reference's recurrence; 1 FP32 iterations
 * with FP64 residual refreshes).  n must be the operator's size. */
void fortfem_operator_solve(int handle, int n, double tolerance, int max_iterations,
                            int precision, const double *rhs, double *solution,
                            int *iterations, int *status);
/* A zero-based complex CSC matrix (values as (re, im) double pairs, the
 * layout of fortfem_complex) kept resident, in place of
 * fortfem_complex_factor_csc (include/fortfem.h); solved from x = 0 by
 * Jacobi-COCG, which needs A = A^T (complex symmetric, as the PML curl-curl
 * operators are), instead of fortfem_complex_solve's direct solve: a matrix
 * that is not complex symmetric (to 1e-12 of its largest entry) is refused
 * with status -1 at upload.  Status as fortfem_operator_solve; -1 also for a
 * real operator's handle.
 * fortfem_operator_free releases either kind. */
void fortfem_complex_operator_csc(int n, int nnz, const int *col_ptr, const int *row_ind,
                                  const double *values, int *handle, int *status);
void fortfem_complex_operator_solve(int handle, int n, double tolerance, int max_iterations,
                                    const double *rhs, double *solution, int *iterations,
                                    int *status);
void fortfem_operator_free(int handle, int *status);

#ifdef __cplusplus
}
#endif

#endif
