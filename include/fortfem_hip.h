/*
 * fortfem_hip.h -- C-ABI of libfortfem_hip.so, the MI355X (gfx950) drop-in
 * for fortfem's sparse-assembly + Jacobi-PCG path.
 *
 * Plain C types only (int, int64_t, double, opaque handles), so Fortran binds
 * it with ISO_C_BINDING and Python with ctypes.  Every entry point names the
 * reference interface it replaces (lazy-fortran/fortfem, path:line).
 *
 * Status codes follow the reference C-API (include/fortfem.h:22-23,141-143):
 *    0 ok, -1 invalid input, -2 capacity too small (sizes still returned),
 *   -3 bad or exhausted handle; this library adds -4 device (HIP) error and
 *   -5 admission rejected (serving).  ffh_last_error() gives the message.
 *
 * Index base: arrays that cross the boundary carry an explicit `base`
 * argument (1 = Fortran/fortsparse, 0 = C), as in the reference's
 * zero-based C export (src/bindings/fortfem_capi_mesh.f90:560-567).
 *
 * Threading: one ffh_context per host thread; every handle belongs to the
 * context that created it and all device work of a context is ordered on its
 * one HIP stream.
 */
#ifndef FORTFEM_HIP_H
#define FORTFEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFH_OK 0
#define FFH_ERR_INVALID (-1)
#define FFH_ERR_CAPACITY (-2)
#define FFH_ERR_HANDLE (-3)
#define FFH_ERR_DEVICE (-4)
#define FFH_ERR_ADMISSION (-5)

/* element families */
#define FFH_FAMILY_H1 1      /* tetrahedral Lagrange, degree 1..4        */
#define FFH_FAMILY_NEDELEC 2 /* tetrahedral first-kind Nedelec, k 1..6    */
#define FFH_FAMILY_P1_2D 3   /* triangular P1 on a 2D mesh               */

/* coefficient kinds for ffh_assemble */
#define FFH_COEFF_SCALAR 0      /* coeff[0]                              */
#define FFH_COEFF_CELL 1        /* coeff[T], one scalar per cell          */
#define FFH_COEFF_CELL_TENSOR 2 /* coeff[9*T], column-major 3x3 per cell  */
#define FFH_COEFF_POINT_TENSOR 3 /* coeff[9*nq*T]: column-major 3x3 per     */
                                 /* point of ffh_nedelec_load_points (the   */
                                 /* weighted element's rule), cell-major    */

/* load kinds for ffh_load */
#define FFH_SOURCE_CONSTANT 0    /* params[0] = f                          */
#define FFH_SOURCE_SIN_PRODUCT 1 /* f = p0 sin(p1 x) sin(p2 y) sin(p3 z)    */
#define FFH_SOURCE_SAMPLES 2     /* params = f(nq, T) at the Duffy points   */
#define FFH_SOURCE_DOF_VECTOR 3  /* params = rhs(n_global) directly         */

/* precision */
#define FFH_FP64 0
#define FFH_MIXED 1 /* FP32 matrix/vectors, FP64 x, reductions and residual refreshes */
#define FFH_FP32 1  /* element contractions in FP32 (ffh_assemble)        */

/* which matrix / vector of a problem */
#define FFH_FULL 0
#define FFH_REDUCED 1

typedef struct ffh_context ffh_context;
typedef struct ffh_problem ffh_problem;
typedef struct ffh_matrix ffh_matrix;
typedef struct ffh_server ffh_server;

/* Mirrors solver_options_t (src/solvers/advanced_solvers.f90:32-44) and the
 * solver_options() defaults (:1430-1465): method 'auto'->'pcg', tolerance
 * 1e-6, max_iterations 1000, preconditioner 'none', tolerance_type
 * 'relative'.  Fields after max_iterations are additions of this library. */
typedef struct ffh_solver_options {
    int method;          /* 0 cg, 1 pcg ('auto' is pcg, :159-161)        */
    int preconditioner;  /* 0 none, 1 jacobi                             */
    int tolerance_type;  /* 0 relative (tol*||b-Ax0||), 1 absolute       */
    int max_iterations;
    double tolerance;
    int precision;       /* FFH_FP64 or FFH_MIXED                          */
    int breakdown_mode;  /* 0 reference |pAp| < 1e-14 (:897),
                            1 scale-invariant |pAp| <= 1e-14 |rz|         */
    double inner_tolerance; /* accepted for compatibility, not used      */
    int max_refinements;    /* accepted for compatibility, not used      */
    int check_interval;     /* iterations per chunk (host poll; MIXED: per FP64 refresh), 0: 16 */
    int use_graph;          /* replay iteration chunks from a HIP graph   */
} ffh_solver_options;

/* Mirrors solver_stats_t (:47-57): converged, iterations, final_residual,
 * solve_time; the rest are additions. */
typedef struct ffh_solver_stats {
    int converged;
    int iterations;        /* Krylov iterations (MIXED: summed inner)    */
    double final_residual; /* recursive ||r|| (FP64) / true ||b-Ax|| (MIXED) */
    double solve_time;     /* seconds, HIP events on the context stream   */
    int refinements;       /* MIXED: FP64 residual refreshes            */
    double true_residual;  /* ||b - A x||_2 recomputed in FP64 at exit    */
    double initial_residual;
} ffh_solver_stats;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the reference is host-only); owns the device,
 * stream and events that every later call uses. */
int ffh_init(int device, ffh_context **ctx);
int ffh_finalize(ffh_context *ctx);
const char *ffh_last_error(const ffh_context *ctx);
/* fills the defaults of solver_options() (advanced_solvers.f90:1430-1465) */
int ffh_solver_options_default(ffh_solver_options *opts);

/* ---- meshes -------------------------------------------------------------- */
/* generate_structured_tetra_box_mesh (src/mesh/structured_tetra_box_mesh.f90:
 * 12-70): bounds(3,2) column-major, counts(3); vertices(3,(n+1)^3), 1-based
 * tetrahedra(4, 6 n^3).  ffh_box_mesh_size gives the sizes. */
int ffh_box_mesh_size(const int counts[3], int *nvert, int *ntet);
int ffh_box_mesh(const double bounds[6], const int counts[3],
                 double *vertices, int *tetrahedra);
/* unit_square_mesh (src/fortfem_api_mesh.f90:128-139; mesh_2d.f90:90-148,
 * 553-603): vertices(2,n^2), triangles(3,2(n-1)^2), boundary flags(n^2). */
int ffh_unit_square_mesh(int n, double *vertices, int *triangles,
                         int *is_boundary_vertex);

/* tetra_duffy_quadrature (src/quadrature/tetra_duffy_quadrature.f90:12-51):
 * the collapsed Gauss rule of polynomial degree p, nq = ((p+4)/2)^3 points in
 * the reference's loop order.  With x == NULL only nq is returned (host
 * routine, used by hosts that sample a source callback at the points). */
int ffh_tetra_duffy(int p, int *nq, double *x, double *y, double *z,
                    double *w);

/* ---- Nedelec host routines ------------------------------------------------ */
/* build_tetra_nedelec_dof_map (src/elements/tetra_nedelec_global_dof_map.f90:
 * 15-105): tets (4 x ntet, 1-based ids); global_dofs (n x ntet,
 * n = k(k+2)(k+3)/2) in base 0 or 1; edge_orient (6 x ntet, +1/-1);
 * face_perm (3 x 4 x ntet, 1..3); any output may be NULL.  *ndofs receives
 * the global count.  Host routine. */
int ffh_nedelec_dof_map(int k, int ntet, const int *tets, int base, int *global_dofs,
                        int *edge_orient, int *face_perm, int *ndofs);
/* evaluate_tetra_nedelec_first_kind (src/elements/tetra_nedelec_arbitrary_order.f90:
 * 75-106) and evaluate_tetra_nedelec_first_order (tetra_nedelec_first_order.f90:
 * 11-51) at npts reference points (3 x npts): values and curls
 * (3 x n x npts).  Host routine. */
int ffh_nedelec_basis(int k, int npts, const double *points, double *values,
                      double *curls);
/* The quadrature of the Nedelec vector load (assemble_tetra_nedelec_load_element,
 * src/assembly/assembly_tetra_nedelec_3d.f90:2516-2553: the 4-point degree-2
 * rule at k = 1; assemble_tetra_nedelec_load_element_order :525-582: the
 * Duffy rule of degree 2k + 2 otherwise), points (4 x nq: x, y, z, weight)
 * in the reference's order.  With points == NULL only nq is returned.  Hosts
 * sample a vector source callback at the mapped points for
 * FFH_SOURCE_SAMPLES.  Host routine. */
int ffh_nedelec_load_points(int k, int *nq, double *points);

/* ---- problems: mesh + space + operator resident in HBM -------------------- */
/* Takes the (mesh_vertices, tetrahedra) arguments of
 * assemble_tetra_lagrange_stiffness_csc (src/assembly/
 * assembly_tetra_lagrange_arbitrary_order_3d.f90:276-285) and
 * assemble_tetra_nedelec_curl_mass_csc (assembly_tetra_nedelec_3d.f90:
 * 2042-2049); tetrahedra are 1-based.  dim = 3 (tets, 4 ids per cell) or
 * 2 (triangles, 3 ids per cell, family FFH_FAMILY_P1_2D). */
int ffh_problem_create(ffh_context *ctx, int dim, int nvert,
                       const double *vertices, int ncell, const int *cells,
                       ffh_problem **out);
int ffh_problem_free(ffh_problem *prob);

/* Symbolic stage shared by the CSC assemblers: global numbering of
 * build_tetra_lagrange_dof_map (src/elements/tetra_lagrange_global_dof_map.f90:
 * 13-79) or build_tetra_nedelec_dof_map (tetra_nedelec_global_dof_map.f90:
 * 15-105), the CSC pattern that csc_from_triplet would produce (ascending
 * rows, structural zeros kept), an element -> value-slot map and an element
 * colouring in which no two cells of one colour share a dof. */
int ffh_pattern_build(ffh_problem *prob, int family, int order,
                      int *n_global, int64_t *nnz);
int ffh_problem_sizes(const ffh_problem *prob, int *n_global, int64_t *nnz,
                      int *n_free, int64_t *nnz_free, int *n_local,
                      int *n_colors);
/* global_dofs(n_local, ncell) as the reference dof map returns it */
int ffh_global_dofs_download(ffh_problem *prob, int base, int *global_dofs);
/* The element colouring of ffh_pattern_build (no reference counterpart: the
 * reference adds its triplets one element at a time): color_start
 * (n_colors + 1 offsets into cells) and cells (ncell cell ids of the given
 * base, colour by colour, in the order the scatter visits them).  No two
 * cells of one colour share a dof; the scatter relies on it. */
int ffh_problem_colouring(ffh_problem *prob, int base, int *color_start, int *cells);

/* Numeric assembly into the fixed pattern: the element loop + triplets +
 * csc_from_triplet of assemble_tetra_lagrange_stiffness_csc (:276-334, per
 * cell kappa as the stiffness_coefficient of :47-48),
 * assemble_tetra_nedelec_curl_mass_csc (:2042-2112; coeff = curl alpha,
 * mass = beta) and assemble_tetra_nedelec_weighted_csc (:1870-2000; coeff =
 * tensor K: constant per cell with FFH_COEFF_CELL_TENSOR, or the reference's
 * coefficient(x_q) sampled at every point x_q = v1 + J xi_q of the rule that
 * ffh_nedelec_load_points returns, FFH_COEFF_POINT_TENSOR: the curl part is
 * then a dense per-cell C^T blockdiag(M_q) C product on MFMA).  Tensors must
 * be symmetric.  precision FFH_FP32 runs the element contraction in FP32. */
int ffh_assemble(ffh_problem *prob, int quadrature_degree, int coeff_kind,
                 const double *coeff, double mass_coefficient, int precision);
/* assemble_tetra_lagrange_stiffness_csc_jvp (src/assembly/
 * assembly_tetra_lagrange_arbitrary_order_3d.f90:336-386) for
 * FFH_FAMILY_H1 and assemble_tetra_nedelec_curl_mass_csc_jvp
 * (assembly_tetra_nedelec_3d.f90:2114-2180) for FFH_FAMILY_NEDELEC
 * (coeff = curl alpha, mass = beta; the quadrature degree is not used):
 * the tangent of the full matrix for tangents of the vertices
 * (vertices_dot(3, nv), NULL = 0), of the stiffness / curl coefficient
 * (coeff_dot: 1 value for FFH_COEFF_SCALAR, one per cell for FFH_COEFF_CELL;
 * NULL = 0) and of the mass coefficient.  values_dot: nnz host values on the
 * problem's pattern (the primal's CSC order; ffh_problem_matrix +
 * ffh_matrix_download give the pattern).  coeff/mass as in ffh_assemble. */
int ffh_assemble_jvp(ffh_problem *prob, int quadrature_degree, int coeff_kind,
                     const double *coeff, double mass, const double *vertices_dot,
                     const double *coeff_dot, double mass_dot, double *values_dot);
/* assemble_tetra_lagrange_stiffness_csc_vjp (:388-471) and
 * assemble_tetra_nedelec_curl_mass_csc_vjp (assembly_tetra_nedelec_3d.f90:
 * 2182-2287): from the cotangent of the full matrix's values (nnz),
 * vertices_bar(3, nv), coeff_bar (1 value, or one per cell for
 * FFH_COEFF_CELL) and mass_bar. */
int ffh_assemble_vjp(ffh_problem *prob, int quadrature_degree, int coeff_kind,
                     const double *coeff, double mass, const double *values_bar,
                     double *vertices_bar, double *coeff_bar, double *mass_bar);

/* Complex curl-curl PML operator on the problem's Nedelec pattern (csc_z_t
 * values): assemble_tetra_nedelec_pml_csc (src/assembly/
 * assembly_tetra_nedelec_3d.f90:1613-1685) for kind 0, stretch = 3 complex
 * per cell; assemble_tetra_nedelec_curvilinear_pml_csc (:1372-1444) for
 * kind 1, stretch = complex 3 x 3 column-major per cell.  Complex numbers
 * are (re, im) double pairs; values receives nnz pairs in the order of the
 * pattern (ffh_problem_matrix + ffh_matrix_download give it). */
int ffh_nedelec_pml_values(ffh_problem *prob, int kind, const double *stretch,
                           double wave_number, double *values);
/* Tangent of that operator (assemble_tetra_nedelec_pml_csc_jvp, :1687-1768;
 * assemble_tetra_nedelec_curvilinear_pml_csc_jvp, :1446-1521) along the
 * stretch tangent (same layout as stretch), the wave-number tangent and the
 * vertex tangents (3 per vertex, or NULL for none); values_dot receives nnz
 * complex pairs on the pattern. */
int ffh_nedelec_pml_values_jvp(ffh_problem *prob, int kind, const double *stretch,
                               const double *stretch_dot, double wave_number,
                               double wave_number_dot, const double *vertices_dot,
                               double *values_dot);
/* A complex CSC matrix resident in HBM (values as (re, im) pairs, base 0 or
 * 1), in place of the reference's retained complex factorisation
 * (fortfem_complex_factor_csc, src/bindings/fortfem_capi_solver.f90:10-70),
 * and its solve A x = b from x = 0 by Jacobi-COCG to
 * ||b - A x|| <= tolerance ||b|| (fortfem_complex_solve, :72-124, which
 * solves directly).  COCG needs a complex-symmetric A (A = A^T, as the PML
 * curl-curl operators are). */
typedef struct ffh_zmatrix ffh_zmatrix;
int ffh_zmatrix_upload(ffh_context *ctx, int n, int64_t nnz, int base, const int *col_ptr,
                       const int *row_idx, const double *values, ffh_zmatrix **out);
int ffh_zmatrix_solve(ffh_zmatrix *mat, const double *b, double *x, double tolerance,
                      int max_iterations, int *iterations, double *relative_residual,
                      int *converged);
int ffh_zmatrix_free(ffh_zmatrix *mat);
/* Cotangent of that operator (assemble_tetra_nedelec_pml_csc_vjp, :1770-1868;
 * assemble_tetra_nedelec_curvilinear_pml_csc_vjp, :1523-1611) for nnz
 * complex seeds on the pattern, in the reference's convention (sum of
 * seed * conj(d value / d input); real part for the vertices, 3 per vertex,
 * and the wave number; stretch_bar complex in the stretch layout). */
int ffh_nedelec_pml_values_vjp(ffh_problem *prob, int kind, const double *stretch,
                               double wave_number, const double *values_bar,
                               double *vertices_bar, double *stretch_bar,
                               double *wave_number_bar);
/* assemble_tetra_lagrange_scalar_load (:705-774) / _load_samples (:473-528)
 * with the source given as a closed form, samples or a dof vector.  For the
 * Nedelec family: assemble_tetra_nedelec_vector_load(_order)
 * (src/assembly/assembly_tetra_nedelec_3d.f90:71-127, 2002-2040) with a
 * constant vector (FFH_SOURCE_CONSTANT, params[0..2]) or samples
 * f(3, nq, ncell) at the ffh_nedelec_load_points mapped into each cell
 * (FFH_SOURCE_SAMPLES); quadrature_degree is ignored (the reference's rule). */
int ffh_load(ffh_problem *prob, int quadrature_degree, int source_kind,
             const double *params, int64_t nparams);

/* build_dirichlet_data (src/solvers/fortfem_tetra_lagrange_solver_3d.f90:
 * 551-610) for H1, the boundary-vertex set of find_boundary (mesh_2d.f90:
 * 553-603) for P1_2D, then the free-dof reduction of
 * sparse_direct_solve_constrained_real (src/solvers/fortfem_sparse_direct.f90:
 * 264-337).  values: NULL for g = 0, else values(n_global) (only constrained
 * entries are read).  For P1_2D, mode 1 applies the symmetric elimination of
 * solve_compiled_p1_problem (fortfem_solvers_laplacian.f90:88-97) to the
 * full matrix instead (identity rows, no reduction). */
int ffh_dirichlet(ffh_problem *prob, int mode, const double *values);
/* constrained flags (n_global), 1 = constrained */
int ffh_constrained_download(ffh_problem *prob, int *constrained);
/* rhs of the reduced system directly (free-dof order), e.g. random RHS */
int ffh_set_rhs(ffh_problem *prob, int which, const double *rhs);
int ffh_rhs_download(ffh_problem *prob, int which, double *rhs);

/* borrowed handle of the problem's full or reduced matrix */
int ffh_problem_matrix(ffh_problem *prob, int which, ffh_matrix **mat);

/* Solve the reduced system and scatter back (sparse_direct_solve_constrained
 * with sparse_solve_once replaced by solve_sparse's CG/PCG, :264-344):
 * solution(n_global), constrained entries = boundary values. */
int ffh_problem_solve(ffh_problem *prob, const ffh_solver_options *opts,
                      double *solution, ffh_solver_stats *stats);
/* The last solve's solution (n_global values, Dirichlet values included),
 * downloaded without solving again. */
int ffh_problem_solution(ffh_problem *prob, double *solution);

/* ---- CSC matrices (fortsparse csc_t) -------------------------------------- */
/* csc_t with col_ptr(n+1), row_idx(nnz), val(nnz) of the given base.  A
 * matrix that is not exactly symmetric also gets a transposed copy so that
 * products stay row-gathers. */
int ffh_matrix_upload(ffh_context *ctx, int n, int64_t nnz, int base,
                      const int *col_ptr, const int *row_idx,
                      const double *val, ffh_matrix **out);
int ffh_matrix_free(ffh_matrix *mat);
/* export with the capacity protocol of export_real_csc
 * (src/bindings/fortfem_capi_mesh.f90:545-571): -2 when capacity < nnz,
 * n and nnz are still returned. */
int ffh_matrix_download(ffh_matrix *mat, int base, int64_t capacity,
                        int *col_ptr, int *row_idx, double *val, int *n,
                        int64_t *nnz);
/* spmv (src/solvers/fortfem_sparse_matrix.f90:46-58): y = A x, host arrays */
int ffh_spmv(ffh_matrix *mat, const double *x, double *y);
/* solve_sparse (src/solvers/advanced_solvers.f90:135-195): x is the initial
 * guess on entry and the solution on exit (host arrays). */
int ffh_solve_sparse(ffh_matrix *mat, const double *b, double *x,
                     const ffh_solver_options *opts, ffh_solver_stats *stats);

/* ---- allocation-free solve -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 Jacobi-PCG / CG of ffh_solve_sparse on a CSC
 * matrix and vectors the caller already holds in device memory, with every
 * work vector in a fixed array of the library (about 56 bytes per row for up
 * to FFH_NOALLOC_MAX_ROWS rows plus 256 KiB, about 113 MiB in all, held by
 * each GPU the library runs on from ffh_init until the process ends).  No
 * hipMalloc, hipFree or allocating library routine is called.
 *   col_ptr(n+1), row_idx(nnz), val(nnz): 0-based device arrays, rows
 *     strictly ascending in each column, the matrix exactly symmetric (it is
 *     read as CSR; checked on the device, -1 otherwise);
 *   b(n) device, x(n) device: initial guess on entry, solution on exit.
 * opts->precision must be FFH_FP64 and use_graph is ignored; n above
 * FFH_NOALLOC_MAX_ROWS is refused with -1.  Work runs on the context's
 * stream.  All calls on one GPU share the array: calls that overlap on one
 * GPU (other threads, other contexts of the device) race and must be
 * serialised by the caller.  For a symmetric matrix the result is that of
 * ffh_solve_sparse with the SpMV variant ffh_noalloc_spmv_variant(n, nnz)
 * (bit for bit). */
#define FFH_NOALLOC_MAX_ROWS (1 << 21)
int ffh_solve_sparse_device_noalloc(ffh_context *ctx, int n, int64_t nnz, const int *col_ptr,
                                    const int *row_idx, const double *val, const double *b,
                                    double *x, const ffh_solver_options *opts,
                                    ffh_solver_stats *stats);
/* The SpMV variant (ffh_matrix_spmv_variant numbering) that
 * ffh_solve_sparse_device_noalloc uses for a matrix of n rows and nnz
 * entries: vector CSR, lanes per row from the mean row length. */
int ffh_noalloc_spmv_variant(int n, int64_t nnz);

/* ---- served solves (continuous batching) ----------------------------------
 * No single reference routine: the request loop serves what the reference
 * does one call at a time (assemble_* + solve_sparse per right-hand side,
 * advanced_solvers.f90:135-195, 831-925).  A server keeps symbolic problems
 * per (mesh_key, family, order) and, per admitted request, its assembled
 * reduced operator and PCG state in HBM.  All active systems and their
 * right-hand sides advance together, one launch per PCG step; requests join
 * and leave between steps; admission is by device-memory footprint.
 * Solutions are over the free dofs (zero Dirichlet data). */
/* budget_bytes: device memory for admitted requests; max_systems: batch
 * slots; chunk: PCG iterations per step between FP64 residual refreshes. */
int ffh_server_create(ffh_context *ctx, int64_t budget_bytes, int max_systems, int chunk,
                      ffh_server **server);
int ffh_server_free(ffh_server *server);
/* Builds (once per key) the pattern, colouring and Dirichlet reduction of a
 * tetrahedral mesh for FFH_FAMILY_H1 or FFH_FAMILY_NEDELEC; returns the
 * free-dof count a request's right-hand sides must have. */
int ffh_server_prepare(ffh_server *server, int mesh_key, int family, int order, int nvert,
                       const double *vertices, int ncell, const int *cells, int *n_free,
                       int64_t *nnz_free);
/* Queues a request: per-cell coefficient (H1 stiffness kappa / Nedelec curl
 * alpha), mass coefficient, nrhs (1..8) right-hand sides (n_free x nrhs,
 * column-major), relative tolerance and iteration cap per right-hand side.
 * FFH_ERR_ADMISSION if the request can never fit the budget. */
int ffh_server_submit(ffh_server *server, int mesh_key, int family, int order,
                      const double *cell_coeff, double mass, int nrhs, const double *rhs,
                      double tol, int max_iterations, int64_t *ticket);
/* Admits queued requests that fit, runs one step (chunk iterations + FP64
 * refresh) of every active system, retires finished requests. */
int ffh_server_step(ffh_server *server, int *n_active, int *n_queued, int *n_finished);
/* 0 and the results (then the ticket is released) if finished, 1 if not yet. */
int ffh_server_poll(ffh_server *server, int64_t ticket, double *solution, int *iterations,
                    int *converged, double *rel_residual);
/* Steps the server (at most max_steps steps, every active request
 * advancing) until the ticket's request has finished, then as
 * ffh_server_poll: 0 and the results, or 1 if it is still running. */
int ffh_server_wait(ffh_server *server, int64_t ticket, int max_steps, double *solution,
                    int *iterations, int *converged, double *rel_residual);
int ffh_server_stats(const ffh_server *server, int64_t *used_bytes, int64_t *iterations_run);
/* Progress of a request as of the last step: per right-hand side the state
 * (-1 queued, 0 iterating, 1 converged, 2 stopped, 4 waiting for the FP64
 * refresh, 5 admitted, not started), iterations and the relative FP64
 * residual of the last refresh. */
int ffh_server_progress(const ffh_server *server, int64_t ticket, int *state, int *iterations,
                        double *rel_residual);
/* Iteration precision of requests admitted from now on: 0 (default) FP32
 * iterations with FP64 refreshes, switching a request to FP64 iterations when
 * its FP64 residual stalls (no refresh in a window of 512 iterations, at
 * least 4 refreshes, falls below 0.9 of its best); 1 FP64 iterations; 2 FP64
 * for Nedelec
 * requests and 0 for H1.  policy = -1 only reads *switches (requests moved to
 * FP64 so far). */
int ffh_server_set_precision(ffh_server *server, int policy, int64_t *switches);
/* Replay each server step's chunk (the iterations and the FP64 refresh of
 * every admitted request) from a HIP graph captured on the context stream,
 * re-captured when the batch's composition or the chunk length changes
 * (enable != 0), or launch the kernels one by one (0, the default).  The
 * results are the same either way.  No reference counterpart. */
int ffh_server_use_graph(ffh_server *server, int enable);
/* Batched SpMV of each system type (pattern, right-hand sides, iteration
 * precision): mode -1 (the default) times, the first time a type is batched,
 * the plain SELL-64 product (one lane per row), the one that stages each
 * item's column tile of p in LDS (offered when the pattern's tiles fit
 * 64 KiB at R values per column) and the row-split ones (2, 4 or 8 lanes per
 * row, each lane a power-of-two share of the R right-hand sides) on that
 * system alone, and keeps the fastest if it beats the plain product by 2 %;
 * 0 always plain; 1 tiled wherever the tiles fit; 2, 4, 8 row-split with
 * that many lanes per row wherever R allows it.  All products give the same
 * bits.  Applies to types chosen from now on.  No reference counterpart (a
 * tuning control). */
int ffh_server_set_tiling(ffh_server *server, int mode);
/* System types chosen so far (*types) and how many of them run the tiled
 * product (*tiled). */
int ffh_server_tiling_stats(const ffh_server *server, int *types, int *tiled);
/* The same, with how many of them run a row-split product (*split); any
 * pointer may be NULL. */
int ffh_server_product_forms(const ffh_server *server, int *types, int *tiled, int *split);

/* ---- differentiable variants --------------------------------------------- */
/* spmv_jvp (src/solvers/fortfem_sparse_matrix.f90:60-91): y_dot = A x_dot +
 * A_dot x with A_dot given by its CSC values (nnz, the matrix's pattern).
 * Host arrays. */
int ffh_spmv_jvp(ffh_matrix *mat, const double *x, const double *values_dot,
                 const double *x_dot, double *y_dot);
/* spmv_vjp (:93-126): values_bar(k) = y_bar(row_k) x(col_k),
 * x_bar = A^T y_bar.  Host arrays. */
int ffh_spmv_vjp(ffh_matrix *mat, const double *x, const double *y_bar,
                 double *values_bar, double *x_bar);
/* Sparse counterparts of pcg_solve_jvp / cg_solve_jvp
 * (src/solvers/advanced_solvers.f90:387-419, 458-489): for a converged
 * x = A^-1 b, x_dot solves A x_dot = b_dot - A_dot x (x_dot from 0) with the
 * options' method, preconditioner and precision; stats report that solve
 * (converged = 0 is the reference's status 2).  A must be symmetric. */
int ffh_solve_sparse_jvp(ffh_matrix *mat, const double *x, const double *values_dot,
                         const double *b_dot, const ffh_solver_options *opts,
                         double *x_dot, ffh_solver_stats *stats);
/* pcg_solve_vjp / cg_solve_vjp (:421-456, 491-522): lambda solves
 * A^T lambda = x_bar; b_bar = lambda and values_bar(k) = -lambda(row_k)
 * x(col_k), the cotangent A_bar = -lambda x^T on the stored entries. */
int ffh_solve_sparse_vjp(ffh_matrix *mat, const double *x, const double *x_bar,
                         const ffh_solver_options *opts, double *values_bar,
                         double *b_bar, ffh_solver_stats *stats);

/* ---- measurement ---------------------------------------------------------- */
/* Times `reps` back-to-back SpMV launches on the context stream with HIP
 * events; precision FFH_FP64 or FFH_FP32.  avg_ms = mean per launch,
 * bytes = ffh_matrix_spmv_bytes of the variant timed. */
int ffh_bench_spmv(ffh_matrix *mat, int precision, int reps, double *avg_ms,
                   double *bytes);
/* Bytes one SpMV launch moves in the storage format of the matrix's SpMV
 * variant for the precision (no reference counterpart: a measurement aid).
 * CSR variants: nnz (s + 4) + 4 (n + 1) + 2 n s for value size s; SELL-64:
 * its padded entries at s + 4 bytes, slice offsets and row map, plus 2 n s;
 * tiled SELL-64: padded entries at s + 2 bytes (16-bit tile positions),
 * slice offsets, row map, tile offsets and tile column lists, plus 2 n s. */
int ffh_matrix_spmv_bytes(ffh_matrix *mat, int precision, double *bytes);
/* Launches per Jacobi-PCG iteration of a solve with n rows in precision
 * FFH_FP32 (the mixed solve's iterations) or FFH_FP64: 3 (direction update,
 * SpMV with A p and p.A p partials, x/r/z update), or 2 (direction formed in
 * the SpMV's gathers, then the update) in a process started with
 * FFH_PCG_LAUNCHES=2.  No reference counterpart: it says which kernels a
 * solve runs.  Returns 2 or 3, or FFH_ERR_INVALID. */
int ffh_pcg_launches_per_iteration(int64_t n, int precision);
/* SpMV variant of a matrix (no reference counterpart: a tuning control).
 * request >= 0 forces variant `request` (0 .. FFH_SPMV_VARIANTS-1) for the
 * precision; request < 0 keeps the current choice, running the autotuner
 * (HIP-event timing of every variant) if none is made yet.  *chosen receives
 * the variant in use. */
#define FFH_SPMV_VARIANTS 13
int ffh_matrix_spmv_variant(ffh_matrix *mat, int precision, int request, int *chosen);
/* Kernel timing inside solves: when enabled, every 8th SpMV launch of the
 * iterations (k_pcg_apdot in the three-launch form, k_pcg_dirspmv in the
 * two-launch form, see ffh_pcg_launches_per_iteration) of a solve (which
 * 0 = FP64 iterations of the FP64 solve, or the FP64 residual refresh of a
 * mixed solve; 1 = FP32 iterations of a mixed solve) is bracketed by HIP
 * events on the context stream (sampling
 * keeps the event cost off the solve's critical path);
 * ffh_last_solve_profile returns the timed launch count and their mean
 * duration (ms) for the last solve. */
int ffh_set_profiling(ffh_context *ctx, int enable);
int ffh_last_solve_profile(ffh_context *ctx, int which, int *count,
                           double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
