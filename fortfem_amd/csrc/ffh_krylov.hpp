// ffh_krylov.hpp -- solver entry points shared by the C-ABI translation units.
#pragma once
#include "ffh_internal.hpp"

namespace ffh {
void ensure_workspace(Context *ctx, Matrix &M);
void ensure_val32(Context *ctx, DevMatrix &A);
void invalidate_graphs(Workspace &w);
void build_rowblocks(Context *ctx, DevMatrix &A, const std::vector<int> &ptr);
void spmv64(Context *ctx, DevMatrix &A, const double *x, double *y);
void spmv32(Context *ctx, DevMatrix &A, const float *x, float *y);
void solve_fp64(Context *ctx, Matrix &M, const double *b, double *x,
                const ffh_solver_options &o, ffh_solver_stats &st);
void solve_mixed(Context *ctx, Matrix &M, const double *b, double *x,
                 const ffh_solver_options &o, ffh_solver_stats &st);
int select_spmv_variant(Context *ctx, Matrix &M, bool f32, int request);
double bench_spmv(Context *ctx, Matrix &M, int precision, int reps,
                  double *bytes);
// allocation-free FP64 solve on device arrays (the scratch array)
void scratch_setup(int device);
int noalloc_variant(int n, int64_t nnz);
void solve_fp64_noalloc(Context *ctx, int n, int64_t nnz, const int *col_ptr,
                        const int *row_idx, const double *val, const double *b, double *x,
                        const ffh_solver_options &o, ffh_solver_stats &st);
} // namespace ffh

namespace ffh {
// problems (ffh_problem.hip)
void problem_pattern(Problem *P, int family, int order);
void problem_nedelec_pml(Problem *P, int kind, const double *stretch, double k0,
                         double *values);
void problem_nedelec_pml_jvp(Problem *P, int kind, const double *stretch,
                             const double *stretch_dot, double k0, double k0_dot,
                             const double *vertices_dot, double *values_dot);
// complex-symmetric operators and their Jacobi-COCG solve (ffh_complex.hip)
struct ZMatrix;
ZMatrix *zmatrix_create(Context *ctx, int n, int64_t nnz, int base, const int *col_ptr,
                        const int *row_idx, const double *values);
void zmatrix_destroy(ZMatrix *m);
int zmatrix_size(const ZMatrix *m);
void zmatrix_solve(ZMatrix *m, const double *b, double *x, double tolerance, int max_iterations,
                   int *iterations, double *relative_residual, int *converged);
void problem_nedelec_pml_vjp(Problem *P, int kind, const double *stretch, double k0,
                             const double *values_bar, double *vertices_bar,
                             double *stretch_bar, double *k0_bar);
void problem_assemble_jvp(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                          double mass, const double *vertices_dot, const double *coeff_dot,
                          double mass_dot, double *values_dot);
void problem_assemble_vjp(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                          double mass, const double *values_bar, double *vertices_bar,
                          double *coeff_bar, double *mass_bar);
void problem_assemble(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                      double mass, int precision);
void problem_load(Problem *P, int qdeg, int kind, const double *params,
                  int64_t np);
void problem_dirichlet(Problem *P, int mode, const double *values);
void problem_reduce(Problem *P);
void problem_solve(Problem *P, const ffh_solver_options &o, double *solution,
                   ffh_solver_stats &st);
} // namespace ffh

namespace ffh {
// SELL-64 copy of a matrix's gather arrays (ffh_krylov.hip): the structure
// (false when the matrix is too uneven for it) and a value permutation into
// caller buffers (either output may be null).
bool sell_structure(Context *ctx, DevMatrix &A);
// the matrix's own SELL value copy of one precision, made current
void ensure_sell(Context *ctx, DevMatrix &A, bool f32);
void sell_values(Context *ctx, const DevMatrix &A, const double *val, double *out64,
                 float *out32);
// column tiles of the SELL copy (DevMatrix::tile_*), built once; tiles_ok
// false when a group reads more than 2^16 columns
void ensure_sell_tiles(Context *ctx, DevMatrix &A);
// bytes of one SpMV launch of variant vi in that variant's format
double variant_bytes(const DevMatrix &A, int vi, int sv);
double spmv_launch_bytes(Context *ctx, Matrix &M, bool f32);
// whether a Jacobi-PCG iteration on n rows takes three launches
bool three_launch(int64_t n, bool f32);
} // namespace ffh

namespace ffh {
// differentiable products over the CSC arrays (ffh_krylov.hip):
// y = A x_dot + A_dot x (either term may be absent), and
// val_bar = scale (y_bar_row x_col) per entry, x_bar = A^T y_bar (either
// output may be absent)
void csc_jvp(Context *ctx, DevMatrix &A, const double *vdot, const double *x,
             const double *xdot, double *y);
void csc_vjp(Context *ctx, const DevMatrix &A, const double *x, const double *ybar, double scale,
             double *vbar, double *xbar);
} // namespace ffh
