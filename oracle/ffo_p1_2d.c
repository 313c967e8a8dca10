/*
 * ffo_p1_2d.c -- oracle restatement of the 2D P1 Poisson path (unit square).
 * Test infrastructure only (see ffo.h).
 */
#include "ffo.h"

#include <float.h>
#include <math.h>

/* assemble_scalar_p1_matrix (src/fortfem_forms_simple.f90:1678-1728):
 * local = kappa area G^T G + mu area M/12 with M = 1 + I; dense n x n
 * column-major accumulation, column-outer / row-inner per triangle. */
int ffo_p1_2d_assemble_dense(int nvert, const double *v, int ntri,
                             const int *tri, double mass, double stiff,
                             double *matrix)
{
    for (size_t i = 0; i < (size_t)nvert * nvert; ++i) matrix[i] = 0.0;
    for (int t = 0; t < ntri; ++t) {
        double x[3], y[3];
        for (int k = 0; k < 3; ++k) {
            int id = tri[3 * t + k];
            if (id < 1 || id > nvert) return 3;
            x[k] = v[2 * (id - 1)];
            y[k] = v[2 * (id - 1) + 1];
        }
        double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
        if (fabs(det) <= 64.0 * DBL_EPSILON) return 3;
        double area = 0.5 * fabs(det);
        double g[2][3] = {{(y[1] - y[2]) / det, (y[2] - y[0]) / det,
                           (y[0] - y[1]) / det},
                          {(x[2] - x[1]) / det, (x[0] - x[2]) / det,
                           (x[1] - x[0]) / det}};
        double local[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double gtg = g[0][a] * g[0][b] + g[1][a] * g[1][b];
                double m = (a == b) ? 2.0 : 1.0;
                local[a][b] = stiff * area * gtg + mass * area * m / 12.0;
            }
        for (int b = 0; b < 3; ++b) {
            int col = tri[3 * t + b] - 1;
            for (int a = 0; a < 3; ++a) {
                int row = tri[3 * t + a] - 1;
                matrix[(size_t)row + (size_t)nvert * col] += local[a][b];
            }
        }
    }
    return 0;
}

/* assemble_scalar_p1_load (:1730-1761): b_i += f area / 3 */
int ffo_p1_2d_load(int nvert, const double *v, int ntri, const int *tri,
                   double load, double *vector)
{
    for (int i = 0; i < nvert; ++i) vector[i] = 0.0;
    for (int t = 0; t < ntri; ++t) {
        double x[3], y[3];
        for (int k = 0; k < 3; ++k) {
            int id = tri[3 * t + k];
            if (id < 1 || id > nvert) return 3;
            x[k] = v[2 * (id - 1)];
            y[k] = v[2 * (id - 1) + 1];
        }
        double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
        if (fabs(det) <= 64.0 * DBL_EPSILON) return 3;
        double area = 0.5 * fabs(det);
        for (int k = 0; k < 3; ++k)
            vector[tri[3 * t + k] - 1] += load * area / 3.0;
    }
    return 0;
}

/* Symmetric Dirichlet elimination of solve_compiled_p1_problem
 * (src/solvers/fortfem_solvers_laplacian.f90:88-97): for each boundary dof in
 * ascending order lift the rhs by its column, zero its row and column, put 1
 * on the diagonal and the boundary value in the rhs. */
int ffo_p1_2d_symmetric_dirichlet(int n, double *matrix, double *rhs,
                                  const int *is_boundary, double value)
{
    for (int d = 0; d < n; ++d) {
        if (!is_boundary[d]) continue;
        for (int i = 0; i < n; ++i)
            rhs[i] = rhs[i] - matrix[(size_t)i + (size_t)n * d] * value;
        for (int i = 0; i < n; ++i) matrix[(size_t)i + (size_t)n * d] = 0.0;
        for (int j = 0; j < n; ++j) matrix[(size_t)d + (size_t)n * j] = 0.0;
        matrix[(size_t)d + (size_t)n * d] = 1.0;
        rhs[d] = value;
    }
    return 0;
}
