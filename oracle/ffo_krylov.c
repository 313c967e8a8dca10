/*
 * ffo_krylov.c -- oracle restatement of solve_sparse's CG / Jacobi-PCG loops.
 * Test infrastructure only (see ffo.h).
 *
 * src/solvers/advanced_solvers.f90:
 *   solve_sparse                 :135-195  ('auto' -> 'pcg' :159-161)
 *   build_sparse_preconditioner  :1124-1155 (1/a_ii if |a_ii| > 1e-14 else 1)
 *   apply_preconditioner         :1185-1203
 *   cg_solve_operator / loop     :709-797  (breakdown -> iterations = iter-1)
 *   pcg_solve_operator / loop    :831-925  (breakdown leaves iterations)
 *   compute_solver_tolerance     :1063-1073
 *   finalize_solver_status       :1075-1086
 * Dot products are sequential sums in index order (Fortran dot_product).
 */
#include "ffo.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Breakdown test.  Mode 0 is the reference's absolute |pAp| < 1e-14
 * (:768, :897).  Mode 1 is the scale-invariant variant: breakdown unless
 * |pAp| > 1e-14 |rz| (alpha = rz/pAp < 1e14), an option added because the
 * absolute test fires on small-h meshes long before a 1e-10 relative
 * residual is reached.  Written as a negated comparison so that pAp = rz = 0
 * (a zero residual: the reference's absolute test exits there too) and a NaN
 * count as breakdown. */
static int breakdown(double pap, double rz, int mode)
{
    if (mode == 1) return !(fabs(pap) > 1.0e-14 * fabs(rz));
    return fabs(pap) < 1.0e-14;
}

int ffo_solve_sparse(int n, const int *col_ptr, const int *row_idx,
                     const double *val, const double *b, double *x,
                     int method, int precond, double tolerance, int tol_type,
                     int max_iterations, int breakdown_mode,
                     ffo_solver_stats *stats)
{
    stats->converged = 0;
    stats->iterations = 0;
    stats->final_residual = 0.0;
    double *d = malloc(sizeof(double) * n);
    double *r = malloc(sizeof(double) * n);
    double *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n);
    double *ap = malloc(sizeof(double) * n);
    /* Jacobi diagonal: linear scan of column i for row i (:1130-1144) */
    for (int i = 0; i < n; ++i) {
        double aii = 0.0;
        for (int k = col_ptr[i] - 1; k < col_ptr[i + 1] - 1; ++k)
            if (row_idx[k] == i + 1) {
                aii = val[k];
                break;
            }
        d[i] = (precond == 1 && fabs(aii) > 1.0e-14) ? 1.0 / aii : 1.0;
    }
    ffo_csc_matvec(n, n, col_ptr, row_idx, val, x, ap);
    for (int i = 0; i < n; ++i) r[i] = b[i] - ap[i];
    double rr = dot(n, r, r);
    double initial_norm = sqrt(rr), residual_norm = initial_norm;
    double tol = tol_type == 1 ? tolerance : tolerance * initial_norm;

    if (method == 0) {
        /* cg_solve_operator: p = r, rr_old = r.r */
        for (int i = 0; i < n; ++i) p[i] = r[i];
        double rr_old = rr;
        for (int iter = 1; iter <= max_iterations; ++iter) {
            ffo_csc_matvec(n, n, col_ptr, row_idx, val, p, ap);
            double pap = dot(n, p, ap);
            if (breakdown(pap, rr_old, breakdown_mode)) {
                stats->iterations = iter - 1;
                break;
            }
            double alpha = rr_old / pap;
            for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
            for (int i = 0; i < n; ++i) r[i] = r[i] - alpha * ap[i];
            double rr_new = dot(n, r, r);
            residual_norm = sqrt(rr_new);
            if (residual_norm <= tol) {
                stats->converged = 1;
                stats->iterations = iter;
                break;
            }
            double beta = rr_new / rr_old;
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
            rr_old = rr_new;
            stats->iterations = iter;
        }
    } else {
        /* initialize_pcg_state: z = M^-1 r, p = z, rz = r.z */
        for (int i = 0; i < n; ++i) z[i] = d[i] * r[i];
        for (int i = 0; i < n; ++i) p[i] = z[i];
        double rz_old = dot(n, r, z);
        for (int iter = 1; iter <= max_iterations; ++iter) {
            ffo_csc_matvec(n, n, col_ptr, row_idx, val, p, ap);
            double pap = dot(n, p, ap);
            if (breakdown(pap, rz_old, breakdown_mode)) break;
            double alpha = rz_old / pap;
            for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
            for (int i = 0; i < n; ++i) r[i] = r[i] - alpha * ap[i];
            residual_norm = sqrt(dot(n, r, r));
            if (residual_norm <= tol) {
                stats->converged = 1;
                stats->iterations = iter;
                break;
            }
            for (int i = 0; i < n; ++i) z[i] = d[i] * r[i];
            double rz_new = dot(n, r, z);
            double beta = rz_new / rz_old;
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            rz_old = rz_new;
            stats->iterations = iter;
        }
    }
    /* finalize_solver_status */
    if (!stats->converged && residual_norm <= tol) stats->converged = 1;
    stats->final_residual = residual_norm;
    free(d);
    free(r);
    free(z);
    free(p);
    free(ap);
    return 0;
}

/* The Jacobi-PCG / CG loop of ffo_solve_sparse on nthreads threads, for
 * timing the CPU baseline on all host cores.  The matrix must be symmetric:
 * the product reads the CSC as CSR (one row per thread, no write conflicts)
 * and the dot products are OpenMP reductions, so the sums differ from the
 * sequential loop's in the last bits (iterations within one or two). */
static double pdot(int T, int n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for num_threads(T) reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void pmatvec(int T, int n, const int *col_ptr, const int *row_idx, const double *val,
                    const double *x, double *y)
{
#pragma omp parallel for num_threads(T) schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = col_ptr[i] - 1; k < col_ptr[i + 1] - 1; ++k) s += val[k] * x[row_idx[k] - 1];
        y[i] = s;
    }
}

int ffo_par_solve_sparse(int nthreads, int n, const int *col_ptr, const int *row_idx,
                         const double *val, const double *b, double *x, int method, int precond,
                         double tolerance, int tol_type, int max_iterations, int breakdown_mode,
                         ffo_solver_stats *stats)
{
    const int T = nthreads < 1 ? 1 : nthreads;
    stats->converged = 0;
    stats->iterations = 0;
    stats->final_residual = 0.0;
    double *d = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n);
    double *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
    double *ap = malloc(sizeof(double) * n);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int i = 0; i < n; ++i) {
        double aii = 0.0;
        for (int k = col_ptr[i] - 1; k < col_ptr[i + 1] - 1; ++k)
            if (row_idx[k] == i + 1) {
                aii = val[k];
                break;
            }
        d[i] = (method == 1 && precond == 1 && fabs(aii) > 1.0e-14) ? 1.0 / aii : 1.0;
    }
    pmatvec(T, n, col_ptr, row_idx, val, x, ap);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - ap[i];
        z[i] = d[i] * r[i];
        p[i] = z[i];
    }
    double rr = pdot(T, n, r, r);
    double residual_norm = sqrt(rr);
    const double tol = tol_type == 1 ? tolerance : tolerance * residual_norm;
    double rz_old = pdot(T, n, r, z);
    for (int iter = 1; iter <= max_iterations; ++iter) {
        pmatvec(T, n, col_ptr, row_idx, val, p, ap);
        const double pap = pdot(T, n, p, ap);
        if (breakdown(pap, rz_old, breakdown_mode)) break;
        const double alpha = rz_old / pap;
        double rrn = 0.0, rzn = 0.0;
#pragma omp parallel for num_threads(T) reduction(+ : rrn, rzn) schedule(static)
        for (int i = 0; i < n; ++i) {
            x[i] = x[i] + alpha * p[i];
            r[i] = r[i] - alpha * ap[i];
            z[i] = d[i] * r[i];
            rrn += r[i] * r[i];
            rzn += r[i] * z[i];
        }
        residual_norm = sqrt(rrn);
        stats->iterations = iter;
        if (residual_norm <= tol) {
            stats->converged = 1;
            break;
        }
        const double beta = rzn / rz_old;
#pragma omp parallel for num_threads(T) schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz_old = rzn;
    }
    if (!stats->converged && residual_norm <= tol) stats->converged = 1;
    stats->final_residual = residual_norm;
    free(d), free(r), free(z), free(p), free(ap);
    return 0;
}
