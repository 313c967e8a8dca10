/*
 * ffo_csc.c -- oracle CSC construction, matvec and Dirichlet reduction.
 * Test infrastructure only (see ffo.h).
 *
 * fortsparse @3c19c978 (csc_from_triplet, csc_matvec) is an absent external
 * dependency.  Its pinned behaviour is duplicate summation
 * (test/test_nedelec_sparse_assembly.f90:36-39: 18 triplets -> nnz 17;
 * test/test_capi_nedelec_assembly.f90:72-83).  Row order inside a column and
 * the explicit-zero policy are not pinned by any reference test; this oracle
 * fixes ascending rows, keeps structural zeros, and sums duplicates in
 * triplet order.
 */
#include "ffo.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* csc_from_triplet: two stable counting sorts (by row, then by column) give
 * (col, row) order with ties in triplet order; equal (row, col) are then
 * merged in triplet order. */
int ffo_csc_from_triplet(int nrow, int ncol, int64_t ntrip, const int *rows,
                         const int *cols, const double *vals, int *col_ptr,
                         int *row_idx, double *val, int64_t *nnz)
{
    if (nrow < 0 || ncol < 0 || ntrip < 0) return 1;
    for (int64_t t = 0; t < ntrip; ++t)
        if (rows[t] < 1 || rows[t] > nrow || cols[t] < 1 || cols[t] > ncol)
            return 1;
    size_t nmax = (size_t)(nrow > ncol ? nrow : ncol) + 1;
    int64_t *start = calloc(nmax + 1, sizeof(int64_t));
    int64_t *by_row = malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
    int64_t *order = malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
    for (int64_t t = 0; t < ntrip; ++t) start[rows[t]]++;
    for (int r = 0; r < nrow; ++r) start[r + 1] += start[r];
    for (int64_t t = 0; t < ntrip; ++t) by_row[start[rows[t] - 1]++] = t;
    memset(start, 0, sizeof(int64_t) * (nmax + 1));
    for (int64_t t = 0; t < ntrip; ++t) start[cols[t]]++;
    for (int c = 0; c < ncol; ++c) start[c + 1] += start[c];
    int64_t *fill = malloc(sizeof(int64_t) * ((size_t)ncol + 1));
    memcpy(fill, start, sizeof(int64_t) * ((size_t)ncol + 1));
    for (int64_t i = 0; i < ntrip; ++i) {
        int64_t t = by_row[i];
        order[fill[cols[t] - 1]++] = t;
    }
    free(by_row);
    int64_t out = 0;
    for (int c = 0; c < ncol; ++c) {
        col_ptr[c] = (int)(out + 1);
        int64_t lo = start[c], hi = start[c + 1];
        for (int64_t i = lo; i < hi;) {
            int r = rows[order[i]];
            double s = vals[order[i]];
            int64_t k = i + 1;
            while (k < hi && rows[order[k]] == r) {
                s = s + vals[order[k]];
                ++k;
            }
            row_idx[out] = r;
            val[out] = s;
            ++out;
            i = k;
        }
    }
    col_ptr[ncol] = (int)(out + 1);
    *nnz = out;
    free(start);
    free(order);
    free(fill);
    return 0;
}

/* csc_matvec (fortsparse), called through spmv at
 * src/solvers/fortfem_sparse_matrix.f90:46-58: column scatter y += a_ij x_j. */
int ffo_csc_matvec(int nrow, int ncol, const int *col_ptr, const int *row_idx,
                   const double *val, const double *x, double *y)
{
    for (int i = 0; i < nrow; ++i) y[i] = 0.0;
    for (int c = 0; c < ncol; ++c) {
        double xc = x[c];
        for (int k = col_ptr[c] - 1; k < col_ptr[c + 1] - 1; ++k)
            y[row_idx[k] - 1] += val[k] * xc;
    }
    return 0;
}

/* spmv_jvp (src/solvers/fortfem_sparse_matrix.f90:60-91): column loop,
 * y_dot(row) += a_k x_dot(col) + a_dot_k x(col), rows summed in column order. */
int ffo_spmv_jvp(int nrow, int ncol, const int *col_ptr, const int *row_idx,
                 const double *val, const double *x, const double *val_dot,
                 const double *x_dot, double *y_dot)
{
    for (int i = 0; i < nrow; ++i) y_dot[i] = 0.0;
    for (int c = 0; c < ncol; ++c)
        for (int k = col_ptr[c] - 1; k < col_ptr[c + 1] - 1; ++k) {
            int r = row_idx[k] - 1;
            y_dot[r] = y_dot[r] + val[k] * x_dot[c] + val_dot[k] * x[c];
        }
    return 0;
}

/* spmv_vjp (:93-126): val_bar(k) = y_bar(row_k) x(col_k),
 * x_bar(col) = sum over the column of a_k y_bar(row_k). */
int ffo_spmv_vjp(int nrow, int ncol, const int *col_ptr, const int *row_idx,
                 const double *val, const double *x, const double *y_bar,
                 double *val_bar, double *x_bar)
{
    (void)nrow;
    for (int c = 0; c < ncol; ++c) {
        double s = 0.0;
        for (int k = col_ptr[c] - 1; k < col_ptr[c + 1] - 1; ++k) {
            int r = row_idx[k] - 1;
            val_bar[k] = y_bar[r] * x[c];
            s = s + val[k] * y_bar[r];
        }
        x_bar[c] = s;
    }
    return 0;
}

/* sparse_from_dense (src/solvers/fortfem_sparse_matrix.f90:16-44): keep
 * |a| > 1e-14, column-major triplets -> CSC.  dense is n x n column-major. */
int ffo_sparse_from_dense(int n, const double *dense, int *col_ptr,
                          int *row_idx, double *val, int64_t *nnz)
{
    int64_t out = 0;
    for (int c = 0; c < n; ++c) {
        col_ptr[c] = (int)(out + 1);
        for (int r = 0; r < n; ++r) {
            double a = dense[(size_t)r + (size_t)n * c];
            if (fabs(a) <= 1.0e-14) continue;
            row_idx[out] = r + 1;
            val[out] = a;
            ++out;
        }
    }
    col_ptr[n] = (int)(out + 1);
    *nnz = out;
    return 0;
}

/* sparse_direct_solve_constrained_real, reduction stage
 * (src/solvers/fortfem_sparse_direct.f90:264-337) and
 * build_constrained_reduction (:782-837): free dofs ascending, reduced CSC
 * keeps the original row order, rhs lifted by -A_fc g_c in column order. */
int ffo_constrained_reduction(int n, const int *col_ptr, const int *row_idx,
                              const double *val, const int *constrained,
                              const double *cval, const double *rhs,
                              int *free_count, int *free_dofs,
                              int *red_col_ptr, int *red_row_idx,
                              double *red_val, int64_t *red_nnz,
                              double *red_rhs)
{
    int *free_index = calloc((size_t)n + 1, sizeof(int));
    int nf = 0;
    for (int c = 0; c < n; ++c) {
        if (constrained[c]) continue;
        free_dofs[nf] = c + 1;
        free_index[c] = ++nf;
    }
    *free_count = nf;
    int64_t out = 0;
    for (int f = 0; f < nf; ++f) {
        int c = free_dofs[f] - 1;
        red_col_ptr[f] = (int)(out + 1);
        for (int k = col_ptr[c] - 1; k < col_ptr[c + 1] - 1; ++k) {
            int r = row_idx[k] - 1;
            if (constrained[r]) continue;
            red_row_idx[out] = free_index[r];
            red_val[out] = val[k];
            ++out;
        }
    }
    red_col_ptr[nf] = (int)(out + 1);
    *red_nnz = out;
    if (rhs && red_rhs) {
        for (int f = 0; f < nf; ++f) red_rhs[f] = rhs[free_dofs[f] - 1];
        for (int c = 0; c < n; ++c) {
            if (!constrained[c]) continue;
            for (int k = col_ptr[c] - 1; k < col_ptr[c + 1] - 1; ++k) {
                int r = row_idx[k] - 1;
                if (constrained[r]) continue;
                red_rhs[free_index[r] - 1] =
                    red_rhs[free_index[r] - 1] - val[k] * cval[c];
            }
        }
    }
    free(free_index);
    return 0;
}

/* ffo_csc_from_triplet on nthreads threads, with the same result: triplets
 * are split into nthreads contiguous ranges whose per-column counts place
 * each range's entries after the earlier ranges' (a stable counting sort by
 * column); each column is then sorted by row with an insertion sort that
 * keeps equal rows in triplet order, and equal (row, col) are summed in
 * triplet order.  For timing the CPU baseline on all host cores. */
int ffo_par_csc_from_triplet(int nthreads, int nrow, int ncol, int64_t ntrip, const int *rows,
                             const int *cols, const double *vals, int *col_ptr, int *row_idx,
                             double *val, int64_t *nnz)
{
    if (nrow < 0 || ncol < 0 || ntrip < 0 || nthreads < 1) return 1;
    int bad = 0;
#pragma omp parallel for num_threads(nthreads) reduction(| : bad) schedule(static)
    for (int64_t t = 0; t < ntrip; ++t)
        if (rows[t] < 1 || rows[t] > nrow || cols[t] < 1 || cols[t] > ncol) bad |= 1;
    if (bad) return 1;
    const int T = nthreads;
    int64_t *hist = calloc((size_t)T * ((size_t)ncol + 1), sizeof(int64_t));
    int64_t *start = calloc((size_t)ncol + 1, sizeof(int64_t));
    int64_t *order = malloc(sizeof(int64_t) * (size_t)(ntrip ? ntrip : 1));
    int64_t *ucount = calloc((size_t)ncol + 1, sizeof(int64_t));
    if (!hist || !start || !order || !ucount) {
        free(hist), free(start), free(order), free(ucount);
        return 1;
    }
#pragma omp parallel num_threads(T)
    {
        const int th = omp_get_thread_num();
        const int64_t lo = ntrip * th / T, hi = ntrip * (th + 1) / T;
        int64_t *h = hist + (size_t)th * ((size_t)ncol + 1);
        for (int64_t t = lo; t < hi; ++t) h[cols[t] - 1]++;
#pragma omp barrier
#pragma omp single
        {
            int64_t s = 0;
            for (int c = 0; c < ncol; ++c) {
                start[c] = s;
                for (int k = 0; k < T; ++k) {
                    int64_t *hk = hist + (size_t)k * ((size_t)ncol + 1);
                    const int64_t cnt = hk[c];
                    hk[c] = s;
                    s += cnt;
                }
            }
            start[ncol] = s;
        }
        for (int64_t t = lo; t < hi; ++t) order[h[cols[t] - 1]++] = t;
#pragma omp barrier
#pragma omp for schedule(dynamic, 1024)
        for (int c = 0; c < ncol; ++c) {
            const int64_t a = start[c], b = start[c + 1];
            for (int64_t i = a + 1; i < b; ++i) { /* stable insertion sort by row */
                const int64_t key = order[i];
                const int r = rows[key];
                int64_t j = i - 1;
                while (j >= a && rows[order[j]] > r) {
                    order[j + 1] = order[j];
                    --j;
                }
                order[j + 1] = key;
            }
            int64_t u = 0;
            for (int64_t i = a; i < b; ++i)
                if (i == a || rows[order[i]] != rows[order[i - 1]]) ++u;
            ucount[c] = u;
        }
#pragma omp single
        {
            int64_t s = 0;
            for (int c = 0; c < ncol; ++c) {
                const int64_t u = ucount[c];
                ucount[c] = s;
                s += u;
            }
            ucount[ncol] = s;
        }
#pragma omp for schedule(dynamic, 1024)
        for (int c = 0; c < ncol; ++c) {
            int64_t out = ucount[c];
            col_ptr[c] = (int)(out + 1);
            const int64_t a = start[c], b = start[c + 1];
            for (int64_t i = a; i < b;) {
                const int r = rows[order[i]];
                double s = vals[order[i]];
                int64_t k = i + 1;
                while (k < b && rows[order[k]] == r) {
                    s = s + vals[order[k]];
                    ++k;
                }
                row_idx[out] = r;
                val[out] = s;
                ++out;
                i = k;
            }
        }
    }
    col_ptr[ncol] = (int)(ucount[ncol] + 1);
    *nnz = ucount[ncol];
    free(hist), free(start), free(order), free(ucount);
    return 0;
}
