/*
 * ffo_quadrature.c -- oracle restatement of the collapsed (Duffy) Gauss rules.
 * Test infrastructure only (see ffo.h).
 *
 * Gauss-Legendre nodes come from Newton's iteration on the Legendre
 * three-term recurrence (the textbook method, e.g. Numerical Recipes
 * "gauleg"), refined to machine precision.  The reference's fortnum
 * gauss_legendre_ab is not in this container; its node ORDER is not pinned
 * and only changes the summation order.  Nodes here are ascending.
 */
#include "ffo.h"

#include <math.h>

/* fortnum gauss_legendre_ab, called at
 * src/quadrature/tetra_duffy_quadrature.f90:27 */
int ffo_gauss_legendre_ab(int n, double a, double b, double *nodes,
                          double *weights)
{
    if (n < 1) return 1;
    const double pi = 3.14159265358979323846264338327950288;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        /* Tricomi initial guess for the i-th largest root on [-1, 1] */
        double x = cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1.0; }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        /* recompute derivative at the converged root */
        {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1.0; }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
        }
        double w = 2.0 / ((1.0 - x * x) * dp * dp);
        /* ascending order on [a, b]: root -x first, +x mirrored */
        int lo = i, hi = n - 1 - i;
        double half = 0.5 * (b - a), mid = 0.5 * (a + b);
        nodes[lo] = mid - half * x;
        nodes[hi] = mid + half * x;
        weights[lo] = half * w;
        weights[hi] = half * w;
        if (lo == hi) nodes[lo] = mid; /* exact centre node for odd n */
    }
    return 0;
}

/* tetra_duffy_quadrature (src/quadrature/tetra_duffy_quadrature.f90:12-51):
 * m = (p + 4)/2 points per direction (integer division, :25). */
int ffo_tetra_duffy_count(int polynomial_degree)
{
    if (polynomial_degree < 0) return -1;
    int m = (polynomial_degree + 4) / 2;
    return m * m * m;
}

int ffo_tetra_duffy(int polynomial_degree, double *x, double *y, double *z,
                    double *w)
{
    if (polynomial_degree < 0) return 1;
    int m = (polynomial_degree + 4) / 2;
    double nodes[64], weights[64];
    if (m > 64) return 1;
    ffo_gauss_legendre_ab(m, 0.0, 1.0, nodes, weights);
    int point = 0;
    for (int a = 0; a < m; ++a) {
        double c1 = 1.0 - nodes[a];
        for (int b = 0; b < m; ++b) {
            double c2 = 1.0 - nodes[b];
            for (int c = 0; c < m; ++c, ++point) {
                /* loop order first -> second -> third (:33-49) */
                x[point] = nodes[a];
                y[point] = c1 * nodes[b];
                z[point] = c1 * c2 * nodes[c];
                w[point] = weights[a] * weights[b] * weights[c] * (c1 * c1) *
                           c2;
            }
        }
    }
    return 0;
}

/* triangle_duffy_quadrature (src/quadrature/triangle_duffy_quadrature.f90:
 * 12-43): m = (p + 3)/2 points per direction. */
int ffo_triangle_duffy_count(int polynomial_degree)
{
    if (polynomial_degree < 0) return -1;
    int m = (polynomial_degree + 3) / 2;
    return m * m;
}

int ffo_triangle_duffy(int polynomial_degree, double *x, double *y, double *w)
{
    if (polynomial_degree < 0) return 1;
    int m = (polynomial_degree + 3) / 2;
    double nodes[64], weights[64];
    if (m > 64) return 1;
    ffo_gauss_legendre_ab(m, 0.0, 1.0, nodes, weights);
    int point = 0;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b, ++point) {
            x[point] = nodes[a];
            y[point] = (1.0 - nodes[a]) * nodes[b];
            w[point] = weights[a] * weights[b] * (1.0 - nodes[a]);
        }
    return 0;
}
