/*
 * ffo_nedelec.c -- oracle restatement of the tetrahedral first-kind Nedelec
 * path: edge / face / cell numbering, the moment-dual basis, face
 * orientation transforms, curl-mass and tensor-weighted element matrices and
 * their CSC assembly.  Test infrastructure only (see ffo.h).
 *
 * Basis (src/elements/tetra_nedelec_arbitrary_order.f90): candidates are the
 * monomial spanning set of the reference's runtime path (:317-397, order <= 5
 * branch); coefficients C = M^-1 with M the moment matrix of the degrees of
 * freedom (:203-315): 6 edges x k shifted-Legendre tangential moments, 4
 * faces x 2 tangents x P_{k-2} monomial moments on the triangle Duffy rule
 * of degree 2k+4, and 3 x P_{k-3} monomial volume moments on the tetra Duffy
 * rule of degree 2k+4, solved after max-row / max-column scaling.  At
 * k >= 6 the face and volume moments use orthogonal Dubiner / Koornwinder
 * polynomials as the reference does (:642-669).  Their implementation lives in
 * the absent fortnum dependency; the reference's tests fix psi_0 = 1 and
 * orthogonality to constants (test_tetra_nedelec_sparse_assembly_slow.f90:
 * 127-135,156-165), and the standard collapsed-coordinate construction
 * (Sherwin-Karniadakis, unnormalised Jacobi polynomials) is used here; any
 * other scaling changes face / interior dofs by an invertible block map only.
 */
#include "ffo.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

int ffo_nedelec_dof_count(int k) { return k * (k + 2) * (k + 3) / 2; }

/* ------------------------------------------------------------ candidates */
/* The coefficient construction and basis evaluation run in IEEE quad precision
 * (__float128): the scaled monomial moment matrix is ill-conditioned enough
 * that a double LU loses ~12 digits at k = 4 (the exact coefficients are
 * rationals such as 72576) and 80-bit arithmetic still loses ~1e-7 at k = 6.
 * Results are rounded to double. */
typedef __float128 ld;

static ld ipow(ld x, int e)
{
    ld r = 1.0L;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

static void add_term_l(ld *v, ld *c, int comp, ld coef, int a, int b, int g, const ld p[3])
{
    /* add_candidate_term (:672-697): component comp (0-based) of coef x^a y^b z^g */
    ld x = p[0], y = p[1], z = p[2];
    ld m = coef * ipow(x, a) * ipow(y, b) * ipow(z, g);
    ld dx = a ? coef * a * ipow(x, a - 1) * ipow(y, b) * ipow(z, g) : 0.0L;
    ld dy = b ? coef * b * ipow(x, a) * ipow(y, b - 1) * ipow(z, g) : 0.0L;
    ld dz = g ? coef * g * ipow(x, a) * ipow(y, b) * ipow(z, g - 1) : 0.0L;
    v[comp] += m;
    if (comp == 0) {
        c[1] += dz;
        c[2] -= dy;
    } else if (comp == 1) {
        c[0] -= dz;
        c[2] += dx;
    } else {
        c[0] += dy;
        c[1] -= dx;
    }
}

/* candidate values / curls, vals[j*3 + c]: the monomial spanning set of
 * evaluate_runtime_candidates (:317-397) */
static void candidates_l(int k, const ld p[3], ld *vals, ld *curls)
{
    int n = ffo_nedelec_dof_count(k);
    memset(vals, 0, sizeof(ld) * 3 * n);
    memset(curls, 0, sizeof(ld) * 3 * n);
    int j = 0;
    for (int comp = 0; comp < 3; ++comp)
        for (int td = 0; td <= k - 1; ++td)
            for (int a = 0; a <= td; ++a)
                for (int b = 0; b <= td - a; ++b, ++j)
                    add_term_l(vals + 3 * j, curls + 3 * j, comp, 1.0L, a, b, td - a - b, p);
    int td = k - 1;
    for (int fam = 0; fam < 2; ++fam)
        for (int a = 0; a <= td; ++a)
            for (int b = 0; b <= td - a; ++b, ++j) {
                int g = td - a - b;
                if (fam == 0) {
                    add_term_l(vals + 3 * j, curls + 3 * j, 0, -1.0L, a, b + 1, g, p);
                    add_term_l(vals + 3 * j, curls + 3 * j, 1, 1.0L, a + 1, b, g, p);
                } else {
                    add_term_l(vals + 3 * j, curls + 3 * j, 0, -1.0L, a, b, g + 1, p);
                    add_term_l(vals + 3 * j, curls + 3 * j, 2, 1.0L, a + 1, b, g, p);
                }
            }
    for (int b = 0; b <= td; ++b, ++j) {
        int g = td - b;
        add_term_l(vals + 3 * j, curls + 3 * j, 1, -1.0L, 0, b, g + 1, p);
        add_term_l(vals + 3 * j, curls + 3 * j, 2, 1.0L, 0, b + 1, g, p);
    }
}

static ld shifted_legendre(int m, ld t)
{
    /* shifted_legendre (:742-763): P_m(2t - 1) by the three-term recurrence */
    ld x = 2.0L * t - 1.0L;
    if (m == 0) return 1.0L;
    ld prev = 1.0L, cur = x;
    for (int d = 2; d <= m; ++d) {
        ld next = ((2.0L * d - 1.0L) * x * cur - (d - 1.0L) * prev) / d;
        prev = cur;
        cur = next;
    }
    return cur;
}

/* Gauss-Legendre on [0, 1] in extended precision (Newton on the recurrence) */
static void gauss_legendre_l(int n, ld *x, ld *w)
{
    const ld pi = M_PIq;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        ld z = cosq(pi * (i + 0.75L) / (n + 0.5L)), pp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            ld p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                ld p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            pp = n * (z * p1 - p0) / (z * z - 1.0L);
            ld dz = p1 / pp;
            z -= dz;
            if (fabsq(dz) < 1e-21L) break;
        }
        ld p0 = 1.0L, p1 = z;
        for (int k = 2; k <= n; ++k) {
            ld p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pp = n * (z * p1 - p0) / (z * z - 1.0L);
        ld wt = 2.0L / ((1.0L - z * z) * pp * pp);
        int lo = i, hi = n - 1 - i;
        x[lo] = 0.5L - 0.5L * z;
        x[hi] = 0.5L + 0.5L * z;
        w[lo] = w[hi] = 0.5L * wt;
        if (lo == hi) x[lo] = 0.5L;
    }
}

static int tri_duffy_l(int p, ld *x, ld *y, ld *w)
{
    int m = (p + 3) / 2;
    ld s[64], sw[64];
    gauss_legendre_l(m, s, sw);
    int q = 0;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b, ++q) {
            x[q] = s[a];
            y[q] = (1.0L - s[a]) * s[b];
            w[q] = sw[a] * sw[b] * (1.0L - s[a]);
        }
    return q;
}

static int tet_duffy_l(int p, ld *x, ld *y, ld *z, ld *w)
{
    int m = (p + 4) / 2;
    ld s[64], sw[64];
    gauss_legendre_l(m, s, sw);
    int q = 0;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            for (int c = 0; c < m; ++c, ++q) {
                ld c1 = 1.0L - s[a], c2 = 1.0L - s[b];
                x[q] = s[a];
                y[q] = c1 * s[b];
                z[q] = c1 * c2 * s[c];
                w[q] = sw[a] * sw[b] * sw[c] * (c1 * c1) * c2;
            }
    return q;
}

/* Jacobi P_n^{(a,0)}(x) by the three-term recurrence */
static ld jacobi(int n, int a, ld x)
{
    if (n == 0) return 1.0Q;
    ld p0 = 1.0Q, p1 = 0.5Q * a + 0.5Q * (a + 2) * x;
    for (int m = 2; m <= n; ++m) {
        ld c = 2.0Q * m + a;
        ld a1 = 2.0Q * m * (m + a) * (c - 2.0Q);
        ld a2 = (c - 1.0Q) * (c * (c - 2.0Q) * x + (ld)a * a);
        ld a3 = 2.0Q * (m + a - 1.0Q) * (m - 1.0Q) * c;
        ld p2 = (a2 * p1 - a3 * p0) / a1;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

/* Dubiner polynomial on the unit triangle: P_p(2x/(1-y) - 1) (1-y)^p
 * P_q^{(2p+1,0)}(2y - 1) */
static ld dubiner(int p, int q, ld x, ld y)
{
    ld s = 1.0Q - y;
    ld e = s > 0 ? 2.0Q * x / s - 1.0Q : -1.0Q;
    return jacobi(p, 0, e) * ipow(s, p) * jacobi(q, 2 * p + 1, 2.0Q * y - 1.0Q);
}

/* Koornwinder polynomial on the unit tetrahedron:
 * P_p(2x/(1-y-z) - 1) (1-y-z)^p P_q^{(2p+1,0)}(2y/(1-z) - 1) (1-z)^q
 * P_r^{(2p+2q+2,0)}(2z - 1) */
static ld koornwinder(int p, int q, int r, ld x, ld y, ld z)
{
    ld s1 = 1.0Q - y - z, s2 = 1.0Q - z;
    ld e1 = s1 > 0 ? 2.0Q * x / s1 - 1.0Q : -1.0Q;
    ld e2 = s2 > 0 ? 2.0Q * y / s2 - 1.0Q : -1.0Q;
    return jacobi(p, 0, e1) * ipow(s1, p) * jacobi(q, 2 * p + 1, e2) * ipow(s2, q) *
           jacobi(r, 2 * p + 2 * q + 2, 2.0Q * z - 1.0Q);
}

/* face_moment_value / volume_moment_value (:642-669) */
static ld face_poly(int k, int a, int b, ld x, ld y)
{
    return k <= 5 ? ipow(x, a) * ipow(y, b) : dubiner(a, b, x, y);
}

static ld volume_poly(int k, int a, int b, int g, ld x, ld y, ld z)
{
    return k <= 5 ? ipow(x, a) * ipow(y, b) * ipow(z, g) : koornwinder(a, b, g, x, y, z);
}

static const ld REF_V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const int EDGE_V[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int FACE_V[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

/* LU solve with partial pivoting: a (n x n, row-major) x = b (n x m) in place */
static int lu_solve(int n, ld *a, int m, ld *b)
{
    for (int c = 0; c < n; ++c) {
        int p = c;
        ld best = fabsq(a[(size_t)c * n + c]);
        for (int r = c + 1; r < n; ++r)
            if (fabsq(a[(size_t)r * n + c]) > best) {
                best = fabsq(a[(size_t)r * n + c]);
                p = r;
            }
        if (best == 0.0L) return 1;
        if (p != c) {
            for (int k = 0; k < n; ++k) {
                ld t = a[(size_t)c * n + k];
                a[(size_t)c * n + k] = a[(size_t)p * n + k];
                a[(size_t)p * n + k] = t;
            }
            for (int k = 0; k < m; ++k) {
                ld t = b[(size_t)c * m + k];
                b[(size_t)c * m + k] = b[(size_t)p * m + k];
                b[(size_t)p * m + k] = t;
            }
        }
        for (int r = c + 1; r < n; ++r) {
            ld f = a[(size_t)r * n + c] / a[(size_t)c * n + c];
            if (f == 0.0L) continue;
            for (int k = c; k < n; ++k) a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
            for (int k = 0; k < m; ++k) b[(size_t)r * m + k] -= f * b[(size_t)c * m + k];
        }
    }
    for (int c = n - 1; c >= 0; --c)
        for (int k = 0; k < m; ++k) {
            ld s = b[(size_t)c * m + k];
            for (int j = c + 1; j < n; ++j) s -= a[(size_t)c * n + j] * b[(size_t)j * m + k];
            b[(size_t)c * m + k] = s / a[(size_t)c * n + c];
        }
    return 0;
}

/* Moment-dual coefficients C (n x n, row-major C[i*n + j] = coefficient of
 * candidate i in basis function j), build_runtime_coefficients (:203-315). */
static int coefficients_l(int k, ld *coeff)
{
    if (k < 1 || k > 8) return 1;
    int n = ffo_nedelec_dof_count(k);
    ld *M = calloc((size_t)n * n, sizeof(ld));
    ld *cv = malloc(sizeof(ld) * 3 * n), *cc = malloc(sizeof(ld) * 3 * n);
    int row = 0;
    /* edge moments: k + 2 Gauss points on [0, 1] */
    ld en[64], ew[64];
    gauss_legendre_l(k + 2, en, ew);
    for (int e = 0; e < 6; ++e) {
        ld tan[3], p[3];
        for (int c = 0; c < 3; ++c) tan[c] = REF_V[EDGE_V[e][1]][c] - REF_V[EDGE_V[e][0]][c];
        for (int m = 0; m < k; ++m, ++row)
            for (int q = 0; q < k + 2; ++q) {
                for (int c = 0; c < 3; ++c) p[c] = REF_V[EDGE_V[e][0]][c] + en[q] * tan[c];
                candidates_l(k, p, cv, cc);
                ld f = ew[q] * shifted_legendre(m, en[q]);
                for (int j = 0; j < n; ++j)
                    M[(size_t)row * n + j] +=
                        f * (tan[0] * cv[3 * j] + tan[1] * cv[3 * j + 1] + tan[2] * cv[3 * j + 2]);
            }
    }
    /* face moments: triangle Duffy rule of degree 2k + 4 */
    int mt = (2 * k + 4 + 3) / 2;
    ld *tx = malloc(sizeof(ld) * mt * mt * 3), *ty = tx + mt * mt, *tw = ty + mt * mt;
    int ntq = tri_duffy_l(2 * k + 4, tx, ty, tw);
    for (int f = 0; f < 4; ++f) {
        ld t[2][3], p[3];
        for (int c = 0; c < 3; ++c) {
            t[0][c] = REF_V[FACE_V[f][1]][c] - REF_V[FACE_V[f][0]][c];
            t[1][c] = REF_V[FACE_V[f][2]][c] - REF_V[FACE_V[f][0]][c];
        }
        for (int comp = 0; comp < 2; ++comp)
            for (int td = 0; td <= k - 2; ++td)
                for (int a = 0; a <= td; ++a, ++row) {
                    int b = td - a;
                    for (int q = 0; q < ntq; ++q) {
                        for (int c = 0; c < 3; ++c)
                            p[c] = REF_V[FACE_V[f][0]][c] + tx[q] * t[0][c] + ty[q] * t[1][c];
                        candidates_l(k, p, cv, cc);
                        ld w = tw[q] * face_poly(k, a, b, tx[q], ty[q]);
                        for (int j = 0; j < n; ++j)
                            M[(size_t)row * n + j] +=
                                w * (t[comp][0] * cv[3 * j] + t[comp][1] * cv[3 * j + 1] +
                                     t[comp][2] * cv[3 * j + 2]);
                    }
                }
    }
    free(tx);
    /* volume moments: tetra Duffy rule of degree 2k + 4 */
    int mv = (2 * k + 4 + 4) / 2;
    size_t nvq0 = (size_t)mv * mv * mv;
    ld *vx = malloc(sizeof(ld) * nvq0 * 4), *vy = vx + nvq0, *vz = vy + nvq0, *vw = vz + nvq0;
    int nvq = tet_duffy_l(2 * k + 4, vx, vy, vz, vw);
    for (int comp = 0; comp < 3; ++comp)
        for (int td = 0; td <= k - 3; ++td)
            for (int a = 0; a <= td; ++a)
                for (int b = 0; b <= td - a; ++b, ++row) {
                    int g = td - a - b;
                    for (int q = 0; q < nvq; ++q) {
                        ld p[3] = {vx[q], vy[q], vz[q]};
                        candidates_l(k, p, cv, cc);
                        ld w = vw[q] * volume_poly(k, a, b, g, p[0], p[1], p[2]);
                        for (int j = 0; j < n; ++j) M[(size_t)row * n + j] += w * cv[3 * j + comp];
                    }
                }
    free(vx);
    free(cv);
    free(cc);
    if (row != n) {
        free(M);
        return 1;
    }
    /* scaling: rows by 1/max|row|, then columns by 1/max|col| (:287-301) */
    ld *rs = malloc(sizeof(ld) * n), *cs = malloc(sizeof(ld) * n);
    for (int i = 0; i < n; ++i) {
        ld mx = 0;
        for (int j = 0; j < n; ++j) mx = fmaxq(mx, fabsq(M[(size_t)i * n + j]));
        rs[i] = 1.0L / mx;
        for (int j = 0; j < n; ++j) M[(size_t)i * n + j] *= rs[i];
    }
    for (int j = 0; j < n; ++j) {
        ld mx = 0;
        for (int i = 0; i < n; ++i) mx = fmaxq(mx, fabsq(M[(size_t)i * n + j]));
        cs[j] = 1.0L / mx;
        for (int i = 0; i < n; ++i) M[(size_t)i * n + j] *= cs[j];
    }
    /* (R M S) X = R  ->  coefficients = S X = M^-1 */
    ld *X = calloc((size_t)n * n, sizeof(ld));
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = rs[i];
    int st = lu_solve(n, M, n, X);
    if (!st)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) coeff[(size_t)i * n + j] = cs[i] * X[(size_t)i * n + j];
    free(X);
    free(rs);
    free(cs);
    free(M);
    return st;
}

/* extended-precision coefficients, built once per order */
static const ld *cached_coefficients(int k)
{
    static ld *cache[9];
    if (k < 1 || k > 8) return NULL;
    if (!cache[k]) {
        int n = ffo_nedelec_dof_count(k);
        ld *c = malloc(sizeof(ld) * n * n);
        if (coefficients_l(k, c)) {
            free(c);
            return NULL;
        }
        cache[k] = c;
    }
    return cache[k];
}

int ffo_nedelec_coefficients(int k, double *coeff)
{
    const ld *c = cached_coefficients(k);
    if (!c) return 1;
    int n = ffo_nedelec_dof_count(k);
    for (int i = 0; i < n * n; ++i) coeff[i] = (double)c[i];
    return 0;
}

/* reference basis values/curls at a point: vals[j*3 + c] (candidates x
 * coefficients, evaluated in extended precision and rounded; coeff is
 * unused and kept for the call signature) */
int ffo_nedelec_eval(int k, const double *coeff, const double p[3], double *vals,
                     double *curls)
{
    (void)coeff;
    int n = ffo_nedelec_dof_count(k);
    const ld *C = cached_coefficients(k);
    if (!C) return 1;
    ld pl[3] = {p[0], p[1], p[2]};
    ld *cv = malloc(sizeof(ld) * 3 * n), *cc = malloc(sizeof(ld) * 3 * n);
    candidates_l(k, pl, cv, cc);
    for (int j = 0; j < n; ++j)
        for (int c = 0; c < 3; ++c) {
            ld sv = 0, sc = 0;
            for (int i = 0; i < n; ++i) {
                ld a = C[(size_t)i * n + j];
                sv += cv[3 * i + c] * a;
                sc += cc[3 * i + c] * a;
            }
            vals[3 * j + c] = (double)sv;
            curls[3 * j + c] = (double)sc;
        }
    free(cv);
    free(cc);
    return 0;
}

/* evaluate_tetra_nedelec_first_order (src/elements/tetra_nedelec_first_order.f90:
 * 11-51): w_ab = l_a grad l_b - l_b grad l_a, curl = 2 grad l_a x grad l_b */
int ffo_whitney(const double p[3], double *vals, double *curls)
{
    static const double G[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double lam[4] = {1.0 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
    for (int e = 0; e < 6; ++e) {
        int a = EDGE_V[e][0], b = EDGE_V[e][1];
        for (int c = 0; c < 3; ++c) vals[3 * e + c] = lam[a] * G[b][c] - lam[b] * G[a][c];
        curls[3 * e + 0] = 2.0 * (G[a][1] * G[b][2] - G[a][2] * G[b][1]);
        curls[3 * e + 1] = 2.0 * (G[a][2] * G[b][0] - G[a][0] * G[b][2]);
        curls[3 * e + 2] = 2.0 * (G[a][0] * G[b][1] - G[a][1] * G[b][0]);
    }
    return 0;
}

/* --------------------------------------------------------- face transform */
/* build_moment_matrix (tetra_face_moment_transforms.f90:335-381), size
 * k(k-1), row-major */
static void face_moment_matrix(int k, const int perm[3], ld *M)
{
    static const int V2[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    int off[2] = {V2[perm[0] - 1][0], V2[perm[0] - 1][1]};
    int aff[2][2]; /* aff[i][j] = i-th coordinate of edge j */
    for (int i = 0; i < 2; ++i) {
        aff[i][0] = V2[perm[1] - 1][i] - off[i];
        aff[i][1] = V2[perm[2] - 1][i] - off[i];
    }
    int s = k * (k - 1);
    memset(M, 0, sizeof(ld) * s * s);
    int mt = (2 * k + 2 + 3) / 2;
    ld *u = malloc(sizeof(ld) * mt * mt * 3), *v = u + mt * mt, *w = v + mt * mt;
    int ntq = tri_duffy_l(2 * k + 2, u, v, w);
    int col = 0;
    for (int bc = 0; bc < 2; ++bc)
        for (int bd = 0; bd <= k - 2; ++bd)
            for (int bx = 0; bx <= bd; ++bx, ++col) {
                int by = bd - bx, row = 0;
                for (int mc = 0; mc < 2; ++mc)
                    for (int md = 0; md <= k - 2; ++md)
                        for (int mx = 0; mx <= md; ++mx, ++row) {
                            int my = md - mx;
                            ld sum = 0;
                            for (int q = 0; q < ntq; ++q) {
                                ld m0 = off[0] + aff[0][0] * u[q] + aff[0][1] * v[q];
                                ld m1 = off[1] + aff[1][0] * u[q] + aff[1][1] * v[q];
                                sum += w[q] * (ld)aff[bc][mc] * face_poly(k, bx, by, m0, m1) *
                                       face_poly(k, mx, my, u[q], v[q]);
                            }
                            M[(size_t)row * s + col] = sum;
                        }
            }
    free(u);
}

/* build_tetra_face_basis_to_local_matrix, runtime route (:296-333):
 * transform = local * canonical^-1 (row-major, k(k-1) square) */
int ffo_face_basis_to_local(int k, const int perm[3], double *T)
{
    int s = k * (k - 1);
    if (s == 0) return 0;
    ld *C = malloc(sizeof(ld) * s * s), *L = malloc(sizeof(ld) * s * s);
    static const int ident[3] = {1, 2, 3};
    face_moment_matrix(k, ident, C);
    face_moment_matrix(k, perm, L);
    /* solve C^T X = L^T, transform = X^T */
    ld *Ct = malloc(sizeof(ld) * s * s), *X = malloc(sizeof(ld) * s * s);
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) {
            Ct[(size_t)i * s + j] = C[(size_t)j * s + i];
            X[(size_t)i * s + j] = L[(size_t)j * s + i];
        }
    int st = lu_solve(s, Ct, s, X);
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) T[(size_t)i * s + j] = (double)X[(size_t)j * s + i];
    free(C);
    free(L);
    free(Ct);
    free(X);
    return st;
}

/* ------------------------------------------------------------ dof maps */
typedef struct {
    int v[3];
} key3;

static uint64_t hash3(const key3 *k)
{
    uint64_t h = 0x51ed270b27f1ull;
    for (int i = 0; i < 3; ++i) {
        h ^= (uint64_t)(uint32_t)k->v[i] * 0x9e3779b97f4a7c15ull;
        h = (h << 27) | (h >> 37);
        h *= 0xbf58476d1ce4e5b9ull;
    }
    return h ^ (h >> 31);
}

typedef struct {
    key3 *keys;
    int *ids;
    size_t cap;
    int count;
} table3;

static void t3_init(table3 *t, size_t n)
{
    t->cap = 1;
    while (t->cap < 2 * n + 2) t->cap <<= 1;
    t->keys = malloc(sizeof(key3) * t->cap);
    t->ids = malloc(sizeof(int) * t->cap);
    for (size_t i = 0; i < t->cap; ++i) t->ids[i] = -1;
    t->count = 0;
}

static int t3_get(table3 *t, key3 k) /* first-appearance id (0-based) */
{
    size_t s = hash3(&k) & (t->cap - 1);
    while (t->ids[s] >= 0 && memcmp(&t->keys[s], &k, sizeof k) != 0) s = (s + 1) & (t->cap - 1);
    if (t->ids[s] < 0) {
        t->keys[s] = k;
        t->ids[s] = t->count++;
    }
    return t->ids[s];
}

static void t3_free(table3 *t)
{
    free(t->keys);
    free(t->ids);
}

/* build_tetra_edge_dof_map (src/elements/tetra_edge_dof_map.f90:10-77):
 * local edges 12,13,14,23,24,34; edge = (min, max) numbered by first
 * appearance; orientation +1 iff first local vertex id < second.
 * edges(2, E), gd(6, T), orient(6, T), all 1-based. */
int ffo_edge_dof_map(int ntet, const int *tets, int *edges, int *gd, int *orient,
                     int *nedges)
{
    *nedges = 0;
    for (int e = 0; e < ntet; ++e)
        for (int a = 0; a < 4; ++a) {
            if (tets[4 * e + a] < 1) return 1;
            for (int b = a + 1; b < 4; ++b)
                if (tets[4 * e + a] == tets[4 * e + b]) return 1;
        }
    table3 t;
    t3_init(&t, (size_t)6 * ntet);
    for (int e = 0; e < ntet; ++e)
        for (int l = 0; l < 6; ++l) {
            int va = tets[4 * e + EDGE_V[l][0]], vb = tets[4 * e + EDGE_V[l][1]];
            key3 k = {{va < vb ? va : vb, va < vb ? vb : va, 0}};
            int before = t.count;
            int id = t3_get(&t, k);
            if (t.count > before) {
                edges[2 * id] = k.v[0];
                edges[2 * id + 1] = k.v[1];
            }
            gd[6 * e + l] = id + 1;
            orient[6 * e + l] = va < vb ? 1 : -1;
        }
    *nedges = t.count;
    t3_free(&t);
    return 0;
}

/* literal linear-search restatement (:42-61) for tiny meshes */
int ffo_edge_dof_map_search(int ntet, const int *tets, int *edges, int *gd, int *orient,
                            int *nedges)
{
    int count = 0;
    for (int e = 0; e < ntet; ++e)
        for (int l = 0; l < 6; ++l) {
            int va = tets[4 * e + EDGE_V[l][0]], vb = tets[4 * e + EDGE_V[l][1]];
            int lo = va < vb ? va : vb, hi = va < vb ? vb : va, found = -1;
            for (int x = 0; x < count; ++x)
                if (edges[2 * x] == lo && edges[2 * x + 1] == hi) {
                    found = x;
                    break;
                }
            if (found < 0) {
                edges[2 * count] = lo;
                edges[2 * count + 1] = hi;
                found = count++;
            }
            gd[6 * e + l] = found + 1;
            orient[6 * e + l] = va < vb ? 1 : -1;
        }
    *nedges = count;
    return 0;
}

/* build_tetra_nedelec_dof_map (src/elements/tetra_nedelec_global_dof_map.f90:
 * 15-105): edge dofs (e-1)k + m, then face dofs by first appearance of the
 * sorted face triple, then k(k-1)(k-2)/2 interior dofs per cell in cell
 * order.  face_perm(3, 4, T) = position of each local face vertex in the
 * sorted triple (1..3).  gd(n, T) 1-based. */
int ffo_nedelec_dof_map(int k, int ntet, const int *tets, int *gd, int *edge_orient,
                        int *face_perm, int *ndofs, int *nedges_out, int *nfaces_out)
{
    if (k < 1) return 1;
    int n = ffo_nedelec_dof_count(k);
    int *edges = malloc(sizeof(int) * 12 * (size_t)ntet);
    int *egd = malloc(sizeof(int) * 6 * (size_t)ntet);
    int ne = 0;
    if (ffo_edge_dof_map(ntet, tets, edges, egd, edge_orient, &ne)) {
        free(edges);
        free(egd);
        return 1;
    }
    int fdc = k * (k - 1), cdc = k * (k - 1) * (k - 2) / 2;
    int edc = k * ne;
    for (int e = 0; e < ntet; ++e)
        for (int l = 0; l < 6; ++l)
            for (int m = 0; m < k; ++m) gd[(size_t)n * e + l * k + m] = (egd[6 * e + l] - 1) * k + m + 1;
    table3 t;
    t3_init(&t, (size_t)4 * ntet);
    for (int e = 0; e < ntet; ++e)
        for (int f = 0; f < 4; ++f) {
            int lv[3], cv[3];
            for (int i = 0; i < 3; ++i) lv[i] = cv[i] = tets[4 * e + FACE_V[f][i]];
            for (int a = 0; a < 2; ++a)
                for (int b = a + 1; b < 3; ++b)
                    if (cv[b] < cv[a]) {
                        int x = cv[a];
                        cv[a] = cv[b];
                        cv[b] = x;
                    }
            key3 key = {{cv[0], cv[1], cv[2]}};
            int face = t3_get(&t, key);
            for (int i = 0; i < 3; ++i)
                for (int pos = 0; pos < 3; ++pos)
                    if (cv[pos] == lv[i]) face_perm[12 * e + 3 * f + i] = pos + 1;
            int start = 6 * k + f * fdc;
            for (int d = 0; d < fdc; ++d) gd[(size_t)n * e + start + d] = edc + face * fdc + d + 1;
        }
    int nf = t.count;
    t3_free(&t);
    int start = 6 * k + 4 * fdc;
    for (int e = 0; e < ntet; ++e)
        for (int d = 0; d < cdc; ++d)
            gd[(size_t)n * e + start + d] = edc + nf * fdc + e * cdc + d + 1;
    *ndofs = edc + nf * fdc + ntet * cdc;
    if (nedges_out) *nedges_out = ne;
    if (nfaces_out) *nfaces_out = nf;
    free(edges);
    free(egd);
    return 0;
}

/* build_tetra_nedelec_basis_transform (:107-155), T row-major n x n */
int ffo_nedelec_basis_transform(int k, const int edge_orient[6], const int face_perm[12],
                                double *T)
{
    int n = ffo_nedelec_dof_count(k);
    memset(T, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) T[(size_t)i * n + i] = 1.0;
    for (int e = 0; e < 6; ++e) {
        if (edge_orient[e] == 1) continue;
        for (int d = 0; d < k; d += 2) {
            int s = e * k + d;
            T[(size_t)s * n + s] = -1.0;
        }
    }
    int fdc = k * (k - 1);
    if (fdc == 0) return 0;
    double *F = malloc(sizeof(double) * fdc * fdc);
    for (int f = 0; f < 4; ++f) {
        int start = 6 * k + f * fdc;
        if (ffo_face_basis_to_local(k, face_perm + 3 * f, F)) {
            free(F);
            return 1;
        }
        for (int i = 0; i < fdc; ++i)
            for (int j = 0; j < fdc; ++j) T[(size_t)(start + i) * n + start + j] = F[(size_t)i * fdc + j];
    }
    free(F);
    return 0;
}

/* ------------------------------------------------------------- elements */
static int geometry(const double v[12], double jac[9], double inv[9], double *det)
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) jac[r + 3 * c] = v[3 * (c + 1) + r] - v[r];
    *det = ffo_det3(jac);
    double mx = 0;
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(jac[i]));
    double s = fmax(1.0, mx);
    if (*det <= 64.0 * DBL_EPSILON * s * s * s) return 1;
    return ffo_inv3(jac, inv);
}

/* element rule: k = 1 uses the 4-point degree-2 rule
 * (assembly_tetra_nedelec_3d.f90:2364-2366,2380-2383), k >= 2 the Duffy rule */
static int element_rule(int k, int qdeg, double **px, double **py, double **pz, double **pw)
{
    if (k == 1) {
        double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
        double *x = malloc(sizeof(double) * 16);
        double pts[4][3] = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
        for (int q = 0; q < 4; ++q) {
            x[q] = pts[q][0];
            x[4 + q] = pts[q][1];
            x[8 + q] = pts[q][2];
            x[12 + q] = 1.0 / 24.0;
        }
        *px = x, *py = x + 4, *pz = x + 8, *pw = x + 12;
        return 4;
    }
    int nq = ffo_tetra_duffy_count(qdeg);
    double *x = malloc(sizeof(double) * 4 * nq);
    ffo_tetra_duffy(qdeg, x, x + nq, x + 2 * nq, x + 3 * nq);
    *px = x, *py = x + nq, *pz = x + 2 * nq, *pw = x + 3 * nq;
    return nq;
}

/* Element matrix (column-major m[r + n c]):
 * assemble_tetra_nedelec_element (:2357-2404) for k = 1,
 * assemble_tetra_nedelec_curl_mass_element (:584-643) for k >= 2 (tensor
 * NULL, alpha = curl coefficient), and the weighted elements (:2406-2514,
 * tensor = K(3x3 column-major, constant on the cell), alpha ignored).
 * Covariant map (src/elements/tetra_piola_maps.f90:18-57): v = J^-T v^,
 * curl = J c^ / det. */
static int element_impl(int k, const double *coeff, const double v[12], int qdeg, double alpha,
                        const double *tensor, int tensor_stride, double beta, double *m);

int ffo_nedelec_element(int k, const double *coeff, const double v[12], int qdeg, double alpha,
                        const double *tensor, double beta, double *m)
{
    return element_impl(k, coeff, v, qdeg, alpha, tensor, 0, beta, m);
}

/* The weighted element with the reference's coefficient(x_q) evaluated at
 * every point of the rule (:2406-2514): point_tensors holds 9 values
 * (column-major) per point of element_rule(k, qdeg), in point order. */
int ffo_nedelec_element_points(int k, const double *coeff, const double v[12], int qdeg,
                               const double *point_tensors, double beta, double *m)
{
    return element_impl(k, coeff, v, qdeg, 0.0, point_tensors, 9, beta, m);
}

/* the points of element_rule(k, qdeg): x, y, z, w per point (npts <= cap) */
int ffo_nedelec_rule_points(int k, int qdeg, int cap, double *out, int *npts)
{
    double *qx, *qy, *qz, *qw;
    int nq = element_rule(k, qdeg, &qx, &qy, &qz, &qw);
    *npts = nq;
    if (nq > cap) {
        free(qx);
        return 1;
    }
    for (int q = 0; q < nq; ++q) {
        out[4 * q] = qx[q];
        out[4 * q + 1] = qy[q];
        out[4 * q + 2] = qz[q];
        out[4 * q + 3] = qw[q];
    }
    free(qx);
    return 0;
}

static int element_impl(int k, const double *coeff, const double v[12], int qdeg, double alpha,
                        const double *tensor, int tensor_stride, double beta, double *m)
{
    int n = ffo_nedelec_dof_count(k);
    double jac[9], inv[9], det;
    if (geometry(v, jac, inv, &det)) return 1;
    double *qx, *qy, *qz, *qw;
    int nq = element_rule(k, qdeg, &qx, &qy, &qz, &qw);
    double *rv = malloc(sizeof(double) * 3 * n), *rc = malloc(sizeof(double) * 3 * n);
    double *pv = malloc(sizeof(double) * 3 * n), *pc = malloc(sizeof(double) * 3 * n);
    double *kc = malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n * n; ++i) m[i] = 0.0;
    for (int q = 0; q < nq; ++q) {
        double p[3] = {qx[q], qy[q], qz[q]};
        if (k == 1) ffo_whitney(p, rv, rc);
        else ffo_nedelec_eval(k, coeff, p, rv, rc);
        for (int j = 0; j < n; ++j)
            for (int r = 0; r < 3; ++r) {
                pv[3 * j + r] = inv[0 + 3 * r] * rv[3 * j] + inv[1 + 3 * r] * rv[3 * j + 1] +
                                inv[2 + 3 * r] * rv[3 * j + 2];
                pc[3 * j + r] = (jac[r] * rc[3 * j] + jac[r + 3] * rc[3 * j + 1] +
                                 jac[r + 6] * rc[3 * j + 2]) / det;
            }
        if (tensor) {
            const double *K = tensor + (size_t)tensor_stride * q;
            for (int j = 0; j < n; ++j)
                for (int r = 0; r < 3; ++r)
                    kc[3 * j + r] = K[r] * pc[3 * j] + K[r + 3] * pc[3 * j + 1] +
                                    K[r + 6] * pc[3 * j + 2];
        }
        double pw = det * qw[q];
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double cc, vv = pv[3 * r] * pv[3 * c] + pv[3 * r + 1] * pv[3 * c + 1] +
                                pv[3 * r + 2] * pv[3 * c + 2];
                if (tensor)
                    cc = pc[3 * r] * kc[3 * c] + pc[3 * r + 1] * kc[3 * c + 1] +
                         pc[3 * r + 2] * kc[3 * c + 2];
                else
                    cc = alpha * (pc[3 * r] * pc[3 * c] + pc[3 * r + 1] * pc[3 * c + 1] +
                                  pc[3 * r + 2] * pc[3 * c + 2]);
                m[r + (size_t)n * c] += pw * (cc + beta * vv);
            }
    }
    free(qx);
    free(rv);
    free(rc);
    free(pv);
    free(pc);
    free(kc);
    return 0;
}

/* CSC triplets of assemble_tetra_nedelec_curl_mass_csc (:2042-2112,
 * 2289-2355) and assemble_tetra_nedelec_weighted_csc (:1870-2000):
 * per cell the oriented matrix T^T E T (k >= 2) or the sign product
 * o_r o_c E (k = 1), column-outer / row-inner.  cell_tensor (9 per cell,
 * optional) selects the weighted variant with quadrature 2k+2; otherwise the
 * curl-mass variant with quadrature 2k.  Returns the triplet count and the
 * global dimension. */
static int triplets_impl(int k, int nvert, const double *vertices, int ntet, const int *tets,
                         double alpha, double beta, const double *cell_tensor, int point_nq,
                         int *rows, int *cols, double *vals, int *ndofs);

int ffo_nedelec_triplets(int k, int nvert, const double *vertices, int ntet, const int *tets,
                         double alpha, double beta, const double *cell_tensor, int *rows,
                         int *cols, double *vals, int *ndofs)
{
    return triplets_impl(k, nvert, vertices, ntet, tets, alpha, beta, cell_tensor, 0, rows, cols,
                         vals, ndofs);
}

/* assemble_tetra_nedelec_weighted_csc (:1870-2000) with the coefficient
 * evaluated at every point of the weighted rule (2k + 2; 4 points at k = 1):
 * point_tensors holds 9 values (column-major) per point, cell-major
 * (cell e, point q at 9 (e nq + q)). */
int ffo_nedelec_triplets_points(int k, int nvert, const double *vertices, int ntet,
                                const int *tets, double beta, const double *point_tensors,
                                int *rows, int *cols, double *vals, int *ndofs)
{
    double *qx, *qy, *qz, *qw;
    int nq = element_rule(k, 2 * k + 2, &qx, &qy, &qz, &qw);
    free(qx);
    return triplets_impl(k, nvert, vertices, ntet, tets, 0.0, beta, point_tensors, nq, rows,
                         cols, vals, ndofs);
}

static int triplets_impl(int k, int nvert, const double *vertices, int ntet, const int *tets,
                         double alpha, double beta, const double *cell_tensor, int point_nq,
                         int *rows, int *cols, double *vals, int *ndofs)
{
    int n = ffo_nedelec_dof_count(k);
    int *gd = malloc(sizeof(int) * (size_t)n * ntet);
    int *eo = malloc(sizeof(int) * 6 * (size_t)ntet);
    int *fp = malloc(sizeof(int) * 12 * (size_t)ntet);
    int nd = 0;
    if (ffo_nedelec_dof_map(k, ntet, tets, gd, eo, fp, &nd, NULL, NULL)) {
        free(gd), free(eo), free(fp);
        return 1;
    }
    *ndofs = nd;
    double *coeff = NULL;
    if (k >= 2) {
        coeff = malloc(sizeof(double) * n * n);
        if (ffo_nedelec_coefficients(k, coeff)) {
            free(gd), free(eo), free(fp), free(coeff);
            return 1;
        }
    }
    int qdeg = cell_tensor ? 2 * k + 2 : 2 * k;
    double *E = malloc(sizeof(double) * n * n), *T = malloc(sizeof(double) * n * n);
    double *ET = malloc(sizeof(double) * n * n), *O = malloc(sizeof(double) * n * n);
    size_t entry = 0;
    int st = 0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12];
        for (int i = 0; i < 4; ++i) {
            int id = tets[4 * e + i];
            if (id < 1 || id > nvert) {
                st = 1;
                break;
            }
            for (int c = 0; c < 3; ++c) v[3 * i + c] = vertices[3 * (id - 1) + c];
        }
        if (st) break;
        const double *K = cell_tensor ? cell_tensor + 9 * (size_t)e * (point_nq ? point_nq : 1)
                                      : NULL;
        if (element_impl(k, coeff, v, qdeg, alpha, K, point_nq ? 9 : 0, beta, E)) {
            st = 1;
            break;
        }
        if (k == 1) {
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r)
                    O[r + n * c] = (double)(eo[6 * e + r] * eo[6 * e + c]) * E[r + n * c];
        } else {
            ffo_nedelec_basis_transform(k, eo + 6 * e, fp + 12 * e, T);
            /* ET = E T (E column-major, T row-major): ET[r + n c] */
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) {
                    double s = 0;
                    for (int j = 0; j < n; ++j) {
                        double t = T[(size_t)j * n + c];
                        if (t != 0.0) s += E[r + (size_t)n * j] * t;
                    }
                    ET[r + (size_t)n * c] = s;
                }
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) {
                    double s = 0;
                    for (int j = 0; j < n; ++j) {
                        double t = T[(size_t)j * n + r];
                        if (t != 0.0) s += t * ET[j + (size_t)n * c];
                    }
                    O[r + (size_t)n * c] = s;
                }
        }
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r, ++entry) {
                rows[entry] = gd[(size_t)n * e + r];
                cols[entry] = gd[(size_t)n * e + c];
                vals[entry] = O[r + (size_t)n * c];
            }
    }
    free(gd), free(eo), free(fp), free(coeff), free(E), free(T), free(ET), free(O);
    return st;
}

/* ------------------------------------------------------ tangents (AD) */
/* reference values and curls of the basis at the 2k + 2 Duffy points (the
 * rule of the curl-mass tangents, :2160, 2266), shared by all cells */
typedef struct {
    int nq;
    double *x, *w, *rv, *rc;
} ned_rule;

static void ned_rule_init(ned_rule *R, int k, const double *coeff)
{
    int n = ffo_nedelec_dof_count(k);
    R->nq = ffo_tetra_duffy_count(2 * k + 2);
    R->x = malloc(sizeof(double) * 4 * R->nq);
    ffo_tetra_duffy(2 * k + 2, R->x, R->x + R->nq, R->x + 2 * R->nq, R->x + 3 * R->nq);
    R->w = R->x + 3 * R->nq;
    R->rv = malloc(sizeof(double) * 3 * n * R->nq);
    R->rc = malloc(sizeof(double) * 3 * n * R->nq);
    for (int q = 0; q < R->nq; ++q) {
        double p[3] = {R->x[q], R->x[R->nq + q], R->x[2 * R->nq + q]};
        double *rv = R->rv + (size_t)3 * n * q, *rc = R->rc + (size_t)3 * n * q;
        if (k == 1) ffo_whitney(p, rv, rc);
        else ffo_nedelec_eval(k, coeff, p, rv, rc);
    }
}

static void ned_rule_free(ned_rule *R)
{
    free(R->x), free(R->rv), free(R->rc);
}

/* Element tangent of assemble_tetra_nedelec_curl_mass_element_jvp
 * (assembly_tetra_nedelec_3d.f90:645-716).  map_tetra_nedelec_covariant_jvp
 * (tetra_piola_maps.f90) is restated from the map's closed form:
 * d(J^-T v^) = -J^-T dJ^T J^-T v^, d(J c^ / det) = dJ c^ / det - J c^ ddet / det^2,
 * ddet = det tr(J^-1 dJ).  The element rule is the reference's 2k + 2. */
static int element_jvp(int k, const ned_rule *R, const double v[12], const double vd[12],
                       double alpha, double beta, double alpha_dot, double beta_dot,
                       double *md)
{
    int n = ffo_nedelec_dof_count(k);
    double jac[9], inv[9], det, jd[9];
    if (geometry(v, jac, inv, &det)) return 1;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) jd[r + 3 * c] = vd[3 * (c + 1) + r] - vd[r];
    double tr = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tr += inv[i + 3 * j] * jd[j + 3 * i];
    double det_dot = det * tr;
    /* A = J^-T dJ^T J^-T, so values_dot = -A v^ */
    double a1[9], A[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += inv[q + 3 * i] * jd[j + 3 * q]; /* (J^-T dJ^T)_ij */
            a1[i + 3 * j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += a1[i + 3 * q] * inv[j + 3 * q]; /* (.. J^-T)_ij */
            A[i + 3 * j] = s;
        }
    const double *qw = R->w;
    double *pv = malloc(sizeof(double) * 3 * n), *pc = malloc(sizeof(double) * 3 * n);
    double *pvd = malloc(sizeof(double) * 3 * n), *pcd = malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n * n; ++i) md[i] = 0.0;
    for (int q = 0; q < R->nq; ++q) {
        const double *rv = R->rv + (size_t)3 * n * q, *rc = R->rc + (size_t)3 * n * q;
        for (int j = 0; j < n; ++j)
            for (int r = 0; r < 3; ++r) {
                pv[3 * j + r] = inv[0 + 3 * r] * rv[3 * j] + inv[1 + 3 * r] * rv[3 * j + 1] +
                                inv[2 + 3 * r] * rv[3 * j + 2];
                pvd[3 * j + r] = -(A[r] * rv[3 * j] + A[r + 3] * rv[3 * j + 1] +
                                   A[r + 6] * rv[3 * j + 2]);
                double jc = jac[r] * rc[3 * j] + jac[r + 3] * rc[3 * j + 1] +
                            jac[r + 6] * rc[3 * j + 2];
                double jdc = jd[r] * rc[3 * j] + jd[r + 3] * rc[3 * j + 1] +
                             jd[r + 6] * rc[3 * j + 2];
                pc[3 * j + r] = jc / det;
                pcd[3 * j + r] = jdc / det - jc * det_dot / (det * det);
            }
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double ce = 0, me = 0, cx = 0, mx = 0;
                for (int i = 0; i < 3; ++i) {
                    ce += pc[3 * r + i] * pc[3 * c + i];
                    me += pv[3 * r + i] * pv[3 * c + i];
                    cx += pcd[3 * r + i] * pc[3 * c + i] + pc[3 * r + i] * pcd[3 * c + i];
                    mx += pvd[3 * r + i] * pv[3 * c + i] + pv[3 * r + i] * pvd[3 * c + i];
                }
                md[r + (size_t)n * c] += qw[q] * (det_dot * (alpha * ce + beta * me) +
                                                  det * (alpha_dot * ce + alpha * cx +
                                                         beta_dot * me + beta * mx));
            }
    }
    free(pv), free(pc), free(pvd), free(pcd);
    return 0;
}

/* Element cotangent of assemble_tetra_nedelec_curl_mass_element_vjp
 * (:718-805); map_tetra_nedelec_covariant_vjp restated from the map:
 * Inv_bar(j, i) += vbar_i v^_j, J_bar(i, j) += cbar_i c^_j / det, and the
 * map's 1/det gives J_bar -= (sum cbar . curl) J^-T; then inv3_vjp, det3_vjp
 * and tetra_jacobian_vjp as in the H1 element. */
static int element_vjp(int k, const ned_rule *R, const double v[12], double alpha,
                       double beta, const double *mb, double vb[12], double *alpha_bar,
                       double *beta_bar)
{
    int n = ffo_nedelec_dof_count(k);
    double jac[9], inv[9], det;
    if (geometry(v, jac, inv, &det)) return 1;
    const double *qw = R->w;
    double *pv = malloc(sizeof(double) * 3 * n), *pc = malloc(sizeof(double) * 3 * n);
    double *vbar = malloc(sizeof(double) * 3 * n), *cbar = malloc(sizeof(double) * 3 * n);
    double det_bar = 0, ab = 0, bb = 0, invb[9] = {0}, jb[9] = {0};
    for (int q = 0; q < R->nq; ++q) {
        const double *rv = R->rv + (size_t)3 * n * q, *rc = R->rc + (size_t)3 * n * q;
        for (int j = 0; j < n; ++j)
            for (int r = 0; r < 3; ++r) {
                pv[3 * j + r] = inv[0 + 3 * r] * rv[3 * j] + inv[1 + 3 * r] * rv[3 * j + 1] +
                                inv[2 + 3 * r] * rv[3 * j + 2];
                pc[3 * j + r] = (jac[r] * rc[3 * j] + jac[r + 3] * rc[3 * j + 1] +
                                 jac[r + 6] * rc[3 * j + 2]) / det;
            }
        for (int i = 0; i < 3 * n; ++i) vbar[i] = cbar[i] = 0.0;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double seed = qw[q] * mb[r + (size_t)n * c];
                double ce = 0, me = 0;
                for (int i = 0; i < 3; ++i) {
                    ce += pc[3 * r + i] * pc[3 * c + i];
                    me += pv[3 * r + i] * pv[3 * c + i];
                }
                det_bar += seed * (alpha * ce + beta * me);
                ab += seed * det * ce;
                bb += seed * det * me;
                for (int i = 0; i < 3; ++i) {
                    cbar[3 * r + i] += seed * det * alpha * pc[3 * c + i];
                    cbar[3 * c + i] += seed * det * alpha * pc[3 * r + i];
                    vbar[3 * r + i] += seed * det * beta * pv[3 * c + i];
                    vbar[3 * c + i] += seed * det * beta * pv[3 * r + i];
                }
            }
        double cdot = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < 3; ++i) {
                cdot += cbar[3 * j + i] * pc[3 * j + i];
                for (int m = 0; m < 3; ++m) {
                    invb[m + 3 * i] += vbar[3 * j + i] * rv[3 * j + m];
                    jb[i + 3 * m] += cbar[3 * j + i] * rc[3 * j + m] / det;
                }
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) jb[i + 3 * j] -= cdot * inv[j + 3 * i];
    }
    /* J_bar += -J^-T Inv_bar J^-T + det_bar det J^-T */
    double w[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += inv[q + 3 * i] * invb[q + 3 * j];
            w[i + 3 * j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += w[i + 3 * q] * inv[j + 3 * q];
            jb[i + 3 * j] += -s + det_bar * det * inv[j + 3 * i];
        }
    for (int r = 0; r < 3; ++r) {
        vb[r] = -(jb[r] + jb[r + 3] + jb[r + 6]);
        for (int c = 0; c < 3; ++c) vb[3 * (c + 1) + r] = jb[r + 3 * c];
    }
    *alpha_bar = ab;
    *beta_bar = bb;
    free(pv), free(pc), free(vbar), free(cbar);
    return 0;
}

/* oriented = T^T E T (k >= 2, T row-major) or o_r o_c E (k = 1) */
static void orient(int k, int n, const int *eo, const int *fp, const double *E, double *O,
                   double *T, double *ET, int transpose)
{
    if (k == 1) {
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) O[r + n * c] = (double)(eo[r] * eo[c]) * E[r + n * c];
        return;
    }
    ffo_nedelec_basis_transform(k, eo, fp, T);
    /* transpose = 0: O = T^T E T; transpose = 1: O = T E T^T */
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) {
            double s = 0;
            for (int j = 0; j < n; ++j) {
                double t = transpose ? T[(size_t)c * n + j] : T[(size_t)j * n + c];
                if (t != 0.0) s += E[r + (size_t)n * j] * t;
            }
            ET[r + (size_t)n * c] = s;
        }
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) {
            double s = 0;
            for (int j = 0; j < n; ++j) {
                double t = transpose ? T[(size_t)r * n + j] : T[(size_t)j * n + r];
                if (t != 0.0) s += t * ET[j + (size_t)n * c];
            }
            O[r + (size_t)n * c] = s;
        }
}

static int ned_setup(int k, int ntet, const int *tets, int **gd, int **eo, int **fp, int *nd,
                     double **coeff)
{
    int n = ffo_nedelec_dof_count(k);
    *gd = malloc(sizeof(int) * (size_t)n * ntet);
    *eo = malloc(sizeof(int) * 6 * (size_t)ntet);
    *fp = malloc(sizeof(int) * 12 * (size_t)ntet);
    *coeff = NULL;
    if (ffo_nedelec_dof_map(k, ntet, tets, *gd, *eo, *fp, nd, NULL, NULL)) return 1;
    if (k >= 2) {
        *coeff = malloc(sizeof(double) * n * n);
        if (ffo_nedelec_coefficients(k, *coeff)) return 1;
    }
    return 0;
}

/* assemble_tetra_nedelec_curl_mass_csc_jvp (:2114-2180): triplets of the
 * oriented element tangents in the primal's order.  cell_alpha /
 * cell_alpha_dot (optional): per-cell curl coefficient and its tangent. */
int ffo_nedelec_jvp_triplets(int k, int nvert, const double *vertices,
                             const double *vertices_dot, int ntet, const int *tets,
                             double alpha, double beta, double alpha_dot, double beta_dot,
                             const double *cell_alpha, const double *cell_alpha_dot,
                             int *rows, int *cols, double *vals, int *ndofs)
{
    int n = ffo_nedelec_dof_count(k), *gd, *eo, *fp, st = 0;
    double *coeff;
    if (ned_setup(k, ntet, tets, &gd, &eo, &fp, ndofs, &coeff)) st = 1;
    double *E = malloc(sizeof(double) * n * n), *T = malloc(sizeof(double) * n * n);
    double *ET = malloc(sizeof(double) * n * n), *O = malloc(sizeof(double) * n * n);
    ned_rule R = {0};
    if (!st) ned_rule_init(&R, k, coeff);
    size_t entry = 0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12], vd[12];
        for (int i = 0; i < 4 && !st; ++i) {
            int id = tets[4 * e + i];
            if (id < 1 || id > nvert) st = 1;
            else
                for (int c = 0; c < 3; ++c) {
                    v[3 * i + c] = vertices[3 * (id - 1) + c];
                    vd[3 * i + c] = vertices_dot[3 * (id - 1) + c];
                }
        }
        if (st) break;
        double a = cell_alpha ? cell_alpha[e] : alpha;
        double ad = cell_alpha_dot ? cell_alpha_dot[e] : alpha_dot;
        if (element_jvp(k, &R, v, vd, a, beta, ad, beta_dot, E)) {
            st = 1;
            break;
        }
        orient(k, n, eo + 6 * e, fp + 12 * e, E, O, T, ET, 0);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r, ++entry) {
                rows[entry] = gd[(size_t)n * e + r];
                cols[entry] = gd[(size_t)n * e + c];
                vals[entry] = O[r + (size_t)n * c];
            }
    }
    if (R.x) ned_rule_free(&R);
    free(gd), free(eo), free(fp), free(coeff), free(E), free(T), free(ET), free(O);
    return st;
}

/* assemble_tetra_nedelec_curl_mass_csc_vjp (:2182-2287): per cell the
 * oriented cotangent read from the primal CSC, element_bar = T O T^T (k >=
 * 2) or o_r o_c O (k = 1), the element cotangent, vertex bars summed in tet
 * order; cell_alpha_bar (optional) gets each cell's alpha bar. */
int ffo_nedelec_vjp(int k, int nvert, const double *vertices, int ntet, const int *tets,
                    double alpha, double beta, const double *cell_alpha, const int *col_ptr,
                    const int *row_idx, const double *values_bar, double *vertices_bar,
                    double *alpha_bar, double *cell_alpha_bar, double *beta_bar)
{
    int n = ffo_nedelec_dof_count(k), *gd, *eo, *fp, nd, st = 0;
    double *coeff;
    if (ned_setup(k, ntet, tets, &gd, &eo, &fp, &nd, &coeff)) st = 1;
    double *Ob = malloc(sizeof(double) * n * n), *Eb = malloc(sizeof(double) * n * n);
    double *T = malloc(sizeof(double) * n * n), *ET = malloc(sizeof(double) * n * n);
    ned_rule R = {0};
    if (!st) ned_rule_init(&R, k, coeff);
    for (int i = 0; i < 3 * nvert; ++i) vertices_bar[i] = 0.0;
    *alpha_bar = 0.0;
    *beta_bar = 0.0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12], vb[12], ab, bb;
        for (int i = 0; i < 4 && !st; ++i) {
            int id = tets[4 * e + i];
            if (id < 1 || id > nvert) st = 1;
            else
                for (int c = 0; c < 3; ++c) v[3 * i + c] = vertices[3 * (id - 1) + c];
        }
        if (st) break;
        const int *g = gd + (size_t)n * e;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double b = 0.0;
                for (int q = col_ptr[g[c] - 1] - 1; q < col_ptr[g[c]] - 1; ++q)
                    if (row_idx[q] == g[r]) {
                        b = values_bar[q];
                        break;
                    }
                Ob[r + (size_t)n * c] = b;
            }
        orient(k, n, eo + 6 * e, fp + 12 * e, Ob, Eb, T, ET, 1);
        double a = cell_alpha ? cell_alpha[e] : alpha;
        if (element_vjp(k, &R, v, a, beta, Eb, vb, &ab, &bb)) {
            st = 1;
            break;
        }
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 3; ++c) vertices_bar[3 * (tets[4 * e + i] - 1) + c] += vb[3 * i + c];
        *alpha_bar += ab;
        *beta_bar += bb;
        if (cell_alpha_bar) cell_alpha_bar[e] = ab;
    }
    if (R.x) ned_rule_free(&R);
    free(gd), free(eo), free(fp), free(coeff), free(Ob), free(Eb), free(T), free(ET);
    return st;
}

/* ------------------------------------------------------------ PML (complex) */
/* Per-cell PML tensors, curl K and value M (complex, column-major):
 * cartesian_curl_curl_pml_coefficients (src/boundary/cartesian_helmholtz_pml.f90:
 * 23-41, 94-108): m_i = s_j s_k / s_i, K = diag(1/m), M = diag(m);
 * curvilinear_curl_curl_pml_coefficients (curvilinear_helmholtz_pml.f90:
 * 113-131 with inverse_and_determinant :202-216): K = S^T S / det S,
 * M = det S S^-1 S^-T.  st: (re, im) pairs. */
static int pml_tensors(int kind, const double *st, double complex K[9], double complex M[9])
{
    for (int i = 0; i < 9; ++i) K[i] = M[i] = 0.0;
    if (kind == 0) {
        double complex s[3];
        for (int i = 0; i < 3; ++i) {
            s[i] = st[2 * i] + I * st[2 * i + 1];
            if (cabs(s[i]) <= DBL_MIN) return 1;
        }
        for (int i = 0; i < 3; ++i) {
            double complex m = s[(i + 1) % 3] * s[(i + 2) % 3] / s[i];
            M[4 * i] = m;
            K[4 * i] = 1.0 / m;
        }
        return 0;
    }
    double complex S[9], cof[9], inv[9];
    double mx = 0.0;
    for (int i = 0; i < 9; ++i) {
        S[i] = st[2 * i] + I * st[2 * i + 1];
        if (cabs(S[i]) > mx) mx = cabs(S[i]);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            cof[r + 3 * c] = S[r1 + 3 * c1] * S[r2 + 3 * c2] - S[r1 + 3 * c2] * S[r2 + 3 * c1];
        }
    double complex det = S[0] * cof[0] + S[3] * cof[3] + S[6] * cof[6];
    double sc = mx > 1.0 ? mx : 1.0;
    if (!(cabs(det) > 64.0 * DBL_EPSILON * sc * sc * sc)) return 1;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inv[r + 3 * c] = cof[c + 3 * r] / det;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double complex a = 0, b = 0;
            for (int t = 0; t < 3; ++t) {
                a += S[t + 3 * r] * S[t + 3 * c];
                b += inv[r + 3 * t] * inv[c + 3 * t];
            }
            K[r + 3 * c] = a / det;
            M[r + 3 * c] = det * b;
        }
    return 0;
}

/* PML elements (assembly_tetra_nedelec_3d.f90:826-887 cartesian, :1092-1150
 * curvilinear): per point the covariant physical values and curls, then
 * E(r,c) += det w (curl_r . K curl_c - k^2 val_r . M val_c). */
static int pml_element(int k, const ned_rule *R, const double v[12], const double complex K[9],
                       const double complex M[9], double k0, double complex *E)
{
    int n = ffo_nedelec_dof_count(k);
    double jac[9], inv[9], det;
    if (geometry(v, jac, inv, &det)) return 1;
    double *pv = malloc(sizeof(double) * 3 * n), *pc = malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n * n; ++i) E[i] = 0.0;
    for (int q = 0; q < R->nq; ++q) {
        const double *rv = R->rv + (size_t)3 * n * q, *rc = R->rc + (size_t)3 * n * q;
        for (int j = 0; j < n; ++j)
            for (int r = 0; r < 3; ++r) {
                pv[3 * j + r] = inv[0 + 3 * r] * rv[3 * j] + inv[1 + 3 * r] * rv[3 * j + 1] +
                                inv[2 + 3 * r] * rv[3 * j + 2];
                pc[3 * j + r] = (jac[r] * rc[3 * j] + jac[r + 3] * rc[3 * j + 1] +
                                 jac[r + 6] * rc[3 * j + 2]) / det;
            }
        double pw = det * R->w[q];
        for (int c = 0; c < n; ++c) {
            double complex kc[3], mc[3];
            for (int i = 0; i < 3; ++i) {
                kc[i] = K[i] * pc[3 * c] + K[i + 3] * pc[3 * c + 1] + K[i + 6] * pc[3 * c + 2];
                mc[i] = M[i] * pv[3 * c] + M[i + 3] * pv[3 * c + 1] + M[i + 6] * pv[3 * c + 2];
            }
            for (int r = 0; r < n; ++r) {
                double complex cc = pc[3 * r] * kc[0] + pc[3 * r + 1] * kc[1] + pc[3 * r + 2] * kc[2];
                double complex mm = pv[3 * r] * mc[0] + pv[3 * r + 1] * mc[1] + pv[3 * r + 2] * mc[2];
                E[r + (size_t)n * c] += pw * (cc - k0 * k0 * mm);
            }
        }
    }
    free(pv), free(pc);
    return 0;
}

/* assemble_tetra_nedelec_pml_csc (:1613-1685, kind 0, stretch 3 complex
 * per cell) and assemble_tetra_nedelec_curvilinear_pml_csc (:1372-1444,
 * kind 1, 3 x 3 complex column-major per cell): oriented complex element
 * matrices as triplets in the primal order; vals holds (re, im) pairs. */
int ffo_nedelec_pml_triplets(int k, int nvert, const double *vertices, int ntet,
                             const int *tets, int kind, const double *stretch, double k0,
                             int *rows, int *cols, double *vals, int *ndofs)
{
    if (k < 1 || k0 <= 0.0 || (kind != 0 && kind != 1)) return 1;
    int n = ffo_nedelec_dof_count(k), *gd, *eo, *fp, st = 0;
    double *coeff;
    if (ned_setup(k, ntet, tets, &gd, &eo, &fp, ndofs, &coeff)) st = 1;
    double complex *E = malloc(sizeof(double complex) * n * n);
    double *Er = malloc(sizeof(double) * n * n), *Ei = malloc(sizeof(double) * n * n);
    double *Or = malloc(sizeof(double) * n * n), *Oi = malloc(sizeof(double) * n * n);
    double *T = malloc(sizeof(double) * n * n), *ET = malloc(sizeof(double) * n * n);
    ned_rule R = {0};
    if (!st) ned_rule_init(&R, k, coeff);
    size_t entry = 0;
    const int per = kind == 0 ? 6 : 18;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12];
        for (int i = 0; i < 4 && !st; ++i) {
            int id = tets[4 * e + i];
            if (id < 1 || id > nvert) st = 1;
            else
                for (int c = 0; c < 3; ++c) v[3 * i + c] = vertices[3 * (id - 1) + c];
        }
        if (st) break;
        double complex K[9], M[9];
        if (pml_tensors(kind, stretch + (size_t)per * e, K, M) ||
            pml_element(k, &R, v, K, M, k0, E)) {
            st = 1;
            break;
        }
        for (int i = 0; i < n * n; ++i) {
            Er[i] = creal(E[i]);
            Ei[i] = cimag(E[i]);
        }
        orient(k, n, eo + 6 * e, fp + 12 * e, Er, Or, T, ET, 0);
        orient(k, n, eo + 6 * e, fp + 12 * e, Ei, Oi, T, ET, 0);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r, ++entry) {
                rows[entry] = gd[(size_t)n * e + r];
                cols[entry] = gd[(size_t)n * e + c];
                vals[2 * entry] = Or[r + (size_t)n * c];
                vals[2 * entry + 1] = Oi[r + (size_t)n * c];
            }
    }
    if (R.x) ned_rule_free(&R);
    free(gd), free(eo), free(fp), free(coeff), free(E), free(Er), free(Ei), free(Or), free(Oi);
    free(T), free(ET);
    return st;
}
