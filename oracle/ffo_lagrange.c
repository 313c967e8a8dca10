/*
 * ffo_lagrange.c -- oracle restatement of the tetrahedral Lagrange H1 path:
 * basis, global DOF numbering, element stiffness, triplets, sampled load and
 * Dirichlet data.  Test infrastructure only (see ffo.h).
 */
#include "ffo.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* (degree+1)(degree+2)(degree+3)/6, tetra_lagrange_arbitrary_order.f90:44 */
int ffo_lagrange_dof_count(int degree)
{
    if (degree < 0) return -1;
    return (degree + 1) * (degree + 2) * (degree + 3) / 6;
}

/* initialize_tetra_lagrange (src/elements/tetra_lagrange_arbitrary_order.f90:
 * 29-66).  barycentric_indices(4,n), nodes(3,n); loops first->second->third,
 * fourth = degree - first - second - third. */
int ffo_lagrange_init(int degree, int *bary, double *nodes)
{
    if (degree < 0) return 1;
    if (degree == 0) {
        bary[0] = bary[1] = bary[2] = bary[3] = 0;
        nodes[0] = nodes[1] = nodes[2] = 0.25;
        return 0;
    }
    int dof = 0;
    for (int a = 0; a <= degree; ++a)
        for (int b = 0; b <= degree - a; ++b)
            for (int c = 0; c <= degree - a - b; ++c, ++dof) {
                int d = degree - a - b - c;
                bary[4 * dof + 0] = a;
                bary[4 * dof + 1] = b;
                bary[4 * dof + 2] = c;
                bary[4 * dof + 3] = d;
                nodes[3 * dof + 0] = (double)b / (double)degree;
                nodes[3 * dof + 1] = (double)c / (double)degree;
                nodes[3 * dof + 2] = (double)d / (double)degree;
            }
    return 0;
}

static double factorial(int n)
{
    double v = 1.0;
    for (int f = 2; f <= n; ++f) v *= (double)f;
    return v;
}

/* cardinal_factor (tetra_lagrange_arbitrary_order.f90:207-249):
 * l_i(lambda) = prod_{f<i} (d*lambda - f) / i! and its derivative. */
static void cardinal_factor(int index, int degree, double lambda,
                            double *value, double *derivative)
{
    double v = 1.0;
    for (int f = 0; f < index; ++f) v = v * ((double)degree * lambda - f);
    *value = v / factorial(index);
    double d = 0.0;
    for (int df = 0; df < index; ++df) {
        double term = (double)degree;
        for (int f = 0; f < index; ++f)
            if (f != df) term = term * ((double)degree * lambda - f);
        d = d + term;
    }
    *derivative = d / factorial(index);
}

/* evaluate_tetra_lagrange (tetra_lagrange_arbitrary_order.f90:68-116):
 * lambda = [1-x-y-z, x, y, z]; d/dx_j = d/dlambda_{j+1} - d/dlambda_1. */
int ffo_lagrange_eval(int degree, const double point[3], double *values,
                      double *gradients)
{
    if (degree < 0) return 1;
    int n = ffo_lagrange_dof_count(degree);
    int *bary = malloc(sizeof(int) * 4 * n);
    double *nodes = malloc(sizeof(double) * 3 * n);
    ffo_lagrange_init(degree, bary, nodes);
    double lam[4] = {1.0 - (point[0] + point[1] + point[2]), point[0],
                     point[1], point[2]};
    int bad = 0;
    for (int c = 0; c < 4; ++c)
        if (lam[c] < -64.0 * DBL_EPSILON) bad = 1;
    if (bad) {
        free(bary);
        free(nodes);
        for (int i = 0; i < n; ++i) {
            values[i] = 0.0;
            gradients[3 * i] = gradients[3 * i + 1] = gradients[3 * i + 2] = 0;
        }
        return 1;
    }
    for (int i = 0; i < n; ++i) {
        double fac[4], der[4], ld[4];
        for (int c = 0; c < 4; ++c)
            cardinal_factor(bary[4 * i + c], degree, lam[c], &fac[c], &der[c]);
        values[i] = ((fac[0] * fac[1]) * fac[2]) * fac[3];
        for (int c = 0; c < 4; ++c) {
            ld[c] = der[c];
            for (int k = 0; k < 4; ++k)
                if (k != c) ld[c] = ld[c] * fac[k];
        }
        for (int k = 0; k < 3; ++k) gradients[3 * i + k] = ld[k + 1] - ld[0];
    }
    free(bary);
    free(nodes);
    return 0;
}

/* ------------------------------------------------------------------------
 * Global DOF numbering: build_tetra_lagrange_dof_map
 * (src/elements/tetra_lagrange_global_dof_map.f90:13-79, sort_key :81-99).
 * A DOF is keyed by the vertex ids with a nonzero barycentric index (sorted
 * ascending, zeros last) plus those indices; numbers are given by first
 * appearance over (cell, local dof).  The reference does a linear search
 * (:62); ffo_lagrange_dof_map_search restates it literally, and
 * ffo_lagrange_dof_map gives the identical numbering with a hash table.
 * ---------------------------------------------------------------------- */
typedef struct {
    int v[4];
    int w[4];
} dof_key;

static void sort_key(int v[4], int w[4])
{
    /* sort_key, tetra_lagrange_global_dof_map.f90:81-99 */
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 4; ++b) {
            if (v[b] == 0) continue;
            if (v[a] != 0 && v[a] <= v[b]) continue;
            int t = v[a];
            v[a] = v[b];
            v[b] = t;
            t = w[a];
            w[a] = w[b];
            w[b] = t;
        }
}

static int validate_cells(int degree, int ntet, const int *tets)
{
    if (degree < 1 || degree > 4) return 1;
    for (int e = 0; e < ntet; ++e)
        for (int k = 0; k < 4; ++k) {
            if (tets[4 * e + k] < 1) return 1;
            int count = 0;
            for (int j = 0; j < 4; ++j)
                if (tets[4 * e + j] == tets[4 * e + k]) ++count;
            if (count != 1) return 1;
        }
    return 0;
}

static dof_key make_key(const int *cell, const int *bary_dof)
{
    dof_key key;
    int existing = 0;
    memset(&key, 0, sizeof key);
    for (int vtx = 0; vtx < 4; ++vtx) {
        if (bary_dof[vtx] == 0) continue;
        key.v[existing] = cell[vtx];
        key.w[existing] = bary_dof[vtx];
        ++existing;
    }
    sort_key(key.v, key.w);
    return key;
}

int ffo_lagrange_dof_map_search(int degree, int ntet, const int *tets,
                                int *global_dofs, int *global_count)
{
    *global_count = 0;
    if (validate_cells(degree, ntet, tets)) return 1;
    int n = ffo_lagrange_dof_count(degree);
    int *bary = malloc(sizeof(int) * 4 * n);
    double *nodes = malloc(sizeof(double) * 3 * n);
    ffo_lagrange_init(degree, bary, nodes);
    dof_key *keys = malloc(sizeof(dof_key) * (size_t)n * ntet);
    int count = 0;
    for (int e = 0; e < ntet; ++e)
        for (int l = 0; l < n; ++l) {
            dof_key key = make_key(tets + 4 * e, bary + 4 * l);
            int found = -1;
            for (int g = 0; g < count; ++g)
                if (memcmp(&keys[g], &key, sizeof key) == 0) {
                    found = g;
                    break;
                }
            if (found < 0) {
                keys[count] = key;
                found = count++;
            }
            global_dofs[(size_t)n * e + l] = found + 1;
        }
    *global_count = count;
    free(keys);
    free(bary);
    free(nodes);
    return 0;
}

static uint64_t hash_key(const dof_key *k)
{
    uint64_t h = 1469598103934665603ull;
    const int *p = (const int *)k;
    for (int i = 0; i < 8; ++i) {
        h ^= (uint64_t)(uint32_t)p[i];
        h *= 1099511628211ull;
        h ^= h >> 29;
    }
    return h;
}

int ffo_lagrange_dof_map(int degree, int ntet, const int *tets,
                         int *global_dofs, int *global_count)
{
    *global_count = 0;
    if (validate_cells(degree, ntet, tets)) return 1;
    int n = ffo_lagrange_dof_count(degree);
    int *bary = malloc(sizeof(int) * 4 * n);
    double *nodes = malloc(sizeof(double) * 3 * n);
    ffo_lagrange_init(degree, bary, nodes);
    size_t cap = 1;
    while (cap < (size_t)n * ntet * 2) cap <<= 1;
    dof_key *table = calloc(cap, sizeof(dof_key));
    int *ids = malloc(sizeof(int) * cap);
    for (size_t i = 0; i < cap; ++i) ids[i] = -1;
    int count = 0;
    for (int e = 0; e < ntet; ++e)
        for (int l = 0; l < n; ++l) {
            dof_key key = make_key(tets + 4 * e, bary + 4 * l);
            size_t slot = hash_key(&key) & (cap - 1);
            while (ids[slot] >= 0 &&
                   memcmp(&table[slot], &key, sizeof key) != 0)
                slot = (slot + 1) & (cap - 1);
            if (ids[slot] < 0) {
                table[slot] = key;
                ids[slot] = count++;
            }
            global_dofs[(size_t)n * e + l] = ids[slot] + 1;
        }
    *global_count = count;
    free(table);
    free(ids);
    free(bary);
    free(nodes);
    return 0;
}

/* ------------------------------------------------------------------------
 * Element stiffness: assemble_tetra_lagrange_stiffness_element
 * (src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:40-105).
 * J(:,c) = v_{c+1} - v_1; reject det <= 64 eps max(1, max|J|)^3;
 * K(r,c) += det w_q (kappa (J^-T g_r).(J^-T g_c) + mu phi_r phi_c).
 * The physical gradients are the same numbers the reference recomputes inside
 * its (col,row) loop (:92-99); they are computed once per point here.
 * ---------------------------------------------------------------------- */
static int tetra_geometry(const double v[12], double jac[9], double inv[9],
                          double *det)
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            jac[r + 3 * c] = v[3 * (c + 1) + r] - v[r];
    *det = ffo_det3(jac);
    double mx = 0.0;
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(jac[i]));
    double s = fmax(1.0, mx);
    if (*det <= 64.0 * DBL_EPSILON * s * s * s) return 1;
    return ffo_inv3(jac, inv);
}

typedef struct {
    int degree, qdeg, n, nq;
    double *qx, *qy, *qz, *qw;
    double *values;    /* nq x n */
    double *gradients; /* nq x 3 x n */
} lagrange_table;

static int table_init(lagrange_table *t, int degree, int qdeg)
{
    t->degree = degree;
    t->qdeg = qdeg;
    t->n = ffo_lagrange_dof_count(degree);
    t->nq = ffo_tetra_duffy_count(qdeg);
    if (t->n <= 0 || t->nq <= 0) return 1;
    t->qx = malloc(sizeof(double) * t->nq);
    t->qy = malloc(sizeof(double) * t->nq);
    t->qz = malloc(sizeof(double) * t->nq);
    t->qw = malloc(sizeof(double) * t->nq);
    ffo_tetra_duffy(qdeg, t->qx, t->qy, t->qz, t->qw);
    t->values = malloc(sizeof(double) * t->nq * t->n);
    t->gradients = malloc(sizeof(double) * t->nq * 3 * t->n);
    for (int q = 0; q < t->nq; ++q) {
        double p[3] = {t->qx[q], t->qy[q], t->qz[q]};
        if (ffo_lagrange_eval(degree, p, t->values + (size_t)q * t->n,
                              t->gradients + (size_t)q * 3 * t->n))
            return 1;
    }
    return 0;
}

static void table_free(lagrange_table *t)
{
    free(t->qx);
    free(t->qy);
    free(t->qz);
    free(t->qw);
    free(t->values);
    free(t->gradients);
}

static int element_with_table(const lagrange_table *t, const double v[12],
                              double kappa, double mu, double *m,
                              double *pg /* 3 x n scratch */)
{
    double jac[9], inv[9], det;
    if (tetra_geometry(v, jac, inv, &det)) return 1;
    int n = t->n;
    for (int i = 0; i < n * n; ++i) m[i] = 0.0;
    for (int q = 0; q < t->nq; ++q) {
        const double *val = t->values + (size_t)q * n;
        const double *g = t->gradients + (size_t)q * 3 * n;
        double pw = det * t->qw[q];
        for (int r = 0; r < n; ++r)
            for (int i = 0; i < 3; ++i)
                pg[3 * r + i] = inv[0 + 3 * i] * g[3 * r + 0] +
                                inv[1 + 3 * i] * g[3 * r + 1] +
                                inv[2 + 3 * i] * g[3 * r + 2];
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double dot = pg[3 * r] * pg[3 * c] +
                             pg[3 * r + 1] * pg[3 * c + 1] +
                             pg[3 * r + 2] * pg[3 * c + 2];
                m[r + n * c] = m[r + n * c] +
                               pw * (kappa * dot + mu * val[r] * val[c]);
            }
    }
    return 0;
}

int ffo_lagrange_stiffness_element(const double vertices[12], int degree,
                                   int quadrature_degree, double kappa,
                                   double mu, double *matrix)
{
    if (degree < 1 || degree > 4 || quadrature_degree < 0) return 1;
    lagrange_table t;
    if (table_init(&t, degree, quadrature_degree)) return 1;
    double *pg = malloc(sizeof(double) * 3 * t.n);
    int st = element_with_table(&t, vertices, kappa, mu, matrix, pg);
    free(pg);
    table_free(&t);
    return st;
}

/* assemble_tetra_lagrange_stiffness_csc (:276-334), triplet stage: per tet,
 * column-outer / row-inner (:322-329).  cell_kappa (optional, per tet)
 * replaces the scalar stiffness coefficient as the per-cell variant of the
 * same composition (element routine called with stiffness_coefficient =
 * kappa_e, :47-48). */
int ffo_lagrange_stiffness_triplets(int nvert, const double *vertices,
                                    int ntet, const int *tets, int degree,
                                    int qdeg, const double *cell_kappa,
                                    double kappa, double mu,
                                    const int *global_dofs, int *rows,
                                    int *cols, double *vals)
{
    lagrange_table t;
    if (table_init(&t, degree, qdeg)) return 1;
    int n = t.n;
    double *m = malloc(sizeof(double) * n * n);
    double *pg = malloc(sizeof(double) * 3 * n);
    size_t entry = 0;
    int st = 0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12];
        for (int k = 0; k < 4; ++k) {
            int id = tets[4 * e + k];
            if (id < 1 || id > nvert) { st = 1; break; }
            for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
        }
        if (st) break;
        double ke = cell_kappa ? cell_kappa[e] : kappa;
        if (element_with_table(&t, v, ke, mu, m, pg)) { st = 1; break; }
        const int *g = global_dofs + (size_t)n * e;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r, ++entry) {
                rows[entry] = g[r];
                cols[entry] = g[c];
                vals[entry] = m[r + n * c];
            }
    }
    free(m);
    free(pg);
    table_free(&t);
    return st;
}

/* fortnum's det3_jvp / inv3_jvp / inv3_vjp / det3_vjp (absent dependency,
 * used at assembly_tetra_lagrange_arbitrary_order_3d.f90:135-139, 265-269),
 * restated from their closed forms: d det = det tr(J^-1 dJ),
 * d J^-1 = -J^-1 dJ J^-1, and the adjoints J_bar = -J^-T Inv_bar J^-T,
 * J_bar = det_bar det J^-T.  Matrices column-major. */
static void mat3_mul(const double *a, const double *b, double *c)
{
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i + 3 * k] * b[k + 3 * j];
            c[i + 3 * j] = s;
        }
}

static void mat3_transpose(const double *a, double *t)
{
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) t[j + 3 * i] = a[i + 3 * j];
}

/* element JVP (:107-174): per point physical gradients and their tangents
 * J^-T g, dJ^-T g; m_dot(r,c) += w (det_dot (k G + mu M) + det (k_dot G +
 * k (dG_r.G_c + G_r.dG_c) + mu_dot M)). */
static int element_jvp_with_table(const lagrange_table *t, const double v[12],
                                  const double vd[12], double kappa, double mu,
                                  double kappa_dot, double mu_dot, double *md,
                                  double *pg, double *pgd)
{
    double jac[9], inv[9], det, jd[9], tmp[9], invd[9];
    if (tetra_geometry(v, jac, inv, &det)) return 1;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) jd[r + 3 * c] = vd[3 * (c + 1) + r] - vd[r];
    mat3_mul(inv, jd, tmp);
    double det_dot = det * (tmp[0] + tmp[4] + tmp[8]);
    mat3_mul(tmp, inv, invd);
    for (int i = 0; i < 9; ++i) invd[i] = -invd[i];
    int n = t->n;
    for (int i = 0; i < n * n; ++i) md[i] = 0.0;
    for (int q = 0; q < t->nq; ++q) {
        const double *val = t->values + (size_t)q * n;
        const double *g = t->gradients + (size_t)q * 3 * n;
        for (int r = 0; r < n; ++r)
            for (int i = 0; i < 3; ++i) {
                pg[3 * r + i] = inv[0 + 3 * i] * g[3 * r + 0] + inv[1 + 3 * i] * g[3 * r + 1] +
                                inv[2 + 3 * i] * g[3 * r + 2];
                pgd[3 * r + i] = invd[0 + 3 * i] * g[3 * r + 0] +
                                 invd[1 + 3 * i] * g[3 * r + 1] +
                                 invd[2 + 3 * i] * g[3 * r + 2];
            }
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double ge = pg[3 * r] * pg[3 * c] + pg[3 * r + 1] * pg[3 * c + 1] +
                            pg[3 * r + 2] * pg[3 * c + 2];
                double me = val[r] * val[c];
                double cross = (pgd[3 * r] * pg[3 * c] + pgd[3 * r + 1] * pg[3 * c + 1] +
                                pgd[3 * r + 2] * pg[3 * c + 2]) +
                               (pg[3 * r] * pgd[3 * c] + pg[3 * r + 1] * pgd[3 * c + 1] +
                                pg[3 * r + 2] * pgd[3 * c + 2]);
                md[r + n * c] = md[r + n * c] +
                                t->qw[q] * (det_dot * (kappa * ge + mu * me) +
                                            det * (kappa_dot * ge + kappa * cross +
                                                   mu_dot * me));
            }
    }
    return 0;
}

/* element VJP (:176-275): accumulate det_bar, kappa_bar, mu_bar and
 * Inv_bar = sum_q g pg_bar^T over the points, then through inv3_vjp,
 * det3_vjp and tetra_jacobian_vjp (:428-437) to the vertices. */
static int element_vjp_with_table(const lagrange_table *t, const double v[12],
                                  double kappa, double mu, const double *mb,
                                  double vb[12], double *kappa_bar, double *mu_bar,
                                  double *pg, double *pgb)
{
    double jac[9], inv[9], det;
    if (tetra_geometry(v, jac, inv, &det)) return 1;
    int n = t->n;
    double det_bar = 0.0, kb = 0.0, mub = 0.0, invb[9] = {0};
    for (int q = 0; q < t->nq; ++q) {
        const double *val = t->values + (size_t)q * n;
        const double *g = t->gradients + (size_t)q * 3 * n;
        for (int r = 0; r < n; ++r)
            for (int i = 0; i < 3; ++i) {
                pg[3 * r + i] = inv[0 + 3 * i] * g[3 * r + 0] + inv[1 + 3 * i] * g[3 * r + 1] +
                                inv[2 + 3 * i] * g[3 * r + 2];
                pgb[3 * r + i] = 0.0;
            }
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double seed = t->qw[q] * mb[r + n * c];
                double ge = pg[3 * r] * pg[3 * c] + pg[3 * r + 1] * pg[3 * c + 1] +
                            pg[3 * r + 2] * pg[3 * c + 2];
                double me = val[r] * val[c];
                det_bar += seed * (kappa * ge + mu * me);
                kb += seed * det * ge;
                mub += seed * det * me;
                for (int i = 0; i < 3; ++i) {
                    pgb[3 * r + i] += seed * det * kappa * pg[3 * c + i];
                    pgb[3 * c + i] += seed * det * kappa * pg[3 * r + i];
                }
            }
        /* Inv_bar += gradients pg_bar^T: Inv_bar(j, i) += sum_r g(j,r) pgb(i,r) */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int r = 0; r < n; ++r) s += g[3 * r + j] * pgb[3 * r + i];
                invb[j + 3 * i] += s;
            }
    }
    double it[9], a[9], jb[9];
    mat3_transpose(inv, it);
    mat3_mul(it, invb, a);
    mat3_mul(a, it, jb);
    for (int i = 0; i < 9; ++i) jb[i] = -jb[i] + det_bar * det * it[i];
    for (int r = 0; r < 3; ++r) {
        vb[r] = -(jb[r] + jb[r + 3] + jb[r + 6]);
        for (int c = 0; c < 3; ++c) vb[3 * (c + 1) + r] = jb[r + 3 * c];
    }
    *kappa_bar = kb;
    *mu_bar = mub;
    return 0;
}

static int gather_vertices(int nvert, const double *vertices, const int *tet, double v[12])
{
    for (int k = 0; k < 4; ++k) {
        int id = tet[k];
        if (id < 1 || id > nvert) return 1;
        for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
    }
    return 0;
}

/* assemble_tetra_lagrange_stiffness_csc_jvp (:336-386), triplet stage in
 * the primal's order; cell_kappa / cell_kappa_dot (optional) are the
 * per-cell form of the scalar coefficient and its tangent. */
int ffo_lagrange_stiffness_jvp_triplets(int nvert, const double *vertices,
                                        const double *vertices_dot, int ntet,
                                        const int *tets, int degree, int qdeg,
                                        const double *cell_kappa,
                                        const double *cell_kappa_dot, double kappa,
                                        double mu, double kappa_dot, double mu_dot,
                                        const int *global_dofs, int *rows, int *cols,
                                        double *vals)
{
    lagrange_table t;
    if (degree < 1 || degree > 4 || qdeg < 0) return 1;
    if (table_init(&t, degree, qdeg)) return 1;
    int n = t.n;
    double *m = malloc(sizeof(double) * n * n);
    double *pg = malloc(sizeof(double) * 3 * n);
    double *pgd = malloc(sizeof(double) * 3 * n);
    size_t entry = 0;
    int st = 0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12], vd[12];
        if (gather_vertices(nvert, vertices, tets + 4 * e, v) ||
            gather_vertices(nvert, vertices_dot, tets + 4 * e, vd)) {
            st = 1;
            break;
        }
        double ke = cell_kappa ? cell_kappa[e] : kappa;
        double ked = cell_kappa_dot ? cell_kappa_dot[e] : kappa_dot;
        if (element_jvp_with_table(&t, v, vd, ke, mu, ked, mu_dot, m, pg, pgd)) {
            st = 1;
            break;
        }
        const int *g = global_dofs + (size_t)n * e;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r, ++entry) {
                rows[entry] = g[r];
                cols[entry] = g[c];
                vals[entry] = m[r + n * c];
            }
    }
    free(m);
    free(pg);
    free(pgd);
    table_free(&t);
    return st;
}

/* assemble_tetra_lagrange_stiffness_csc_vjp (:388-471): element cotangents
 * read from the primal CSC (csc_value_bar_at, :836-851: the entry of
 * (row, col), 0 when absent), element VJPs, vertex bars summed per tet in
 * tet order.  cell_kappa_bar (optional) receives each cell's kappa bar;
 * kappa_bar is their sum in tet order. */
int ffo_lagrange_stiffness_vjp(int nvert, const double *vertices, int ntet,
                               const int *tets, int degree, int qdeg,
                               const double *cell_kappa, double kappa, double mu,
                               const int *global_dofs, const int *col_ptr,
                               const int *row_idx, const double *values_bar,
                               double *vertices_bar, double *kappa_bar,
                               double *cell_kappa_bar, double *mu_bar)
{
    lagrange_table t;
    if (degree < 1 || degree > 4 || qdeg < 0) return 1;
    if (table_init(&t, degree, qdeg)) return 1;
    int n = t.n;
    double *eb = malloc(sizeof(double) * n * n);
    double *pg = malloc(sizeof(double) * 3 * n);
    double *pgb = malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < 3 * nvert; ++i) vertices_bar[i] = 0.0;
    *kappa_bar = 0.0;
    *mu_bar = 0.0;
    int st = 0;
    for (int e = 0; e < ntet && !st; ++e) {
        double v[12], vb[12], kb, mb;
        if (gather_vertices(nvert, vertices, tets + 4 * e, v)) {
            st = 1;
            break;
        }
        const int *g = global_dofs + (size_t)n * e;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                double b = 0.0;
                int col = g[c];
                for (int k = col_ptr[col - 1] - 1; k < col_ptr[col] - 1; ++k)
                    if (row_idx[k] == g[r]) {
                        b = values_bar[k];
                        break;
                    }
                eb[r + n * c] = b;
            }
        double ke = cell_kappa ? cell_kappa[e] : kappa;
        if (element_vjp_with_table(&t, v, ke, mu, eb, vb, &kb, &mb, pg, pgb)) {
            st = 1;
            break;
        }
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c)
                vertices_bar[3 * (tets[4 * e + k] - 1) + c] += vb[3 * k + c];
        *kappa_bar += kb;
        *mu_bar += mb;
        if (cell_kappa_bar) cell_kappa_bar[e] = kb;
    }
    free(eb);
    free(pg);
    free(pgb);
    table_free(&t);
    return st;
}

/* assemble_tetra_lagrange_scalar_load_samples (:473-528):
 * b(g) += det w_q f(q,e) phi(q).  source_values is (nq, T) column-major. */
int ffo_lagrange_load_samples(int nvert, const double *vertices, int ntet,
                              const int *tets, int degree, int qdeg,
                              const double *src, const int *global_dofs,
                              int global_count, double *rhs)
{
    lagrange_table t;
    if (table_init(&t, degree, qdeg)) return 1;
    int n = t.n;
    for (int i = 0; i < global_count; ++i) rhs[i] = 0.0;
    for (int e = 0; e < ntet; ++e) {
        double v[12], jac[9], inv[9], det;
        for (int k = 0; k < 4; ++k) {
            int id = tets[4 * e + k];
            if (id < 1 || id > nvert) { table_free(&t); return 1; }
            for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
        }
        if (tetra_geometry(v, jac, inv, &det)) { table_free(&t); return 1; }
        const int *g = global_dofs + (size_t)n * e;
        for (int q = 0; q < t.nq; ++q) {
            const double *val = t.values + (size_t)q * n;
            double f = src[(size_t)t.nq * e + q];
            for (int d = 0; d < n; ++d)
                rhs[g[d] - 1] = rhs[g[d] - 1] + det * t.qw[q] * f * val[d];
        }
    }
    table_free(&t);
    return 0;
}

/* Physical quadrature points x_q = v_1 + J xi_q
 * (assemble_tetra_lagrange_scalar_load, :761-763), points(3, nq, T), so a
 * caller can sample a source f at exactly the points the reference's
 * callback sees. */
int ffo_tetra_physical_points(int nvert, const double *vertices, int ntet,
                              const int *tets, int qdeg, double *points)
{
    int nq = ffo_tetra_duffy_count(qdeg);
    if (nq <= 0) return 1;
    double *qx = malloc(sizeof(double) * nq * 4);
    double *qy = qx + nq, *qz = qy + nq, *qw = qz + nq;
    ffo_tetra_duffy(qdeg, qx, qy, qz, qw);
    for (int e = 0; e < ntet; ++e) {
        double v[12], jac[9];
        for (int k = 0; k < 4; ++k) {
            int id = tets[4 * e + k];
            if (id < 1 || id > nvert) { free(qx); return 1; }
            for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
        }
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r)
                jac[r + 3 * c] = v[3 * (c + 1) + r] - v[r];
        for (int q = 0; q < nq; ++q)
            for (int r = 0; r < 3; ++r)
                points[3 * ((size_t)nq * e + q) + r] =
                    v[r] + (jac[r] * qx[q] + jac[r + 3] * qy[q] +
                            jac[r + 6] * qz[q]);
    }
    free(qx);
    return 0;
}

/* ------------------------------------------------------------------------
 * Dirichlet data: build_dirichlet_data
 * (src/solvers/fortfem_tetra_lagrange_solver_3d.f90:551-610).  A boundary
 * face is a face contained in exactly one cell (:587-597); the constrained
 * local dofs of that cell are those with barycentric index 0 at the opposite
 * vertex (:598-606).  node_points(3, G) receives v_1 + J node for every
 * constrained dof (last writer in cell order, as in the reference loop) so a
 * caller can evaluate the boundary value g there.
 * ---------------------------------------------------------------------- */
typedef struct {
    int v[3];
} face_key;

static uint64_t hash_face(const face_key *k)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int i = 0; i < 3; ++i) {
        h ^= (uint64_t)(uint32_t)k->v[i] + 0x9e3779b97f4a7c15ull + (h << 6) +
             (h >> 2);
        h *= 0xbf58476d1ce4e5b9ull;
    }
    return h ^ (h >> 31);
}

static face_key sorted_face(const int *cell, int opposite)
{
    face_key f;
    int k = 0;
    for (int i = 0; i < 4; ++i)
        if (i != opposite) f.v[k++] = cell[i];
    for (int a = 0; a < 2; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (f.v[b] < f.v[a]) {
                int t = f.v[a];
                f.v[a] = f.v[b];
                f.v[b] = t;
            }
    return f;
}

static void mark_constrained(int degree, const int *bary, int n,
                             const int *g, int opposite, int *constrained,
                             const double *v, const double *nodes,
                             double *node_points)
{
    for (int d = 0; d < n; ++d) {
        if (bary[4 * d + opposite] != 0) continue;
        constrained[g[d] - 1] = 1;
        if (node_points) {
            double jac[9];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    jac[r + 3 * c] = v[3 * (c + 1) + r] - v[r];
            const double *nd = nodes + 3 * d;
            for (int r = 0; r < 3; ++r)
                node_points[3 * (g[d] - 1) + r] =
                    v[r] + (jac[r] * nd[0] + jac[r + 3] * nd[1] +
                            jac[r + 6] * nd[2]);
        }
    }
    (void)degree;
}

int ffo_lagrange_dirichlet(int nvert, const double *vertices, int ntet,
                           const int *tets, int degree,
                           const int *global_dofs, int global_count,
                           int *constrained, double *node_points)
{
    int n = ffo_lagrange_dof_count(degree);
    if (degree < 1 || degree > 4) return 1;
    int *bary = malloc(sizeof(int) * 4 * n);
    double *nodes = malloc(sizeof(double) * 3 * n);
    ffo_lagrange_init(degree, bary, nodes);
    for (int i = 0; i < global_count; ++i) constrained[i] = 0;
    size_t cap = 1;
    while (cap < (size_t)ntet * 8) cap <<= 1;
    face_key *keys = malloc(sizeof(face_key) * cap);
    int *counts = malloc(sizeof(int) * cap);
    for (size_t i = 0; i < cap; ++i) counts[i] = 0;
    for (int e = 0; e < ntet; ++e)
        for (int o = 0; o < 4; ++o) {
            face_key f = sorted_face(tets + 4 * e, o);
            size_t s = hash_face(&f) & (cap - 1);
            while (counts[s] && memcmp(&keys[s], &f, sizeof f) != 0)
                s = (s + 1) & (cap - 1);
            keys[s] = f;
            counts[s]++;
        }
    for (int e = 0; e < ntet; ++e) {
        double v[12];
        for (int k = 0; k < 4; ++k) {
            int id = tets[4 * e + k];
            if (id < 1 || id > nvert) {
                free(keys); free(counts); free(bary); free(nodes);
                return 1;
            }
            for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
        }
        for (int o = 0; o < 4; ++o) {
            face_key f = sorted_face(tets + 4 * e, o);
            size_t s = hash_face(&f) & (cap - 1);
            while (memcmp(&keys[s], &f, sizeof f) != 0) s = (s + 1) & (cap - 1);
            if (counts[s] != 1) continue;
            mark_constrained(degree, bary, n, global_dofs + (size_t)n * e, o,
                             constrained, v, nodes, node_points);
        }
    }
    free(keys);
    free(counts);
    free(bary);
    free(nodes);
    return 0;
}

/* Literal O(T^2) restatement of the incidence scan (:587-597) for tiny
 * meshes: incidence = number of cells containing all three face vertices. */
int ffo_lagrange_dirichlet_search(int ntet, const int *tets, int degree,
                                  const int *global_dofs, int global_count,
                                  int *constrained)
{
    int n = ffo_lagrange_dof_count(degree);
    if (degree < 1 || degree > 4) return 1;
    int *bary = malloc(sizeof(int) * 4 * n);
    double *nodes = malloc(sizeof(double) * 3 * n);
    ffo_lagrange_init(degree, bary, nodes);
    for (int i = 0; i < global_count; ++i) constrained[i] = 0;
    for (int e = 0; e < ntet; ++e)
        for (int o = 0; o < 4; ++o) {
            int face[3], k = 0;
            for (int i = 0; i < 4; ++i)
                if (i != o) face[k++] = tets[4 * e + i];
            int incidence = 0;
            for (int other = 0; other < ntet; ++other) {
                int all = 1;
                for (int a = 0; a < 3; ++a) {
                    int any = 0;
                    for (int b = 0; b < 4; ++b)
                        if (tets[4 * other + b] == face[a]) any = 1;
                    all = all && any;
                }
                if (all) ++incidence;
            }
            if (incidence != 1) continue;
            mark_constrained(degree, bary, n, global_dofs + (size_t)n * e, o,
                             constrained, NULL, nodes, NULL);
        }
    free(bary);
    free(nodes);
    return 0;
}

/* The triplet stage of ffo_lagrange_stiffness_triplets on nthreads threads
 * (OpenMP over cells, each thread with its own element scratch): the same
 * numbers in the same positions.  For timing the CPU baseline on all host
 * cores; the sequential routine above is the one the tests pin. */
int ffo_par_lagrange_stiffness_triplets(int nthreads, int nvert, const double *vertices,
                                        int ntet, const int *tets, int degree, int qdeg,
                                        const double *cell_kappa, double kappa, double mu,
                                        const int *global_dofs, int *rows, int *cols,
                                        double *vals)
{
    lagrange_table t;
    if (table_init(&t, degree, qdeg)) return 1;
    const int n = t.n;
    int st = 0;
#pragma omp parallel num_threads(nthreads) reduction(| : st)
    {
        double *m = malloc(sizeof(double) * n * n);
        double *pg = malloc(sizeof(double) * 3 * n);
#pragma omp for schedule(static)
        for (int e = 0; e < ntet; ++e) {
            double v[12];
            int bad = 0;
            for (int k = 0; k < 4; ++k) {
                int id = tets[4 * e + k];
                if (id < 1 || id > nvert) { bad = 1; break; }
                for (int c = 0; c < 3; ++c) v[3 * k + c] = vertices[3 * (id - 1) + c];
            }
            double ke = cell_kappa ? cell_kappa[e] : kappa;
            if (bad || element_with_table(&t, v, ke, mu, m, pg)) {
                st |= 1;
                continue;
            }
            const int *g = global_dofs + (size_t)n * e;
            size_t entry = (size_t)n * n * e;
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r, ++entry) {
                    rows[entry] = g[r];
                    cols[entry] = g[c];
                    vals[entry] = m[r + n * c];
                }
        }
        free(m);
        free(pg);
    }
    table_free(&t);
    return st;
}
