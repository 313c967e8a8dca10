// ffh_problem.hip -- problems resident in HBM: symbolic stage, element
// kernels with the coloured atomic-free CSC scatter, loads, Dirichlet
// reduction and the constrained solve.
//
// Reference: src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90
// (element :40-105, CSC :276-334, load :473-528 / :705-774),
// src/solvers/fortfem_tetra_lagrange_solver_3d.f90 (:64-129, boundary
// :551-610), src/solvers/fortfem_sparse_direct.f90 (:264-344).
//
// Element kernels use the reference-tensor form of the element integral:
// for affine cells the reference's quadrature sum
//   K_rc = sum_q det w_q kappa (J^-T g_r(q)).(J^-T g_c(q))
// equals kappa det sum_ab (J^-1 J^-T)_ab S^ab_rc with
// S^ab_rc = sum_q w_q d_a phi_r(q) d_b phi_c(q), tabulated once in FP64 with
// the same Duffy rule.  Only the six distinct entries of the symmetric
// G = J^-1 J^-T enter, so an element costs 6 n(n+1)/2 multiply-adds instead of
// the reference's n_q n^2 (3x3 matvecs).  Jacobian, determinant and inverse
// are always FP64.  The upper triangle is computed once and written to both
// (r,c) and (c,r) slots, so the assembled matrix is exactly symmetric and
// products can read the CSC arrays as CSR.
#include <algorithm>
#include <cfloat>
#include <complex>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "ffh_internal.hpp"
#include "ffh_krylov.hpp"

namespace ffh {

constexpr int kBlockA = 128;

// ---------------------------------------------------------------- kernels --
__device__ __forceinline__ bool tet_geometry(const double *__restrict__ verts, int v0,
                                             int v1, int v2, int v3, double jac[9],
                                             double inv[9], double &det)
{
    const int id[4] = {v0, v1, v2, v3};
    double x0 = verts[3 * (size_t)v0], y0 = verts[3 * (size_t)v0 + 1],
           z0 = verts[3 * (size_t)v0 + 2];
    for (int c = 0; c < 3; ++c) {
        const double *v = verts + 3 * (size_t)id[c + 1];
        jac[0 + 3 * c] = v[0] - x0;
        jac[1 + 3 * c] = v[1] - y0;
        jac[2 + 3 * c] = v[2] - z0;
    }
#define J(r, c) jac[(r) + 3 * (c)]
    det = J(0, 0) * (J(1, 1) * J(2, 2) - J(1, 2) * J(2, 1)) -
          J(0, 1) * (J(1, 0) * J(2, 2) - J(1, 2) * J(2, 0)) +
          J(0, 2) * (J(1, 0) * J(2, 1) - J(1, 1) * J(2, 0));
    double mx = 0.0;
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(jac[i]));
    double s = fmax(1.0, mx);
    // degenerate guard of the reference element (:65-66)
    if (!(det > 64.0 * 2.220446049250313e-16 * s * s * s)) return false;
    double ri = 1.0 / det;
    inv[0 + 3 * 0] = (J(1, 1) * J(2, 2) - J(1, 2) * J(2, 1)) * ri;
    inv[0 + 3 * 1] = (J(0, 2) * J(2, 1) - J(0, 1) * J(2, 2)) * ri;
    inv[0 + 3 * 2] = (J(0, 1) * J(1, 2) - J(0, 2) * J(1, 1)) * ri;
    inv[1 + 3 * 0] = (J(1, 2) * J(2, 0) - J(1, 0) * J(2, 2)) * ri;
    inv[1 + 3 * 1] = (J(0, 0) * J(2, 2) - J(0, 2) * J(2, 0)) * ri;
    inv[1 + 3 * 2] = (J(0, 2) * J(1, 0) - J(0, 0) * J(1, 2)) * ri;
    inv[2 + 3 * 0] = (J(1, 0) * J(2, 1) - J(1, 1) * J(2, 0)) * ri;
    inv[2 + 3 * 1] = (J(0, 1) * J(2, 0) - J(0, 0) * J(2, 1)) * ri;
    inv[2 + 3 * 2] = (J(0, 0) * J(1, 1) - J(0, 1) * J(1, 0)) * ri;
#undef J
    return true;
}

// G = J^-1 J^-T (symmetric) as g[0..5] = G00, G11, G22, G01, G02, G12
// (inv column-major).
__device__ __forceinline__ void h1_gterms(const double inv[9], double g[6])
{
    const int a_[6] = {0, 1, 2, 0, 0, 1}, b_[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int a = a_[k], b = b_[k];
        g[k] = inv[a + 0] * inv[b + 0] + inv[a + 3] * inv[b + 3] + inv[a + 6] * inv[b + 6];
    }
}

// Upper triangle of sum_k coef[k] S_k (+ cm M when mass) into E[u][pos].
template <int NL, typename CT>
__device__ __forceinline__ void h1_emit(const CT coef[6], CT cm, bool mass,
                                        const CT *__restrict__ tab, CT *__restrict__ E,
                                        size_t T, int pos)
{
    constexpr int NN = NL * NL;
    int u = 0;
    for (int r = 0; r < NL; ++r) {
        for (int cc = r; cc < NL; ++cc, ++u) {
            const int rc = r * NL + cc;
            CT s = coef[0] * tab[rc] + coef[1] * tab[NN + rc] + coef[2] * tab[2 * NN + rc] +
                   coef[3] * tab[3 * NN + rc] + coef[4] * tab[4 * NN + rc] +
                   coef[5] * tab[5 * NN + rc];
            if (mass) s += cm * tab[6 * NN + rc];
            E[(size_t)u * T + pos] = s;
        }
    }
}

// Phase 1 of the assembly: one thread per cell (all cells, colour order),
// the upper triangle of the element matrix into E[u][pos] (u = packed upper
// index, lanes consecutive in pos).  tab = 7 tables of NL x NL:
// S00, S11, S22, S01+S10, S02+S20, S12+S21, mass.
template <int NL, typename CT>
__global__ __launch_bounds__(kBlockA) void k_h1_elem(
    int ncell, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const int *__restrict__ perm, const double *__restrict__ kappa_cell, double kappa,
    double mu, const CT *__restrict__ tab, CT *__restrict__ E, int *__restrict__ err)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    double g[6];
    h1_gterms(inv, g);
    const double ke = kappa_cell ? kappa_cell[perm[pos]] : kappa;
    const CT kd = (CT)(ke * det);
    CT coef[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) coef[k] = kd * (CT)g[k];
    h1_emit<NL, CT>(coef, (CT)(mu * det), mu != 0.0, tab, E, (size_t)ncell, pos);
}

// Tangent of the H1 element matrices (the derivative of the element
// integral of assembly_tetra_lagrange_arbitrary_order_3d.f90:107-174) in the
// same reference-tensor form: with c_k = kappa det G_k and c_m = mu det,
//   dc_k = dkappa det G_k + kappa ddet G_k + kappa det dG_k,
//   dc_m = dmu det + mu ddet,
// ddet = det tr(J^-1 dJ), dJ^-1 = -J^-1 dJ J^-1, dG = dJ^-1 J^-T + J^-1 dJ^-T.
// verts_dot (3 per vertex) and the per-cell kappa tangents may be null.
template <int NL>
__global__ __launch_bounds__(kBlockA) void k_h1_elem_jvp(
    int ncell, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const double *__restrict__ verts_dot, const int *__restrict__ perm,
    const double *__restrict__ kappa_cell, const double *__restrict__ kappa_cell_dot,
    double kappa, double kappa_dot, double mu, double mu_dot, const double *__restrict__ tab,
    double *__restrict__ E, int *__restrict__ err)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    double jd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (verts_dot) {
        const int id[4] = {c.x, c.y, c.z, c.w};
        const double *v0 = verts_dot + 3 * (size_t)c.x;
        for (int cc = 0; cc < 3; ++cc) {
            const double *v = verts_dot + 3 * (size_t)id[cc + 1];
            for (int r = 0; r < 3; ++r) jd[r + 3 * cc] = v[r] - v0[r];
        }
    }
    // m = J^-1 dJ; ddet = det tr(m); dinv = -m J^-1
    double m[9], invd[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            m[i + 3 * j] = inv[i] * jd[3 * j] + inv[i + 3] * jd[1 + 3 * j] +
                           inv[i + 6] * jd[2 + 3 * j];
    const double detd = det * (m[0] + m[4] + m[8]);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            invd[i + 3 * j] = -(m[i] * inv[3 * j] + m[i + 3] * inv[1 + 3 * j] +
                                m[i + 6] * inv[2 + 3 * j]);
    double g[6], gd[6];
    h1_gterms(inv, g);
    {
        const int a_[6] = {0, 1, 2, 0, 0, 1}, b_[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int a = a_[k], b = b_[k];
            double s = 0.0;
            for (int q = 0; q < 3; ++q)
                s += invd[a + 3 * q] * inv[b + 3 * q] + inv[a + 3 * q] * invd[b + 3 * q];
            gd[k] = s;
        }
    }
    const int cell = perm[pos];
    const double ke = kappa_cell ? kappa_cell[cell] : kappa;
    const double ked = kappa_cell_dot ? kappa_cell_dot[cell] : kappa_dot;
    double coef[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) coef[k] = ked * det * g[k] + ke * detd * g[k] + ke * det * gd[k];
    h1_emit<NL, double>(coef, mu_dot * det + mu * detd, true, tab, E, (size_t)ncell, pos);
}

// Cotangent of the H1 element matrices (assembly_tetra_lagrange_arbitrary_
// order_3d.f90:176-275, 388-471), one colour per launch.  The element
// cotangent is gathered from the CSC cotangent through the entry slots
// (csc_value_bar_at), contracted with the 7 tables, t_k = sum_rc Eb_rc S_k,rc,
// and taken back through c_k = kappa det G_k, c_m = mu det:
//   kappa_bar = det sum_k t_k G_k, mu_bar = det t_m,
//   det_bar = kappa sum_k t_k G_k + mu t_m, G_bar_k = kappa det t_k,
//   Inv_bar from G = Inv Inv^T, J_bar = -Inv^T Inv_bar Inv^T + det_bar det Inv^T,
// then to the four vertices (tetra_jacobian_vjp).  Cells of one colour
// share no vertex (every vertex carries a dof), so the vertex sums need no
// atomics.  Per-cell kappa / mu bars go to kb / mb [pos].
template <int NL>
__global__ __launch_bounds__(kBlockA) void k_h1_elem_vjp(
    int pos0, int cnt, int ncell, const int *__restrict__ cells_c,
    const double *__restrict__ verts, const int *__restrict__ perm,
    const double *__restrict__ kappa_cell, double kappa, double mu,
    const double *__restrict__ tab, const int *__restrict__ slots,
    const double *__restrict__ vbar, double *__restrict__ verts_bar, double *__restrict__ kb,
    double *__restrict__ mb, int *__restrict__ err)
{
    const int i = blockIdx.x * kBlockA + threadIdx.x;
    if (i >= cnt) return;
    const int pos = pos0 + i;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    constexpr int NN = NL * NL;
    const size_t T = (size_t)ncell;
    double t[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int rc = 0; rc < NN; ++rc) {
        const double e = vbar[slots[(size_t)rc * T + pos]];
#pragma unroll
        for (int k = 0; k < 7; ++k) t[k] += e * tab[k * NN + rc];
    }
    double g[6];
    h1_gterms(inv, g);
    const double ke = kappa_cell ? kappa_cell[perm[pos]] : kappa;
    double tg = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) tg += t[k] * g[k];
    kb[pos] = det * tg;
    mb[pos] = det * t[6];
    const double det_bar = ke * tg + mu * t[6];
    double ib[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    {
        const int a_[6] = {0, 1, 2, 0, 0, 1}, b_[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int a = a_[k], b = b_[k];
            const double gb = ke * det * t[k];
            for (int q = 0; q < 3; ++q) {
                ib[a + 3 * q] += gb * inv[b + 3 * q];
                ib[b + 3 * q] += gb * inv[a + 3 * q];
            }
        }
    }
    // J_bar = -Inv^T ib Inv^T + det_bar det Inv^T
    double w[9], jb[9];
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            w[r + 3 * j] = inv[3 * r] * ib[3 * j] + inv[1 + 3 * r] * ib[1 + 3 * j] +
                           inv[2 + 3 * r] * ib[2 + 3 * j];
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            jb[r + 3 * j] = -(w[r] * inv[j] + w[r + 3] * inv[j + 3] + w[r + 6] * inv[j + 6]) +
                            det_bar * det * inv[j + 3 * r];
    const int id[4] = {c.x, c.y, c.z, c.w};
    for (int r = 0; r < 3; ++r) {
        verts_bar[3 * (size_t)id[0] + r] -= jb[r] + jb[r + 3] + jb[r + 6];
        for (int q = 0; q < 3; ++q) verts_bar[3 * (size_t)id[q + 1] + r] += jb[r + 3 * q];
    }
}

// ------------------------------------------------- gather assembly (H1) --
// The per-cell reference-tensor coefficients of the element matrix,
// c_k = kappa det G_k (k = 0..5, the G pairs of h1_gterms) and c_6 = mu det,
// in cell order, 8 values per cell (c_7 unused).
// One thread per cell in the mesh's own cell order (cells_e: 0-based
// vertex ids), so the coefficient reads and the term writes are coalesced.
template <typename CT>
__global__ __launch_bounds__(kBlockA) void k_h1_terms(int ncell, const int *__restrict__ cells_e,
                                                      const double *__restrict__ verts,
                                                      const double *__restrict__ kappa_cell,
                                                      double kappa, double mu,
                                                      CT *__restrict__ terms,
                                                      int *__restrict__ err)
{
    const int e = blockIdx.x * kBlockA + threadIdx.x;
    if (e >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_e)[e];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    double g[6];
    h1_gterms(inv, g);
    const double ke = kappa_cell ? kappa_cell[e] : kappa;
    const double kd = ke * det;
    CT *t = terms + (size_t)8 * e;
#pragma unroll
    for (int k = 0; k < 6; ++k) t[k] = (CT)(kd * g[k]);
    t[6] = (CT)(mu * det);
    t[7] = CT(0);
}

// Largest value run of one gather block (values of CT in LDS) and most
// columns per block (one thread per column, 64-column slices per wave).
constexpr int kGatherNz = 8192;
constexpr int kGatherCols = 256;

// Assembly by columns: a block takes a run of CSC columns whose values fit
// LDS; one lane per column walks the column's occurrences (the cells that
// hold the column's dof, ascending) and adds, for each local row index l_r,
// the element entry sum_k c_k S_k(l_r, l_c) (+ c_6 M) at its value position;
// the block then writes the run's values once.  No colours, no atomics, no
// element matrices in HBM.  The occurrences and value positions are stored
// step-major per 64-column slice, so each step of a wave reads 256 + 512 Q
// contiguous bytes.  Entries (i, j) and (j, i) add the same numbers (the
// upper-triangle table entry) over the same cells in the same order, so the
// matrix is exactly symmetric and the result does not depend on the launch.
template <int NL, typename CT>
__global__ __launch_bounds__(kGatherCols) void k_h1_gather(
    const int *__restrict__ blk, const int *__restrict__ bslice,
    const int *__restrict__ soff, const int *__restrict__ col_ptr,
    const int *__restrict__ occ, const ushort4 *__restrict__ gslot,
    const CT *__restrict__ terms, const CT *__restrict__ tab, int mass,
    CT *__restrict__ val)
{
    constexpr int NN = NL * NL;
    constexpr int Q = (NL + 3) / 4;
    extern __shared__ __align__(16) unsigned char gather_lds[];
    CT *acc = reinterpret_cast<CT *>(gather_lds);
    __shared__ CT st[7 * NN];
    const int b = xcd_block();
    const int base = col_ptr[blk[b]], nz = col_ptr[blk[b + 1]] - base;
    for (int i = threadIdx.x; i < nz; i += kGatherCols) acc[i] = CT(0);
    for (int i = threadIdx.x; i < 7 * NN; i += kGatherCols) st[i] = tab[i];
    __syncthreads();
    const int s = bslice[b] + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s < bslice[b + 1]) {
        const int k1 = soff[s + 1];
        for (int k = soff[s]; k < k1; ++k) {
            const int o = occ[(size_t)k * 64 + lane];
            if (o < 0) continue;
            const int e = o / NL, lc = o - e * NL;
            const CT *t = terms + (size_t)8 * e;
            const CT t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3], t4 = t[4], t5 = t[5],
                     t6 = t[6];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const ushort4 g = gslot[((size_t)k * Q + q) * 64 + lane];
                const unsigned short gl[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int lr = 4 * q + i;
                    if (lr >= NL) break;
                    const int rc = lr < lc ? lr * NL + lc : lc * NL + lr;
                    CT v = t0 * st[rc] + t1 * st[NN + rc] + t2 * st[2 * NN + rc] +
                           t3 * st[3 * NN + rc] + t4 * st[4 * NN + rc] + t5 * st[5 * NN + rc];
                    if (mass) v += t6 * st[6 * NN + rc];
                    acc[gl[i]] += v;
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nz; i += kGatherCols) val[base + i] = acc[i];
}

// Value positions of the gather layout, one thread per (slice, lane), from
// the element slots ([r n + c][pos], absolute) less the block's first value.
__global__ __launch_bounds__(256) void k_h1_gather_slots(int nslice, int ncell, int nl,
                                                         const int *__restrict__ soff,
                                                         const int *__restrict__ sbase,
                                                         const int *__restrict__ occ,
                                                         const int *__restrict__ invperm,
                                                         const int *__restrict__ slots,
                                                         uint16_t *__restrict__ gslot)
{
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (int64_t)nslice * 64) return;
    const int s = (int)(t >> 6), lane = (int)(t & 63);
    const int q4 = 4 * ((nl + 3) / 4);
    for (int k = soff[s]; k < soff[s + 1]; ++k) {
        const int o = occ[(size_t)k * 64 + lane];
        uint16_t *g = gslot + (size_t)k * q4 * 64;
        if (o < 0) {
            for (int lr = 0; lr < q4; ++lr) g[((lr >> 2) * 64 + lane) * 4 + (lr & 3)] = 0;
            continue;
        }
        const int e = o / nl, lc = o - e * nl;
        const int pos = invperm[e];
        for (int lr = 0; lr < q4; ++lr)
            g[((lr >> 2) * 64 + lane) * 4 + (lr & 3)] =
                lr < nl ? (uint16_t)(slots[(size_t)(lr * nl + lc) * ncell + pos] - sbase[s]) : 0;
    }
}

__global__ void k_invert_perm(int n, const int *__restrict__ perm, int *__restrict__ inv)
{
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos < n) inv[perm[pos]] = pos;
}

// Phase 2: scatter-add of one colour, one thread per (cell, entry).  Cells of
// a colour share no dof, so every value slot has a single writer.  The upper
// entry u(r,c) goes to both (r,c) and (c,r), so the matrix is exactly
// symmetric.
template <typename CT, typename OT>
__global__ __launch_bounds__(256) void k_scatter_colour(int pos0, int cnt, int ncell, int nl,
                                                         const CT *__restrict__ E,
                                                         const int *__restrict__ slots,
                                                         OT *__restrict__ val)
{
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (int64_t)cnt * nl * nl) return;
    const int rc = (int)(t / cnt);
    const int pos = pos0 + (int)(t - (int64_t)rc * cnt);
    int r = rc / nl, c = rc - r * nl;
    if (r > c) {
        const int tmp = r;
        r = c;
        c = tmp;
    }
    const int u = r * nl - (r * (r - 1)) / 2 + (c - r);
    const size_t T = (size_t)ncell;
    val[slots[(size_t)rc * T + pos]] += (OT)E[(size_t)u * T + pos];
}

// Load phase 2: scatter-add of the element vectors of one colour.
__global__ __launch_bounds__(256) void k_scatter_vec_colour(int pos0, int cnt, int ncell,
                                                             int nl,
                                                             const double *__restrict__ EL,
                                                             const int *__restrict__ gd,
                                                             double *__restrict__ rhs)
{
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (int64_t)cnt * nl) return;
    const int l = (int)(t / cnt);
    const int pos = pos0 + (int)(t - (int64_t)l * cnt);
    const size_t T = (size_t)ncell;
    rhs[gd[(size_t)l * T + pos]] += EL[(size_t)l * T + pos];
}

// 2D P1 (assemble_scalar_p1_matrix, src/fortfem_forms_simple.f90:1678-1728):
// local = kappa area G^T G + mu area (1 + I)/12.
__global__ __launch_bounds__(kBlockA) void k_p1_2d_cells(
    int pos0, int pos1, int ncell, const int *__restrict__ cells_c,
    const double *__restrict__ verts, const int *__restrict__ perm,
    const double *__restrict__ kappa_cell, double kappa, double mu,
    const int *__restrict__ slots, double *__restrict__ val, int *__restrict__ err)
{
    const int pos = pos0 + blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= pos1) return;
    const int *c = cells_c + 3 * (size_t)pos;
    double x[3], y[3];
    for (int k = 0; k < 3; ++k) {
        x[k] = verts[2 * (size_t)c[k]];
        y[k] = verts[2 * (size_t)c[k] + 1];
    }
    const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
    if (!(fabs(det) > 64.0 * 2.220446049250313e-16)) {
        atomicOr(err, 1);
        return;
    }
    const double area = 0.5 * fabs(det);
    const double gx[3] = {(y[1] - y[2]) / det, (y[2] - y[0]) / det, (y[0] - y[1]) / det};
    const double gy[3] = {(x[2] - x[1]) / det, (x[0] - x[2]) / det, (x[1] - x[0]) / det};
    const double ke = kappa_cell ? kappa_cell[perm[pos]] : kappa;
    const size_t T = (size_t)ncell;
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) {
            double v = ke * area * (gx[a] * gx[b] + gy[a] * gy[b]) +
                       mu * area * (a == b ? 2.0 : 1.0) / 12.0;
            val[slots[(size_t)(a * 3 + b) * T + pos]] += v;
            if (a != b) val[slots[(size_t)(b * 3 + a) * T + pos]] += v;
        }
}

// Nedelec terms of one cell (colour order): the 12 coefficients of the
// reference tensors.  Curl part J^T K J / det with K = alpha I (scalar or per
// cell) or the cell's tensor; value part beta det J^-1 J^-T, or
// det J^-1 M J^-T with the cell's value tensor M.  Terms follow the table
// order (00, 11, 22, 01, 02, 12), curl then value.
__global__ __launch_bounds__(kBlockA) void k_ned_terms(
    int ncell, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const int *__restrict__ perm, const double *__restrict__ alpha_cell,
    const double *__restrict__ tensor, double alpha, double beta,
    double *__restrict__ gterm, int *__restrict__ err,
    const double *__restrict__ mtensor = nullptr)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    const int e = perm[pos];
    double K[9];
    if (tensor) {
        for (int i = 0; i < 9; ++i) K[i] = tensor[9 * (size_t)e + i];
    } else {
        const double a = alpha_cell ? alpha_cell[e] : alpha;
        for (int i = 0; i < 9; ++i) K[i] = (i % 4 == 0) ? a : 0.0;
    }
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    const size_t T = (size_t)ncell;
    const double rdet = 1.0 / det;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int a = pa[s], b = pb[s];
        // (J^T K J)_ab = sum_tu J_ta K_tu J_ub (column-major J, K)
        double cm = 0.0;
        for (int t = 0; t < 3; ++t) {
            double kj = K[t + 0] * jac[0 + 3 * b] + K[t + 3] * jac[1 + 3 * b] + K[t + 6] * jac[2 + 3 * b];
            cm += jac[t + 3 * a] * kj;
        }
        double mm;
        if (mtensor) {
            // (J^-1 M J^-T)_ab with the cell's symmetric value tensor M
            const double *M = mtensor + 9 * (size_t)e;
            mm = 0.0;
            for (int t = 0; t < 3; ++t) {
                const double mi = M[t + 0] * inv[b + 0] + M[t + 3] * inv[b + 3] + M[t + 6] * inv[b + 6];
                mm += inv[a + 3 * t] * mi;
            }
        } else {
            mm = beta * (inv[a + 0] * inv[b + 0] + inv[a + 3] * inv[b + 3] + inv[a + 6] * inv[b + 6]);
        }
        gterm[(size_t)s * T + pos] = cm * rdet;
        gterm[(size_t)(6 + s) * T + pos] = det * mm;
    }
}

// Nedelec assembly of one colour: lanes run over element entries
// t = c * n + r of a chunk (rows of one element column side by side), each
// thread walks kNedCells consecutive cells of the colour (sorted by orientation
// class, so the 12 table entries of (min(r,c), max(r,c)) stay in registers
// across cells) and stores (first contribution, negative slot) or adds the
// entry to its value slot.  Cells of a colour share no dof: every slot has one
// writer.  Both triangles come from the same packed
// table entry, so the matrix is exactly symmetric.
constexpr int kNedCells = 16;
template <typename CT, typename OT>
__global__ __launch_bounds__(256) void k_ned_assemble(
    int pos0, int cnt, int ncell, int n, const unsigned char *__restrict__ cls_c,
    const double *__restrict__ gterm, const CT *__restrict__ tab,
    const int *__restrict__ slots_cm, OT *__restrict__ val)
{
    const int nn = n * n;
    const int t = blockIdx.y * 256 + threadIdx.x;
    if (t >= nn) return;
    const int c = t / n, r = t - c * n;
    const int lo = min(r, c), hi = max(r, c);
    const int u = lo * n - (lo * (lo - 1)) / 2 + (hi - lo);
    const int np = n * (n + 1) / 2;
    const int p0 = pos0 + blockIdx.x * kNedCells;
    const int p1 = min(pos0 + cnt, p0 + kNedCells);
    const size_t T = (size_t)ncell;
    int cur = -1;
    CT rt[12];
    for (int pos = p0; pos < p1; ++pos) {
        const int cl = cls_c[pos];
        if (cl != cur) {
            cur = cl;
#pragma unroll
            for (int s = 0; s < 12; ++s) rt[s] = tab[((size_t)cl * 12 + s) * np + u];
        }
        CT e = CT(0);
#pragma unroll
        for (int s = 0; s < 12; ++s) e += (CT)gterm[(size_t)s * T + pos] * rt[s];
        const int sl = slots_cm[(size_t)pos * nn + t];
        if (sl < 0)
            val[~sl] = (OT)e; // first contribution to this slot
        else
            val[sl] += (OT)e;
    }
}

// ------------------------------------------- per-point weighted Nedelec --
// The curl part of assemble_tetra_nedelec_weighted_element(_order)
// (assembly_tetra_nedelec_3d.f90:2406-2514) with its tensor evaluated at
// every point x_q of the rule: with physical curls J c^ / det, in the class's
// oriented basis the element curl matrix is the dense product
//   E = C^T blockdiag(M_q) C,  M_q = w_q J^T K(x_q) J / det,
// C[(q, a), j] the oriented reference curls (3 rows per point).  One
// workgroup per cell; each wave owns up to SPW regions of 32 x 32 entries
// (2 x 2 MFMA tiles of 16 x 16) of E's upper triangle.  The rule is walked in
// chunks of PC points: their rows of C and of D = blockdiag(M) C are staged
// in LDS and every 4 rows are one MFMA per tile (FP32 or FP64 accumulation).
// E's upper entries are added through the cell's slots to both (r, c) and
// (c, r) from the same number (the mass part and the first-touch stores come
// before, from the class-table kernel), so the matrix stays exactly
// symmetric.  A colour's cells share no dof, so the adds do not race.
template <typename CT>
struct MfmaTile;
template <>
struct MfmaTile<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // row of result register r in lane l (column l % 16)
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <>
struct MfmaTile<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};

template <typename CT>
constexpr int kPointCurlPad = sizeof(CT) == 4 ? 16 : 0;

template <typename CT, int W, int SPW, int PC>
__global__ __launch_bounds__(64 * W) void k_ned_point_curl(
    int pos0, int ncell, int n, int ldn, int nq, const int *__restrict__ cells_c,
    const double *__restrict__ verts, const int *__restrict__ perm,
    const unsigned char *__restrict__ cls_c, const double *__restrict__ ptens,
    const double *__restrict__ pts, const CT *__restrict__ ctab,
    const int *__restrict__ slots_cm, CT *__restrict__ val, int *__restrict__ err)
{
    using MT = MfmaTile<CT>;
    using v4 = typename MT::v4;
    constexpr int NT = 64 * W, ROWS = 3 * PC;
    static_assert(ROWS % 4 == 0 && ROWS < 128, "chunk rows must fill MFMA steps");
    // LDS row stride: FP32 rows are padded by 16 words so the four rows an
    // MFMA operand read touches fall on two halves of the banks
    const int ls = ldn + kPointCurlPad<CT>;
    extern __shared__ __align__(16) unsigned char pc_lds[];
    CT *sC = reinterpret_cast<CT *>(pc_lds); // [ROWS][ls]
    CT *sD = sC + (size_t)ROWS * ls;         // [ROWS][ls]
    __shared__ double sJ[9], sdet;
    __shared__ int sok;
    __shared__ CT sM[PC][9];
    const int pos = pos0 + blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) {
        const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
        double jac[9], inv[9], det;
        sok = tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det) ? 1 : 0;
        if (!sok) atomicOr(err, 1);
        for (int i = 0; i < 9; ++i) sJ[i] = jac[i];
        sdet = det;
    }
    __syncthreads();
    if (!sok) return;
    const int e = perm[pos];
    const CT *C = ctab + (size_t)cls_c[pos] * nq * 3 * ldn;
    // this wave's 32 x 32 regions (bi <= bj), in row-major upper order
    const int ns = ldn >> 5;
    int rb[SPW], cb[SPW];
#pragma unroll
    for (int t = 0; t < SPW; ++t) {
        int sidx = wv + W * t, bi = 0;
        rb[t] = cb[t] = -1;
        while (bi < ns && sidx >= ns - bi) {
            sidx -= ns - bi;
            ++bi;
        }
        if (bi < ns) {
            rb[t] = 32 * bi;
            cb[t] = 32 * (bi + sidx);
        }
    }
    v4 acc[SPW][2][2];
#pragma unroll
    for (int t = 0; t < SPW; ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[t][a][b] = v4{CT(0), CT(0), CT(0), CT(0)};
    for (int q0 = 0; q0 < nq; q0 += PC) {
        const int nqc = min(PC, nq - q0);
        // M_q = w_q J^T K_q J / det (K column-major), in FP64
        if (tid < nqc) {
            const double *K = ptens + ((size_t)e * nq + q0 + tid) * 9;
            const double w = pts[4 * (q0 + tid) + 3] / sdet;
            double KJ[9]; // K J, column-major
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    KJ[i + 3 * j] = K[i] * sJ[3 * j] + K[i + 3] * sJ[1 + 3 * j] +
                                    K[i + 6] * sJ[2 + 3 * j];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    sM[tid][i + 3 * j] = (CT)(w * (sJ[3 * i] * KJ[3 * j] + sJ[1 + 3 * i] * KJ[1 + 3 * j] +
                                                   sJ[2 + 3 * i] * KJ[2 + 3 * j]));
        }
        // (row, column) of element i = r ldn + j advance without divisions:
        // i += NT moves by NT / ldn rows and NT % ldn columns
        const int dr = NT / ldn, dj = NT - dr * ldn;
        // rows of C for the chunk, contiguous in the table (zero past the rule)
        {
            const CT *src = C + (size_t)3 * q0 * ldn;
            const int live = 3 * nqc * ldn;
            int r = tid / ldn, j = tid - r * ldn;
            for (int i = tid; i < ROWS * ldn; i += NT) {
                sC[r * ls + j] = i < live ? src[i] : CT(0);
                r += dr;
                j += dj;
                if (j >= ldn) {
                    j -= ldn;
                    ++r;
                }
            }
        }
        __syncthreads();
        // D = blockdiag(M) C
        {
            int r = tid / ldn, j = tid - r * ldn;
            for (int i = tid; i < ROWS * ldn; i += NT) {
                const int qq = (r * 43) >> 7, a = r - 3 * qq; // r / 3 for r < 128
                CT d = CT(0);
                if (qq < nqc) {
                    const CT *c3 = sC + (size_t)(3 * qq) * ls + j;
                    d = sM[qq][a] * c3[0] + sM[qq][a + 3] * c3[ls] + sM[qq][a + 6] * c3[2 * ls];
                }
                sD[r * ls + j] = d;
                r += dr;
                j += dj;
                if (j >= ldn) {
                    j -= ldn;
                    ++r;
                }
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int k0 = 0; k0 < ROWS; k0 += 4) {
            const int kr = k0 + (lane >> 4), li = lane & 15;
            const CT *ca = sC + (size_t)kr * ls + li;
            const CT *db = sD + (size_t)kr * ls + li;
#pragma unroll
            for (int t = 0; t < SPW; ++t) {
                if (rb[t] < 0) continue;
                const CT a0 = ca[rb[t]], a1 = ca[rb[t] + 16];
                const CT b0 = db[cb[t]], b1 = db[cb[t] + 16];
                acc[t][0][0] = MT::op(a0, b0, acc[t][0][0]);
                acc[t][0][1] = MT::op(a0, b1, acc[t][0][1]);
                acc[t][1][0] = MT::op(a1, b0, acc[t][1][0]);
                acc[t][1][1] = MT::op(a1, b1, acc[t][1][1]);
            }
        }
        __syncthreads();
    }
    // upper entries into the CSC values, both (r, c) and (c, r)
    const int nn = n * n;
    const int *sl = slots_cm + (size_t)pos * nn;
#pragma unroll
    for (int t = 0; t < SPW; ++t) {
        if (rb[t] < 0) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rb[t] + 16 * a + MT::row(lane, r);
                    const int col = cb[t] + 16 * b + (lane & 15);
                    if (row >= n || col >= n || row > col) continue;
                    const CT v = acc[t][a][b][r];
                    const int s1 = sl[col * n + row];
                    val[s1 < 0 ? ~s1 : s1] += v;
                    if (row != col) {
                        const int s2 = sl[row * n + col];
                        val[s2 < 0 ? ~s2 : s2] += v;
                    }
                }
    }
}

// Tangent of the 12 Nedelec terms of the curl-mass variant
// (assemble_tetra_nedelec_curl_mass_element_jvp,
// assembly_tetra_nedelec_3d.f90:645-716, in the reference-tensor form):
// c_s = alpha H_s / det with H = J^T J and v_s = beta det G_s with
// G = J^-1 J^-T, so
//   dc_s = (dalpha H_s + alpha dH_s) / det - alpha H_s ddet / det^2,
//   dv_s = dbeta det G_s + beta ddet G_s + beta det dG_s,
// dH = dJ^T J + J^T dJ, ddet = det tr(J^-1 dJ), dJ^-1 = -J^-1 dJ J^-1.
__global__ __launch_bounds__(kBlockA) void k_ned_terms_jvp(
    int ncell, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const double *__restrict__ verts_dot, const int *__restrict__ perm,
    const double *__restrict__ alpha_cell, const double *__restrict__ alpha_cell_dot,
    double alpha, double alpha_dot, double beta, double beta_dot, double *__restrict__ gterm,
    int *__restrict__ err)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    double jd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (verts_dot) {
        const int id[4] = {c.x, c.y, c.z, c.w};
        const double *v0 = verts_dot + 3 * (size_t)c.x;
        for (int cc = 0; cc < 3; ++cc) {
            const double *v = verts_dot + 3 * (size_t)id[cc + 1];
            for (int r = 0; r < 3; ++r) jd[r + 3 * cc] = v[r] - v0[r];
        }
    }
    double m[9], invd[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            m[i + 3 * j] = inv[i] * jd[3 * j] + inv[i + 3] * jd[1 + 3 * j] +
                           inv[i + 6] * jd[2 + 3 * j];
    const double detd = det * (m[0] + m[4] + m[8]);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            invd[i + 3 * j] = -(m[i] * inv[3 * j] + m[i + 3] * inv[1 + 3 * j] +
                                m[i + 6] * inv[2 + 3 * j]);
    const int e = perm[pos];
    const double a = alpha_cell ? alpha_cell[e] : alpha;
    const double ad = alpha_cell_dot ? alpha_cell_dot[e] : alpha_dot;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    const size_t T = (size_t)ncell;
    const double rdet = 1.0 / det;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int ia = pa[s], ib = pb[s];
        double H = 0, Hd = 0, G = 0, Gd = 0;
        for (int t = 0; t < 3; ++t) {
            H += jac[t + 3 * ia] * jac[t + 3 * ib];
            Hd += jd[t + 3 * ia] * jac[t + 3 * ib] + jac[t + 3 * ia] * jd[t + 3 * ib];
            G += inv[ia + 3 * t] * inv[ib + 3 * t];
            Gd += invd[ia + 3 * t] * inv[ib + 3 * t] + inv[ia + 3 * t] * invd[ib + 3 * t];
        }
        gterm[(size_t)s * T + pos] = (ad * H + a * Hd) * rdet - a * H * detd * rdet * rdet;
        gterm[(size_t)(6 + s) * T + pos] = beta_dot * det * G + beta * detd * G + beta * det * Gd;
    }
}

// Cotangent of the Nedelec curl-mass matrix
// (assemble_tetra_nedelec_curl_mass_csc_vjp, :2182-2287), one colour per
// launch, one wave per cell: the lanes contract the cell's oriented element
// cotangent (read through the entry slots) with the class's 12 tables,
// t_s = sum_rc Eb_rc Tab_s,rc; lane 0 takes the 12 sums back through
// c_s = alpha H_s / det and v_s = beta det G_s to alpha, beta and the
// vertices.  Cells of one colour may share a vertex (the dofs sit on edges,
// faces and interiors), so the vertex sums are atomic.
// Cotangents of the twelve per-cell terms: gbar[s * T + pos] = the sum over
// the cell's element entries of vbar(slot) times the class table of term s
// (one wave per cell, as k_ned_vjp).
__global__ __launch_bounds__(256) void k_ned_term_bar(int cnt, int ncell, int n,
                                                      const unsigned char *__restrict__ cls_c,
                                                      const double *__restrict__ tab,
                                                      const int *__restrict__ slots_cm,
                                                      const double *__restrict__ vbar,
                                                      double *__restrict__ gbar)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pos = blockIdx.x * 4 + wave;
    if (pos >= cnt) return; // uniform per wave
    const int nn = n * n, np = n * (n + 1) / 2;
    const double *tb = tab + (size_t)cls_c[pos] * 12 * np;
    double acc[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) acc[s] = 0.0;
    for (int t = lane; t < nn; t += 64) {
        int sl = slots_cm[(size_t)pos * nn + t];
        if (sl < 0) sl = ~sl;
        const double e = vbar[sl];
        const int cc = t / n, r = t - cc * n;
        const int lo = min(r, cc), hi = max(r, cc);
        const int u = lo * n - (lo * (lo - 1)) / 2 + (hi - lo);
#pragma unroll
        for (int s = 0; s < 12; ++s) acc[s] += e * tb[(size_t)s * np + u];
    }
#pragma unroll
    for (int s = 0; s < 12; ++s)
        for (int off = 32; off > 0; off >>= 1) acc[s] += __shfl_xor(acc[s], off);
    if (lane == 0)
#pragma unroll
        for (int s = 0; s < 12; ++s) gbar[(size_t)s * ncell + pos] = acc[s];
}

__global__ __launch_bounds__(256) void k_ned_vjp(
    int pos0, int cnt, int ncell, int n, const unsigned char *__restrict__ cls_c,
    const int *__restrict__ cells_c, const double *__restrict__ verts,
    const int *__restrict__ perm, const double *__restrict__ alpha_cell, double alpha,
    double beta, const double *__restrict__ tab, const int *__restrict__ slots_cm,
    const double *__restrict__ vbar, double *__restrict__ cell_vbar, double *__restrict__ ab,
    double *__restrict__ bb, int *__restrict__ err)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wave;
    if (i >= cnt) return; // uniform per wave
    const int pos = pos0 + i;
    const int nn = n * n, np = n * (n + 1) / 2;
    const int cl = cls_c[pos];
    const double *tb = tab + (size_t)cl * 12 * np;
    double acc[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) acc[s] = 0.0;
    for (int t = lane; t < nn; t += 64) {
        int sl = slots_cm[(size_t)pos * nn + t];
        if (sl < 0) sl = ~sl;
        const double e = vbar[sl];
        const int cc = t / n, r = t - cc * n;
        const int lo = min(r, cc), hi = max(r, cc);
        const int u = lo * n - (lo * (lo - 1)) / 2 + (hi - lo);
#pragma unroll
        for (int s = 0; s < 12; ++s) acc[s] += e * tb[(size_t)s * np + u];
    }
#pragma unroll
    for (int s = 0; s < 12; ++s)
        for (int off = 32; off > 0; off >>= 1) acc[s] += __shfl_xor(acc[s], off);
    if (lane != 0) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    const double a = alpha_cell ? alpha_cell[perm[pos]] : alpha;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    const double rdet = 1.0 / det;
    double jb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ibar[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double abar = 0.0, bbar = 0.0, det_bar = 0.0;
    for (int s = 0; s < 6; ++s) {
        const int ia = pa[s], ib = pb[s];
        double H = 0, G = 0;
        for (int t = 0; t < 3; ++t) {
            H += jac[t + 3 * ia] * jac[t + 3 * ib];
            G += inv[ia + 3 * t] * inv[ib + 3 * t];
        }
        abar += acc[s] * H * rdet;
        bbar += acc[6 + s] * det * G;
        det_bar += -acc[s] * a * H * rdet * rdet + acc[6 + s] * beta * G;
        const double hb = a * acc[s] * rdet, gb = beta * det * acc[6 + s];
        for (int t = 0; t < 3; ++t) {
            jb[t + 3 * ia] += hb * jac[t + 3 * ib];
            jb[t + 3 * ib] += hb * jac[t + 3 * ia];
            ibar[ia + 3 * t] += gb * inv[ib + 3 * t];
            ibar[ib + 3 * t] += gb * inv[ia + 3 * t];
        }
    }
    ab[pos] = abar;
    bb[pos] = bbar;
    double w[9];
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            w[r + 3 * j] = inv[3 * r] * ibar[3 * j] + inv[1 + 3 * r] * ibar[1 + 3 * j] +
                           inv[2 + 3 * r] * ibar[2 + 3 * j];
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r)
            jb[r + 3 * j] += -(w[r] * inv[j] + w[r + 3] * inv[j + 3] + w[r + 6] * inv[j + 6]) +
                             det_bar * det * inv[j + 3 * r];
    // the cell's four vertex cotangents (summed per vertex in cell order by
    // the host, so the result does not depend on the colouring or timing)
    double *o = cell_vbar + (size_t)pos * 12;
    for (int r = 0; r < 3; ++r) {
        o[r] = -(jb[r] + jb[r + 3] + jb[r + 6]);
        for (int q = 0; q < 3; ++q) o[3 * (q + 1) + r] = jb[r + 3 * q];
    }
}

// Nedelec vector load (assemble_tetra_nedelec_vector_load(_order),
// src/assembly/assembly_tetra_nedelec_3d.f90:71-127, 525-582, 2002-2040,
// 2516-2553), step 1, one thread per cell: g_q = det w_q J^-1 f(x_q) at the
// reference's load points, so that f . (J^-T v_r) = (J^-1 f) . v_r.  f is a
// constant vector, or sampled at the mapped points (samples(3, nq, ncell) in
// the reference's cell order).
__global__ __launch_bounds__(kBlockA) void k_ned_load_g(
    int ncell, int nq, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const int *__restrict__ perm, const double *__restrict__ pts, int kind, double f0, double f1,
    double f2, const double *__restrict__ samples, double *__restrict__ g, int *__restrict__ err)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    const size_t e = (size_t)perm[pos];
    for (int q = 0; q < nq; ++q) {
        double f[3] = {f0, f1, f2};
        if (kind == FFH_SOURCE_SAMPLES)
            for (int t = 0; t < 3; ++t) f[t] = samples[(e * nq + q) * 3 + t];
        const double s = det * pts[4 * q + 3];
        for (int a = 0; a < 3; ++a)
            g[((size_t)pos * nq + q) * 3 + a] =
                s * (inv[a] * f[0] + inv[a + 3] * f[1] + inv[a + 6] * f[2]);
    }
}

// step 2, one thread per (cell, oriented local dof r):
// b_r = sum_(q,c) g_qc tab[class][q][c][r] (the class table folds in T^T)
__global__ __launch_bounds__(256) void k_ned_load_elem(int ncell, int n, int nq3,
                                                        const unsigned char *__restrict__ cls_c,
                                                        const double *__restrict__ g,
                                                        const double *__restrict__ tab,
                                                        double *__restrict__ elem)
{
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= (int64_t)ncell * n) return;
    const int pos = (int)(t / n), r = (int)(t - (int64_t)pos * n);
    const double *gp = g + (size_t)pos * nq3;
    const double *tp = tab + (size_t)cls_c[pos] * nq3 * n + r;
    double acc = 0.0;
    for (int j = 0; j < nq3; ++j) acc += gp[j] * tp[(size_t)j * n];
    elem[(size_t)r * ncell + pos] = acc;
}

// 2D P1 load (assemble_scalar_p1_load, src/fortfem_forms_simple.f90:1730-1761):
// f area / 3 to each vertex, one colour per launch.
__global__ __launch_bounds__(kBlockA) void k_p1_2d_load(int pos0, int pos1,
                                                         const int *__restrict__ cells_c,
                                                         const double *__restrict__ verts,
                                                         double f, double *__restrict__ rhs,
                                                         int *__restrict__ err)
{
    const int pos = pos0 + blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= pos1) return;
    const int *c = cells_c + 3 * (size_t)pos;
    double x[3], y[3];
    for (int k = 0; k < 3; ++k) {
        x[k] = verts[2 * (size_t)c[k]];
        y[k] = verts[2 * (size_t)c[k] + 1];
    }
    const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
    if (!(fabs(det) > 64.0 * 2.220446049250313e-16)) {
        atomicOr(err, 1);
        return;
    }
    const double v = f * 0.5 * fabs(det) / 3.0;
    for (int k = 0; k < 3; ++k) rhs[c[k]] += v;
}

// Load phase 1: element vectors det sum_q w_q f(x_q) phi_l(q) into
// EL[l][pos]; tab per point = xi(3), w phi(NL).
template <int NL>
__global__ __launch_bounds__(kBlockA) void k_h1_load(
    int ncell, const int *__restrict__ cells_c, const double *__restrict__ verts,
    const int *__restrict__ perm, const double *__restrict__ tab, int nq, int kind,
    double p0, double p1, double p2, double p3, const double *__restrict__ samples,
    double *__restrict__ EL, int *__restrict__ err)
{
    const int pos = blockIdx.x * kBlockA + threadIdx.x;
    if (pos >= ncell) return;
    const int4 c = reinterpret_cast<const int4 *>(cells_c)[pos];
    double jac[9], inv[9], det;
    if (!tet_geometry(verts, c.x, c.y, c.z, c.w, jac, inv, det)) {
        atomicOr(err, 1);
        return;
    }
    const double *v0 = verts + 3 * (size_t)c.x;
    double acc[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) acc[l] = 0.0;
    const int e = perm[pos];
    for (int q = 0; q < nq; ++q) {
        const double *t = tab + (size_t)q * (3 + NL);
        double f;
        if (kind == FFH_SOURCE_CONSTANT) {
            f = p0;
        } else if (kind == FFH_SOURCE_SIN_PRODUCT) {
            double x = v0[0] + (jac[0] * t[0] + jac[3] * t[1] + jac[6] * t[2]);
            double y = v0[1] + (jac[1] * t[0] + jac[4] * t[1] + jac[7] * t[2]);
            double z = v0[2] + (jac[2] * t[0] + jac[5] * t[1] + jac[8] * t[2]);
            f = p0 * sin(p1 * x) * sin(p2 * y) * sin(p3 * z);
        } else {
            f = samples[(size_t)nq * e + q];
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) acc[l] += f * t[3 + l];
    }
    const size_t T = (size_t)ncell;
#pragma unroll
    for (int l = 0; l < NL; ++l) EL[(size_t)l * T + pos] = det * acc[l];
}

__global__ void k_widen(int64_t n, const float *__restrict__ a, double *__restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

__global__ void k_gather_vals(int64_t n, const int *__restrict__ orig,
                              const double *__restrict__ src, double *__restrict__ dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[orig[i]];
}

// reduced_rhs(f) = rhs(free(f)) - sum_c A(free(f), c) g(c), c ascending
// (sparse_direct_solve_constrained_real, :327-337)
__global__ void k_reduce_rhs(int nf, const int *__restrict__ free_dofs,
                             const double *__restrict__ rhs,
                             const int *__restrict__ lift_ptr,
                             const int *__restrict__ lift_k,
                             const int *__restrict__ lift_c,
                             const double *__restrict__ val,
                             const double *__restrict__ g, double *__restrict__ out)
{
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf;
         f += gridDim.x * blockDim.x) {
        double b = rhs[free_dofs[f]];
        if (g)
            for (int k = lift_ptr[f]; k < lift_ptr[f + 1]; ++k)
                b = b - val[lift_k[k]] * g[lift_c[k]];
        out[f] = b;
    }
}

__global__ void k_scatter_solution(int n, int nf, const int *__restrict__ free_dofs,
                                   const double *__restrict__ xr,
                                   const double *__restrict__ g, double *__restrict__ x)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        x[i] = g ? g[i] : 0.0;
}

__global__ void k_scatter_free(int nf, const int *__restrict__ free_dofs,
                               const double *__restrict__ xr, double *__restrict__ x)
{
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf;
         f += gridDim.x * blockDim.x)
        x[free_dofs[f]] = xr[f];
}

// symmetric elimination of solve_compiled_p1_problem
// (src/solvers/fortfem_solvers_laplacian.f90:88-97), value part: rows and
// columns of boundary dofs become identity, rhs lifted by the column.
__global__ void k_eliminate(int n, const int *__restrict__ ptr,
                            const int *__restrict__ idx, double *__restrict__ val,
                            const int *__restrict__ bnd, double g,
                            const double *__restrict__ rhs_in, double *__restrict__ rhs)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        // row i: b_i - sum over boundary columns d (ascending) of a_id g
        double b = rhs_in[i];
        if (bnd[i]) {
            b = g;
        } else {
            for (int k = ptr[i]; k < ptr[i + 1]; ++k)
                if (bnd[idx[k]]) b = b - val[k] * g;
        }
        rhs[i] = b;
    }
}

__global__ void k_eliminate_vals(int n, const int *__restrict__ ptr,
                                 const int *__restrict__ idx, double *__restrict__ val,
                                 const int *__restrict__ bnd)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += gridDim.x * blockDim.x)
        for (int k = ptr[j]; k < ptr[j + 1]; ++k) {
            int i = idx[k];
            if (bnd[i] || bnd[j]) val[k] = (i == j) ? 1.0 : 0.0;
        }
}

// ---------------------------------------------------------------- helpers --
static int grid1(const Context *ctx, int64_t n, int block = 256)
{
    int64_t need = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ctx->cu_count * 16));
}

void ensure_val64(Context *ctx, DevMatrix &A)
{
    if (!A.val64_stale) return;
    hipLaunchKernelGGL(k_widen, dim3(grid1(ctx, A.nnz)), dim3(256), 0, ctx->stream, A.nnz,
                       A.val32.p, A.val.p);
    FFH_HIP(hipGetLastError());
    A.val64_stale = false;
}

// S tables of the reference-tensor form with the reference's Duffy rule
static void h1_tables(int d, int qdeg, std::vector<double> &tab)
{
    std::vector<double> x, y, z, w;
    tetra_duffy(qdeg, x, y, z, w);
    std::vector<int> bary;
    std::vector<double> nodes;
    lagrange_indices(d, bary, nodes);
    const int n = lagrange_dof_count(d), NN = n * n;
    tab.assign(7 * NN, 0.0);
    std::vector<double> val(n), grad(3 * n);
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (size_t q = 0; q < w.size(); ++q) {
        double pt[3] = {x[q], y[q], z[q]};
        lagrange_eval(d, bary, pt, val.data(), grad.data());
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) {
                const int rc = r * n + c;
                for (int k = 0; k < 6; ++k) {
                    const int a = pa[k], b = pb[k];
                    double s = grad[3 * r + a] * grad[3 * c + b];
                    if (a != b) s += grad[3 * r + b] * grad[3 * c + a];
                    tab[k * NN + rc] += w[q] * s;
                }
                tab[6 * NN + rc] += w[q] * val[r] * val[c];
            }
    }
}

static void h1_load_table(int d, int qdeg, std::vector<double> &tab, int &nq)
{
    std::vector<double> x, y, z, w;
    tetra_duffy(qdeg, x, y, z, w);
    std::vector<int> bary;
    std::vector<double> nodes;
    lagrange_indices(d, bary, nodes);
    const int n = lagrange_dof_count(d);
    nq = (int)w.size();
    tab.assign((size_t)nq * (3 + n), 0.0);
    std::vector<double> val(n), grad(3 * n);
    for (int q = 0; q < nq; ++q) {
        double pt[3] = {x[q], y[q], z[q]};
        lagrange_eval(d, bary, pt, val.data(), grad.data());
        double *t = tab.data() + (size_t)q * (3 + n);
        t[0] = x[q], t[1] = y[q], t[2] = z[q];
        for (int l = 0; l < n; ++l) t[3 + l] = w[q] * val[l];
    }
}

#define FFH_NL_DISPATCH(NLv, ...)                                             \
    switch (NLv) {                                                           \
    case 4: { constexpr int NL = 4; __VA_ARGS__; } break;                     \
    case 10: { constexpr int NL = 10; __VA_ARGS__; } break;                   \
    case 20: { constexpr int NL = 20; __VA_ARGS__; } break;                   \
    case 35: { constexpr int NL = 35; __VA_ARGS__; } break;                   \
    default: throw Error{FFH_ERR_INVALID, "unsupported local dof count"};    \
    }

static void check_err(Problem *P, const char *what)
{
    int h = 0;
    P->d_err.download(&h, 1, P->ctx->stream);
    FFH_HIP(hipStreamSynchronize(P->ctx->stream));
    if (h) {
        FFH_HIP(hipMemsetAsync(P->d_err.p, 0, sizeof(int), P->ctx->stream));
        throw Error{FFH_ERR_INVALID, std::string(what) + ": degenerate cell"};
    }
}

// -------------------------------------------------------------- symbolic --
// The device symbolic phase holds 16 bytes per element entry (keys and their
// sorted copy) besides the 4-byte slots; it runs when that is under a
// quarter of the free device memory (FFH_SYMBOLIC=host forces the host
// path, which builds the same pattern and slots).
static bool distinct_cell_vertices(int64_t T, const int *tets)
{
    for (int64_t e = 0; e < T; ++e)
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b)
                if (tets[4 * e + a] == tets[4 * e + b]) return false;
    return true;
}

static bool device_symbolic_fits(int64_t T, int nl)
{
    static const bool host = [] {
        const char *e = std::getenv("FFH_SYMBOLIC");
        return e && std::string(e) == "host";
    }();
    if (host) return false;
    size_t free_b = 0, total_b = 0;
    FFH_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)device_pattern_bytes(T, nl) + 4.0 * T * nl * nl;
    return need < 0.25 * (double)free_b;
}

// Structures of the H1 gather assembly (k_h1_gather): column blocks, their
// 64-column slices, per slice the step-major occurrences (ascending cell
// order within a column) and value positions.  Left off (gat_ok false, the
// colour scatter assembles) when a column holds more values than a block's
// LDS takes.
static void build_h1_gather(Problem *P, const std::vector<int> &gd,
                            const std::vector<int> &col_ptr)
{
    hipStream_t s = P->ctx->stream;
    const int T = P->ncell, nl = P->n_local, G = P->n_global;
    P->gat_ok = false;
    P->n_gblk = P->gat_nz_max = 0;
    P->d_gat_occ.reset();
    P->d_gat_blk.reset();
    P->d_gat_bslice.reset();
    P->d_gat_soff.reset();
    P->d_gat_slot.reset();
    P->d_h1_terms.reset();
    P->d_cells_e.reset();
    if (G == 0 || T == 0 || (int64_t)T * nl >= INT_MAX) return;
    std::vector<int> blk(1, 0);
    int nz_max = 0;
    for (int j = 0; j < G; ++j) {
        if (col_ptr[j + 1] - col_ptr[j] > kGatherNz) return;
        const int b0 = blk.back();
        if (j > b0 && (col_ptr[j + 1] - col_ptr[b0] > kGatherNz || j - b0 + 1 > kGatherCols)) {
            nz_max = std::max(nz_max, col_ptr[j] - col_ptr[b0]);
            blk.push_back(j);
        }
    }
    nz_max = std::max(nz_max, col_ptr[G] - col_ptr[blk.back()]);
    blk.push_back(G);
    const int nblk = (int)blk.size() - 1;
    // occurrences per column, ascending cell (counting sort)
    std::vector<int> ptr((size_t)G + 1, 0);
    for (size_t i = 0; i < gd.size(); ++i) ++ptr[(size_t)gd[i] + 1];
    for (int j = 0; j < G; ++j) ptr[j + 1] += ptr[j];
    std::vector<int> occ(gd.size()), next(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < T; ++e)
        for (int l = 0; l < nl; ++l) occ[next[gd[(size_t)nl * e + l]]++] = e * nl + l;
    std::vector<int>().swap(next);
    // slices of 64 columns inside each block, step-major occurrences
    std::vector<int> bslice(1, 0), soff(1, 0), sbase;
    std::vector<int> scol, send;
    for (int b = 0; b < nblk; ++b) {
        for (int c = blk[b]; c < blk[b + 1]; c += 64) {
            const int ce = std::min(c + 64, blk[b + 1]);
            int kmax = 0;
            for (int j = c; j < ce; ++j) kmax = std::max(kmax, ptr[j + 1] - ptr[j]);
            soff.push_back(soff.back() + kmax);
            sbase.push_back(col_ptr[blk[b]]);
            scol.push_back(c);
            send.push_back(ce);
        }
        bslice.push_back((int)scol.size());
    }
    const int nslice = (int)scol.size();
    const size_t nent = (size_t)soff.back() * 64;
    std::vector<int> occ_s(nent, -1);
    parallel_for(nslice, [&](int64_t lo, int64_t hi) {
        for (int64_t si = lo; si < hi; ++si) {
            const int c = scol[si];
            for (int j = c; j < send[si]; ++j)
                for (int k = ptr[j]; k < ptr[j + 1]; ++k)
                    occ_s[(size_t)(soff[si] + k - ptr[j]) * 64 + (j - c)] = occ[k];
        }
    });
    std::vector<int>().swap(occ);
    P->d_gat_occ.upload(occ_s, s);
    std::vector<int>().swap(occ_s);
    std::vector<int> cells_e(P->cells.size());
    for (size_t i = 0; i < cells_e.size(); ++i) cells_e[i] = P->cells[i] - 1;
    P->d_cells_e.upload(cells_e, s);
    P->d_gat_blk.upload(blk, s);
    P->d_gat_bslice.upload(bslice, s);
    P->d_gat_soff.upload(soff, s);
    DBuf<int> inv, dsbase;
    dsbase.upload(sbase, s);
    inv.alloc(T);
    hipLaunchKernelGGL(k_invert_perm, dim3((T + 255) / 256), dim3(256), 0, s, T, P->d_perm.p,
                       inv.p);
    const int q4 = 4 * ((nl + 3) / 4);
    P->d_gat_slot.alloc(nent * q4);
    hipLaunchKernelGGL(k_h1_gather_slots, dim3((int)(((int64_t)nslice * 64 + 255) / 256)),
                       dim3(256), 0, s, nslice, T, nl, P->d_gat_soff.p, dsbase.p,
                       P->d_gat_occ.p, inv.p, P->d_slots.p, P->d_gat_slot.p);
    FFH_HIP(hipGetLastError());
    FFH_HIP(hipStreamSynchronize(s)); // host vectors and inv go out of scope
    P->n_gblk = nblk;
    P->gat_nz_max = nz_max;
    P->gat_ok = true;
}

// meshes from this many cells are coloured on the device (see below)
constexpr int64_t kDeviceColourCells = 1 << 20;

void problem_pattern(Problem *P, int family, int order)
{
    Context *ctx = P->ctx;
    const int T = P->ncell;
    std::vector<int> gd;
    int G = 0;
    int nl = 0;
    // incidence lists of the device colouring (vertices, or Nedelec edges)
    DBuf<int> d_inc;
    int inc_m = 0, inc_base = 0, n_ent = 0;
    bool dev = false;
    if (family == FFH_FAMILY_H1) {
        FFH_REQUIRE(P->dim == 3, FFH_ERR_INVALID, "H1 family needs a tetrahedral mesh");
        FFH_REQUIRE(order >= 1 && order <= 4, FFH_ERR_INVALID, "H1 degree must be 1..4");
        for (size_t i = 0; i < P->cells.size(); ++i)
            FFH_REQUIRE(P->cells[i] >= 1 && P->cells[i] <= P->nvert, FFH_ERR_INVALID,
                        "cell vertex id out of range");
        nl = lagrange_dof_count(order);
        dev = device_symbolic_fits(T, nl);
        if (dev) {
            FFH_REQUIRE(distinct_cell_vertices(T, P->cells.data()), FFH_ERR_INVALID,
                        "invalid tetrahedra for the dof map");
            std::vector<int> bary;
            std::vector<double> nodes;
            lagrange_indices(order, bary, nodes);
            d_inc.upload(P->cells, ctx->stream);
            device_lagrange_dofs(ctx, order, T, d_inc.p, bary, gd, G);
            inc_m = 4;
            inc_base = 1;
            n_ent = P->nvert;
        } else {
            FFH_REQUIRE(lagrange_dof_map(order, T, P->cells.data(), gd, G) == 0,
                        FFH_ERR_INVALID, "invalid tetrahedra for the dof map");
        }
    } else if (family == FFH_FAMILY_NEDELEC) {
        FFH_REQUIRE(P->dim == 3, FFH_ERR_INVALID, "Nedelec family needs a tetrahedral mesh");
        FFH_REQUIRE(order >= 1 && order <= 6, FFH_ERR_INVALID, "Nedelec order must be 1..6");
        for (size_t i = 0; i < P->cells.size(); ++i)
            FFH_REQUIRE(P->cells[i] >= 1 && P->cells[i] <= P->nvert, FFH_ERR_INVALID,
                        "cell vertex id out of range");
        nl = nedelec_dof_count(order);
        dev = device_symbolic_fits(T, nl);
        if (dev) {
            FFH_REQUIRE(distinct_cell_vertices(T, P->cells.data()), FFH_ERR_INVALID,
                        "invalid tetrahedra for the Nedelec dof map");
            nedelec_cell_classes(T, P->cells.data(), P->cell_class);
            DBuf<int> dt;
            dt.upload(P->cells, ctx->stream);
            device_nedelec_dofs(ctx, order, T, dt.p, gd, G, P->n_edges, P->n_faces, &d_inc);
            inc_m = 6;
            inc_base = 0;
            n_ent = P->n_edges;
        } else {
            FFH_REQUIRE(nedelec_dof_map(order, T, P->cells.data(), gd, P->cell_class, G,
                                        P->n_edges, P->n_faces) == 0,
                        FFH_ERR_INVALID, "invalid tetrahedra for the Nedelec dof map");
        }
    } else if (family == FFH_FAMILY_P1_2D) {
        FFH_REQUIRE(P->dim == 2, FFH_ERR_INVALID, "P1_2D family needs a triangle mesh");
        nl = 3;
        gd.resize((size_t)3 * T);
        for (size_t i = 0; i < gd.size(); ++i) {
            FFH_REQUIRE(P->cells[i] >= 1 && P->cells[i] <= P->nvert, FFH_ERR_INVALID,
                        "triangle vertex id out of range");
            gd[i] = P->cells[i] - 1;
        }
        G = P->nvert; // dof = vertex index (mesh_2d)
        dev = device_symbolic_fits(T, nl);
        if (dev) {
            d_inc.upload(P->cells, ctx->stream);
            inc_m = 3;
            inc_base = 1;
            n_ent = P->nvert;
        }
    } else {
        throw Error{FFH_ERR_INVALID, "unsupported element family"};
    }
    P->family = family;
    P->order = order;
    P->n_local = nl;
    P->n_global = G;
    std::vector<int> cstart, perm;
    // Colouring: the host's greedy first fit in cell order, or on the device
    // (Jones-Plassmann rounds, first fit) for meshes of kDeviceColourCells
    // cells and more.  Greedy in cell order finds fewer colours on ordered
    // meshes (box meshes: 6 against 12-13 for Nedelec edges, 32 against 40
    // for P1 vertices) and costs under ~20 ms below that size; the device
    // rounds take about half the host time at 12.6 M cells (145 ms against
    // 300 ms).  FFH_COLOUR=device or host forces either.
    const char *cenv = std::getenv("FFH_COLOUR");
    const bool force_dev = cenv && std::string(cenv) == "device";
    const bool force_host = cenv && std::string(cenv) == "host";
    const bool use_dev = dev && !force_host && (force_dev || (int64_t)T >= kDeviceColourCells);
    if (use_dev && device_colour(ctx, T, inc_m, d_inc.p, inc_base, n_ent, cstart, perm))
        P->n_colors = (int)cstart.size() - 1;
    else
        P->n_colors = color_cells(G, T, nl, gd, cstart, perm);
    d_inc.reset();
    std::vector<unsigned char> cls_c;
    if (family == FFH_FAMILY_NEDELEC) {
        // inside a colour, cells sorted by orientation class
        for (int c = 0; c < P->n_colors; ++c)
            std::stable_sort(perm.begin() + cstart[c], perm.begin() + cstart[c + 1],
                             [&](int a, int b) { return P->cell_class[a] < P->cell_class[b]; });
        int slot_of[kNedelecClasses];
        std::fill(slot_of, slot_of + kNedelecClasses, -1);
        P->class_list.clear();
        cls_c.resize(T);
        for (int pos = 0; pos < T; ++pos) {
            const int c = P->cell_class[perm[pos]];
            if (slot_of[c] < 0) {
                slot_of[c] = (int)P->class_list.size();
                P->class_list.push_back(c);
            }
            cls_c[pos] = (unsigned char)slot_of[c];
        }
    }
    P->color_start = cstart;
    P->perm = perm;
    const int vpc = P->dim == 3 ? 4 : 3;
    std::vector<int> cells_c((size_t)vpc * T), gdofs_c((size_t)nl * T);
    parallel_for(T, [&](int64_t lo, int64_t hi) {
        for (int64_t pos = lo; pos < hi; ++pos) {
            const int e = perm[pos];
            for (int k = 0; k < vpc; ++k)
                cells_c[(size_t)vpc * pos + k] = P->cells[(size_t)vpc * e + k] - 1;
            for (int l = 0; l < nl; ++l)
                gdofs_c[(size_t)l * T + pos] = gd[(size_t)nl * e + l];
        }
    });
    hipStream_t s = ctx->stream;
    P->d_perm.upload(perm, s);
    P->d_cells_c.upload(cells_c, s);
    P->d_gdofs_c.upload(gdofs_c, s);
    if (family == FFH_FAMILY_NEDELEC) {
        P->d_cls_c.upload(cls_c, s);
        P->ned_rule = -1;
        P->d_ned_load_tab.reset();
        P->ned_pt_ready = false;
        P->d_ned_ct64.reset();
        P->d_ned_ct32.reset();
        P->d_ptens.reset();
    }
    // CSC pattern and element-entry slots: on the device (radix sort of the
    // entry keys) when the keys fit comfortably, else on the host
    DevMatrix &A = P->full.m;
    std::vector<int> col_ptr;
    if (device_symbolic_fits(T, nl)) {
        device_pattern(ctx, G, T, nl, P->d_gdofs_c.p, A.col_ptr, A.row_idx, col_ptr);
        A.nnz = (int64_t)A.row_idx.n;
        if (family == FFH_FAMILY_NEDELEC) {
            P->d_slots.reset();
            device_slots_cm(ctx, T, nl, P->d_gdofs_c.p, A.col_ptr, A.row_idx, A.nnz, cstart,
                            P->d_slots_cm);
        } else {
            P->d_slots_cm.reset();
            device_slots_h1(ctx, T, nl, P->d_gdofs_c.p, A.col_ptr, A.row_idx, P->d_slots);
        }
    } else {
        std::vector<int> row_idx, slots;
        build_pattern(G, T, nl, gd, col_ptr, row_idx);
        if (family == FFH_FAMILY_NEDELEC) {
            build_entry_slots(T, nl, gd, perm, cstart, col_ptr, row_idx, slots);
            P->d_slots.reset();
            P->d_slots_cm.upload(slots, s);
        } else {
            build_slots(T, nl, gd, perm, col_ptr, row_idx, slots);
            P->d_slots.upload(slots, s);
            P->d_slots_cm.reset();
        }
        A.nnz = (int64_t)row_idx.size();
        A.col_ptr.upload(col_ptr, s);
        A.row_idx.upload(row_idx, s);
    }
    A.n = G;
    if (family == FFH_FAMILY_H1) {
        build_h1_gather(P, gd, col_ptr);
    } else {
        P->gat_ok = false;
        P->d_gat_slot.reset();
        P->d_gat_occ.reset();
        P->d_h1_terms.reset();
        P->d_cells_e.reset();
    }
    A.val.alloc(A.nnz);
    FFH_HIP(hipMemsetAsync(A.val.p, 0, sizeof(double) * A.nnz, s));
    A.symmetric = true;
    A.values_changed();
    A.val32.reset();
    build_rowblocks(ctx, A, col_ptr);
    P->full.ctx = ctx;
    P->full.borrowed = true;
    invalidate_graphs(P->full.ws);
    P->rhs_full.alloc(G);
    FFH_HIP(hipMemsetAsync(P->rhs_full.p, 0, sizeof(double) * G, s));
    P->global_dofs.swap(gd);
    P->d_err.alloc(1);
    FFH_HIP(hipMemsetAsync(P->d_err.p, 0, sizeof(int), s));
    P->tab_qdeg = -1;
    P->load_qdeg = -1;
    P->have_dirichlet = false;
    P->red_dirty = true;
    FFH_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- numeric --
// Nedelec curl-mass assembly (assemble_tetra_nedelec_curl_mass_csc,
// assembly_tetra_nedelec_3d.f90:2042-2112 / 2289-2355: coefficient = curl
// coefficient alpha, mass = beta) and the tensor-weighted variant
// (assemble_tetra_nedelec_weighted_csc, :1870-2000: a symmetric 3x3 tensor per
// cell on the curl term).  Each cell's matrix is T^T M T of the reference, so
// the assembled values match its triplet sums.
// class tables for the element rule of the variant (k = 1: 4 points)
static void ensure_ned_tables(Problem *P, bool weighted)
{
    const int k = P->order;
    const int rule = (k == 1) ? 0 : (weighted ? 2 : 1);
    if (P->ned_rule == rule) return;
    hipStream_t s = P->ctx->stream;
    std::vector<double> tab;
    nedelec_class_tables(k, weighted, P->class_list, tab);
    P->d_ned64.upload(tab, s);
    std::vector<float> t32(tab.begin(), tab.end());
    P->d_ned32.upload(t32, s);
    FFH_HIP(hipStreamSynchronize(s));
    P->ned_rule = rule;
}

// Oriented reference curls at the weighted rule for the classes present
// (k_ned_point_curl), once per pattern.
static void ensure_ned_point_tables(Problem *P)
{
    if (P->ned_pt_ready) return;
    hipStream_t s = P->ctx->stream;
    const int k = P->order, n = P->n_local;
    const int ldn = 32 * ((n + 31) / 32);
    std::vector<double> tab, pts;
    nedelec_curl_point_tables(k, P->class_list, ldn, tab, pts);
    P->d_ned_ct64.upload(tab, s);
    std::vector<float> t32(tab.begin(), tab.end());
    P->d_ned_ct32.upload(t32, s);
    P->d_ned_pt_pts.upload(pts, s);
    FFH_HIP(hipStreamSynchronize(s));
    P->ned_pt_nq = (int)(pts.size() / 4);
    P->ned_pt_ldn = ldn;
    P->ned_pt_ready = true;
}

// The curl part of the per-point weighted element for the cells of one
// colour (positions p0 .. p0 + cnt - 1): waves and regions per wave by order.
template <typename CT>
static void launch_point_curl(Problem *P, int p0, int cnt, CT *val, const CT *ctab)
{
    hipStream_t s = P->ctx->stream;
    const int n = P->n_local, ldn = P->ned_pt_ldn, nq = P->ned_pt_nq;
    constexpr int PC = sizeof(CT) == 8 ? 8 : 16;
    const size_t lds = (size_t)2 * 3 * PC * (ldn + kPointCurlPad<CT>) * sizeof(CT);
    const int ns = ldn / 32, regions = ns * (ns + 1) / 2;
#define FFH_PC_LAUNCH(W, SPW)                                                                   \
    FFH_REQUIRE((W) * (SPW) >= regions && lds <= 98304, FFH_ERR_INVALID,                        \
                "per-point Nedelec kernel shape does not cover the element");                  \
    hipLaunchKernelGGL((k_ned_point_curl<CT, W, SPW, PC>), dim3(cnt), dim3(64 * W), lds, s, p0, \
                       P->ncell, n, ldn, nq, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p,     \
                       P->d_cls_c.p, P->d_ptens.p, P->d_ned_pt_pts.p, ctab, P->d_slots_cm.p,   \
                       val, P->d_err.p)
    // regions of 32 x 32 in the upper triangle: 1, 1, 3, 6, 15, 28; waves in
    // multiples of the CU's 4 SIMDs where the regions allow it
    switch (P->order) {
    case 1:
    case 2: FFH_PC_LAUNCH(1, 1); break;
    case 3: FFH_PC_LAUNCH(3, 1); break;
    case 4: FFH_PC_LAUNCH(8, 1); break;
    case 5: FFH_PC_LAUNCH(8, 2); break;
    case 6: FFH_PC_LAUNCH(8, 4); break;
    default: throw Error{FFH_ERR_INVALID, "Nedelec order must be 1..6"};
    }
#undef FFH_PC_LAUNCH
}

static void assemble_nedelec(Problem *P, int coeff_kind, const double *coeff, double mass,
                             int precision)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    const int T = P->ncell, k = P->order, nl = P->n_local;
    const int np = nl * (nl + 1) / 2;
    double alpha = 1.0;
    const double *dalpha = nullptr, *dtensor = nullptr;
    bool weighted = false;
    if (coeff_kind == FFH_COEFF_SCALAR) {
        if (coeff) alpha = coeff[0];
    } else if (coeff_kind == FFH_COEFF_CELL) {
        if (coeff) {
            P->d_coeff.upload(coeff, (size_t)T, s);
            P->coeff_kind = FFH_COEFF_CELL;
        }
        FFH_REQUIRE(P->coeff_kind == FFH_COEFF_CELL && P->d_coeff.p, FFH_ERR_INVALID,
                    "per-cell coefficients were never uploaded");
        dalpha = P->d_coeff.p;
    } else if (coeff_kind == FFH_COEFF_CELL_TENSOR) {
        if (coeff) {
            for (int64_t e = 0; e < T; ++e) {
                const double *K = coeff + 9 * e;
                double mx = 0.0;
                for (int i = 0; i < 9; ++i) mx = std::max(mx, std::fabs(K[i]));
                FFH_REQUIRE(std::fabs(K[1] - K[3]) <= 1e-12 * mx &&
                                std::fabs(K[2] - K[6]) <= 1e-12 * mx &&
                                std::fabs(K[5] - K[7]) <= 1e-12 * mx,
                            FFH_ERR_INVALID, "cell tensors must be symmetric");
            }
            P->d_coeff.upload(coeff, (size_t)9 * T, s);
            P->coeff_kind = FFH_COEFF_CELL_TENSOR;
        }
        FFH_REQUIRE(P->coeff_kind == FFH_COEFF_CELL_TENSOR && P->d_coeff.p, FFH_ERR_INVALID,
                    "cell tensors were never uploaded");
        dtensor = P->d_coeff.p;
        weighted = true;
    } else if (coeff_kind == FFH_COEFF_POINT_TENSOR) {
        ensure_ned_point_tables(P);
        const int nq = P->ned_pt_nq;
        if (coeff) {
            const int64_t cnt = (int64_t)T * nq;
            for (int64_t i = 0; i < cnt; ++i) {
                const double *K = coeff + 9 * i;
                double mx = 0.0;
                for (int j = 0; j < 9; ++j) {
                    FFH_REQUIRE(std::isfinite(K[j]), FFH_ERR_INVALID, "non-finite point tensor");
                    mx = std::max(mx, std::fabs(K[j]));
                }
                FFH_REQUIRE(std::fabs(K[1] - K[3]) <= 1e-12 * mx &&
                                std::fabs(K[2] - K[6]) <= 1e-12 * mx &&
                                std::fabs(K[5] - K[7]) <= 1e-12 * mx,
                            FFH_ERR_INVALID, "point tensors must be symmetric");
            }
            P->d_ptens.upload(coeff, (size_t)9 * cnt, s);
        }
        FFH_REQUIRE(P->d_ptens.p && P->d_ptens.n == (size_t)9 * T * nq, FFH_ERR_INVALID,
                    "point tensors were never uploaded");
        alpha = 0.0; // the class-table pass adds the mass part only
        weighted = true;
    } else {
        throw Error{FFH_ERR_INVALID, "unknown coefficient kind"};
    }
    ensure_ned_tables(P, weighted);
    P->d_gterm.alloc((size_t)12 * T);
    hipLaunchKernelGGL(k_ned_terms, dim3((T + kBlockA - 1) / kBlockA), dim3(kBlockA), 0, s, T,
                       P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, dalpha, dtensor, alpha,
                       mass, P->d_gterm.p, P->d_err.p);
    FFH_HIP(hipGetLastError());
    check_err(P, "Nedelec assembly");
    DevMatrix &A = P->full.m;
    const bool f32 = precision == FFH_FP32;
    if (f32) A.val32.alloc(A.nnz); // every slot is written (first touch stores)
    const int gy = (nl * nl + 255) / 256;
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        if (cnt == 0) continue;
        const dim3 grid((cnt + kNedCells - 1) / kNedCells, gy);
        if (f32)
            hipLaunchKernelGGL((k_ned_assemble<float, float>), grid, dim3(256), 0, s, p0, cnt, T,
                               nl, P->d_cls_c.p, P->d_gterm.p, P->d_ned32.p, P->d_slots_cm.p,
                               A.val32.p);
        else
            hipLaunchKernelGGL((k_ned_assemble<double, double>), grid, dim3(256), 0, s, p0, cnt,
                               T, nl, P->d_cls_c.p, P->d_gterm.p, P->d_ned64.p, P->d_slots_cm.p,
                               A.val.p);
    }
    if (coeff_kind == FFH_COEFF_POINT_TENSOR)
        for (int c = 0; c < P->n_colors; ++c) {
            const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
            if (cnt == 0) continue;
            if (f32)
                launch_point_curl<float>(P, p0, cnt, A.val32.p, P->d_ned_ct32.p);
            else
                launch_point_curl<double>(P, p0, cnt, A.val.p, P->d_ned_ct64.p);
        }
    FFH_HIP(hipGetLastError());
    check_err(P, "Nedelec assembly");
    A.values_changed();
    A.val32_valid = f32;
    A.val64_stale = f32;
    FFH_HIP(hipStreamSynchronize(s));
    P->assembled = true;
    P->red_dirty = true;
}

static void ensure_h1_tables(Problem *P, int qdeg)
{
    FFH_REQUIRE(qdeg >= 0, FFH_ERR_INVALID, "quadrature degree must be >= 0");
    if (P->tab_qdeg == qdeg) return;
    hipStream_t s = P->ctx->stream;
    std::vector<double> tab;
    h1_tables(P->order, qdeg, tab);
    P->d_tab64.upload(tab, s);
    std::vector<float> t32(tab.begin(), tab.end());
    P->d_tab32.upload(t32, s);
    P->tab_qdeg = qdeg;
}

// Scalar stiffness coefficient: FFH_COEFF_SCALAR sets kappa (coeff[0], 1 when
// null) and returns null; FFH_COEFF_CELL uploads coeff (T values, reusing the
// last upload when null) and returns the device array.
static const double *scalar_coefficients(Problem *P, int coeff_kind, const double *coeff,
                                         double &kappa)
{
    kappa = 1.0;
    if (coeff_kind == FFH_COEFF_SCALAR) {
        if (coeff) kappa = coeff[0];
        return nullptr;
    }
    FFH_REQUIRE(coeff_kind == FFH_COEFF_CELL, FFH_ERR_INVALID,
                "unsupported coefficient kind for this family");
    if (coeff) {
        P->d_coeff.upload(coeff, (size_t)P->ncell, P->ctx->stream);
        P->coeff_kind = FFH_COEFF_CELL;
    }
    FFH_REQUIRE(P->coeff_kind == FFH_COEFF_CELL && P->d_coeff.p, FFH_ERR_INVALID,
                "per-cell coefficients were never uploaded");
    return P->d_coeff.p;
}

void problem_assemble(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                      double mass, int precision)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
    if (P->family == FFH_FAMILY_NEDELEC) {
        assemble_nedelec(P, coeff_kind, coeff, mass, precision);
        return;
    }
    const int T = P->ncell;
    double kappa = 1.0;
    const double *dk = scalar_coefficients(P, coeff_kind, coeff, kappa);
    DevMatrix &A = P->full.m;
    const int nc = P->n_colors;
    if (P->family == FFH_FAMILY_P1_2D) {
        A.val64_stale = false;
        A.values_changed();
        FFH_HIP(hipMemsetAsync(A.val.p, 0, sizeof(double) * A.nnz, s));
        for (int c = 0; c < nc; ++c) {
            int p0 = P->color_start[c], p1 = P->color_start[c + 1];
            int grid = (p1 - p0 + kBlockA - 1) / kBlockA;
            if (grid == 0) continue;
            hipLaunchKernelGGL(k_p1_2d_cells, dim3(grid), dim3(kBlockA), 0, s, p0, p1, T,
                               P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, dk, kappa,
                               mass, P->d_slots.p, A.val.p, P->d_err.p);
        }
        FFH_HIP(hipGetLastError());
        check_err(P, "2D P1 assembly");
    } else {
        ensure_h1_tables(P, qdeg);
        const bool f32 = precision == FFH_FP32;
        const int nl = P->n_local;
        if (f32) A.val32.alloc(A.nnz);
        if (P->gat_ok) {
            // per-cell coefficients, then the column gather (every value
            // written, no clearing)
            P->d_h1_terms.alloc((size_t)8 * T);
            const int gridc = (T + kBlockA - 1) / kBlockA;
            const size_t lds = (size_t)P->gat_nz_max * (f32 ? sizeof(float) : sizeof(double));
            const ushort4 *gsl = reinterpret_cast<const ushort4 *>(P->d_gat_slot.p);
            FFH_NL_DISPATCH(nl, {
                if (f32) {
                    float *t32 = reinterpret_cast<float *>(P->d_h1_terms.p);
                    hipLaunchKernelGGL((k_h1_terms<float>), dim3(gridc), dim3(kBlockA), 0, s, T,
                                       P->d_cells_e.p, P->d_vertices.p, dk, kappa, mass, t32,
                                       P->d_err.p);
                    hipLaunchKernelGGL((k_h1_gather<NL, float>), dim3(P->n_gblk),
                                       dim3(kGatherCols), lds, s, P->d_gat_blk.p, P->d_gat_bslice.p,
                                       P->d_gat_soff.p, A.col_ptr.p, P->d_gat_occ.p, gsl, t32,
                                       P->d_tab32.p, mass != 0.0 ? 1 : 0, A.val32.p);
                } else {
                    hipLaunchKernelGGL((k_h1_terms<double>), dim3(gridc), dim3(kBlockA), 0, s, T,
                                       P->d_cells_e.p, P->d_vertices.p, dk, kappa, mass,
                                       P->d_h1_terms.p, P->d_err.p);
                    hipLaunchKernelGGL((k_h1_gather<NL, double>), dim3(P->n_gblk),
                                       dim3(kGatherCols), lds, s, P->d_gat_blk.p, P->d_gat_bslice.p,
                                       P->d_gat_soff.p, A.col_ptr.p, P->d_gat_occ.p, gsl,
                                       P->d_h1_terms.p, P->d_tab64.p, mass != 0.0 ? 1 : 0,
                                       A.val.p);
                }
            });
            FFH_HIP(hipGetLastError());
            check_err(P, "tetrahedral H1 assembly");
            A.values_changed();
            A.val32_valid = f32;
            A.val64_stale = f32;
            P->assembled = true;
            P->red_dirty = true;
            return;
        }
        if (f32)
            FFH_HIP(hipMemsetAsync(A.val32.p, 0, sizeof(float) * A.nnz, s));
        else
            FFH_HIP(hipMemsetAsync(A.val.p, 0, sizeof(double) * A.nnz, s));
        const size_t nup = (size_t)nl * (nl + 1) / 2;
        P->d_elem.alloc(nup * T);
        double *E64 = P->d_elem.p;
        float *E32 = reinterpret_cast<float *>(P->d_elem.p);
        const int gridc = (T + kBlockA - 1) / kBlockA;
        FFH_NL_DISPATCH(nl, {
            if (f32)
                hipLaunchKernelGGL((k_h1_elem<NL, float>), dim3(gridc), dim3(kBlockA), 0, s,
                                   T, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, dk,
                                   kappa, mass, P->d_tab32.p, E32, P->d_err.p);
            else
                hipLaunchKernelGGL((k_h1_elem<NL, double>), dim3(gridc), dim3(kBlockA), 0, s,
                                   T, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, dk,
                                   kappa, mass, P->d_tab64.p, E64, P->d_err.p);
        });
        for (int c = 0; c < nc; ++c) {
            const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
            const int64_t work = (int64_t)cnt * nl * nl;
            if (work == 0) continue;
            const int grid = (int)((work + 255) / 256);
            if (f32)
                hipLaunchKernelGGL((k_scatter_colour<float, float>), dim3(grid), dim3(256), 0,
                                   s, p0, cnt, T, nl, E32, P->d_slots.p, A.val32.p);
            else
                hipLaunchKernelGGL((k_scatter_colour<double, double>), dim3(grid), dim3(256),
                                   0, s, p0, cnt, T, nl, E64, P->d_slots.p, A.val.p);
        }
        FFH_HIP(hipGetLastError());
        check_err(P, "tetrahedral H1 assembly");
        // FP32 contraction: the FP64 view (export, FP64 solves) is made on
        // first use
        A.values_changed();
        A.val32_valid = f32;
        A.val64_stale = f32;
    }
    P->assembled = true;
    P->red_dirty = true;
}

// Nedelec curl-mass tangent (assemble_tetra_nedelec_curl_mass_csc_jvp,
// assembly_tetra_nedelec_3d.f90:2114-2180): tangent terms, then the
// primal's class-table assembly, colour by colour (every slot gets a first
// store, so the output needs no clearing).
static void nedelec_jvp(Problem *P, int coeff_kind, const double *coeff, double beta,
                        const double *vertices_dot, const double *coeff_dot, double beta_dot,
                        double *values_dot)
{
    hipStream_t s = P->ctx->stream;
    const int T = P->ncell, nl = P->n_local;
    double alpha = 1.0;
    const double *da = scalar_coefficients(P, coeff_kind, coeff, alpha);
    ensure_ned_tables(P, false);
    double alpha_dot = 0.0;
    DBuf<double> dad, dvd, vals;
    if (coeff_kind == FFH_COEFF_SCALAR) {
        if (coeff_dot) alpha_dot = coeff_dot[0];
    } else if (coeff_dot) {
        dad.upload(coeff_dot, (size_t)T, s);
    }
    if (vertices_dot) dvd.upload(vertices_dot, 3 * (size_t)P->nvert, s);
    P->d_gterm.alloc((size_t)12 * T);
    hipLaunchKernelGGL(k_ned_terms_jvp, dim3((T + kBlockA - 1) / kBlockA), dim3(kBlockA), 0, s,
                       T, P->d_cells_c.p, P->d_vertices.p, dvd.p, P->d_perm.p, da, dad.p, alpha,
                       alpha_dot, beta, beta_dot, P->d_gterm.p, P->d_err.p);
    const int64_t nnz = P->full.m.nnz;
    vals.alloc(std::max<int64_t>(nnz, 1));
    const int gy = (nl * nl + 255) / 256;
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        if (cnt == 0) continue;
        hipLaunchKernelGGL((k_ned_assemble<double, double>),
                           dim3((cnt + kNedCells - 1) / kNedCells, gy), dim3(256), 0, s, p0, cnt,
                           T, nl, P->d_cls_c.p, P->d_gterm.p, P->d_ned64.p, P->d_slots_cm.p,
                           vals.p);
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "Nedelec assembly tangent");
    vals.download(values_dot, nnz, s);
    FFH_HIP(hipStreamSynchronize(s));
}

// Nedelec curl-mass cotangent (:2182-2287): k_ned_vjp per colour, per-cell
// alpha / beta bars summed on the host in cell order.
static void nedelec_vjp(Problem *P, int coeff_kind, const double *coeff, double beta,
                        const double *values_bar, double *vertices_bar, double *coeff_bar,
                        double *beta_bar)
{
    hipStream_t s = P->ctx->stream;
    const int T = P->ncell, nl = P->n_local;
    double alpha = 1.0;
    const double *da = scalar_coefficients(P, coeff_kind, coeff, alpha);
    ensure_ned_tables(P, false);
    const int64_t nnz = P->full.m.nnz;
    DBuf<double> vb, xb, ab, bb;
    vb.upload(values_bar, (size_t)nnz, s);
    xb.alloc(12 * (size_t)std::max(T, 1)); // per cell: 4 vertices x 3
    ab.alloc(std::max(T, 1));
    bb.alloc(std::max(T, 1));
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_ned_vjp, dim3((cnt + 3) / 4), dim3(256), 0, s, p0, cnt, T, nl,
                           P->d_cls_c.p, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, da,
                           alpha, beta, P->d_ned64.p, P->d_slots_cm.p, vb.p, xb.p, ab.p, bb.p,
                           P->d_err.p);
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "Nedelec assembly cotangent");
    std::vector<double> ha(T), hb(T), hx(12 * (size_t)T);
    ab.download(ha.data(), T, s);
    bb.download(hb.data(), T, s);
    xb.download(hx.data(), 12 * (size_t)T, s);
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<double> ca(T), cb(T);
    std::vector<int> pos_of(T);
    for (int pos = 0; pos < T; ++pos) {
        ca[P->perm[pos]] = ha[pos];
        cb[P->perm[pos]] = hb[pos];
        pos_of[P->perm[pos]] = pos;
    }
    // vertex cotangents summed over the cells in ascending cell order
    std::fill(vertices_bar, vertices_bar + 3 * (size_t)P->nvert, 0.0);
    for (int e = 0; e < T; ++e) {
        const double *c = hx.data() + (size_t)pos_of[e] * 12;
        for (int q = 0; q < 4; ++q) {
            double *vbq = vertices_bar + 3 * (size_t)(P->cells[(size_t)4 * e + q] - 1);
            for (int r = 0; r < 3; ++r) vbq[r] += c[3 * q + r];
        }
    }
    double asum = 0.0, bsum = 0.0;
    for (int e = 0; e < T; ++e) {
        asum += ca[e];
        bsum += cb[e];
    }
    if (coeff_kind == FFH_COEFF_CELL)
        std::copy(ca.begin(), ca.end(), coeff_bar);
    else
        coeff_bar[0] = asum;
    *beta_bar = bsum;
}

// Tangent of the tetrahedral H1 matrix on the problem's pattern
// (assemble_tetra_lagrange_stiffness_csc_jvp,
// assembly_tetra_lagrange_arbitrary_order_3d.f90:336-386): the element
// tangents of k_h1_elem_jvp through the primal's colour scatter, so the
// result has the primal's CSC order.  values_dot: host, nnz of the full
// matrix.
void problem_assemble_jvp(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                          double mass, const double *vertices_dot, const double *coeff_dot,
                          double mass_dot, double *values_dot)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
    if (P->family == FFH_FAMILY_NEDELEC) {
        nedelec_jvp(P, coeff_kind, coeff, mass, vertices_dot, coeff_dot, mass_dot, values_dot);
        return;
    }
    FFH_REQUIRE(P->family == FFH_FAMILY_H1, FFH_ERR_INVALID,
                "assembly tangents are built for the tetrahedral H1 and Nedelec families");
    const int T = P->ncell;
    double kappa = 1.0;
    const double *dk = scalar_coefficients(P, coeff_kind, coeff, kappa);
    ensure_h1_tables(P, qdeg);
    double kappa_dot = 0.0;
    DBuf<double> dkd, dvd;
    if (coeff_kind == FFH_COEFF_SCALAR) {
        if (coeff_dot) kappa_dot = coeff_dot[0];
    } else if (coeff_dot) {
        dkd.upload(coeff_dot, (size_t)T, s);
    }
    if (vertices_dot) dvd.upload(vertices_dot, 3 * (size_t)P->nvert, s);
    const int nl = P->n_local;
    const size_t nup = (size_t)nl * (nl + 1) / 2;
    P->d_elem.alloc(nup * T);
    const int64_t nnz = P->full.m.nnz;
    DBuf<double> vals;
    vals.alloc(std::max<int64_t>(nnz, 1));
    FFH_HIP(hipMemsetAsync(vals.p, 0, sizeof(double) * nnz, s));
    const int gridc = (T + kBlockA - 1) / kBlockA;
    FFH_NL_DISPATCH(nl, {
        hipLaunchKernelGGL((k_h1_elem_jvp<NL>), dim3(gridc), dim3(kBlockA), 0, s, T,
                           P->d_cells_c.p, P->d_vertices.p, dvd.p, P->d_perm.p, dk, dkd.p,
                           kappa, kappa_dot, mass, mass_dot, P->d_tab64.p, P->d_elem.p,
                           P->d_err.p);
    });
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        const int64_t work = (int64_t)cnt * nl * nl;
        if (work == 0) continue;
        hipLaunchKernelGGL((k_scatter_colour<double, double>), dim3((int)((work + 255) / 256)),
                           dim3(256), 0, s, p0, cnt, T, nl, P->d_elem.p, P->d_slots.p, vals.p);
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "tetrahedral H1 assembly tangent");
    vals.download(values_dot, nnz, s);
    FFH_HIP(hipStreamSynchronize(s));
}

// Cotangent of the tetrahedral H1 matrix
// (assemble_tetra_lagrange_stiffness_csc_vjp, :388-471): one k_h1_elem_vjp
// launch per colour; the coefficient bars are summed on the host in cell
// order.  vertices_bar: 3 nv; coeff_bar: 1 value (FFH_COEFF_SCALAR) or one
// per cell (FFH_COEFF_CELL); mass_bar: 1 value.
void problem_assemble_vjp(Problem *P, int qdeg, int coeff_kind, const double *coeff,
                          double mass, const double *values_bar, double *vertices_bar,
                          double *coeff_bar, double *mass_bar)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
    if (P->family == FFH_FAMILY_NEDELEC) {
        nedelec_vjp(P, coeff_kind, coeff, mass, values_bar, vertices_bar, coeff_bar, mass_bar);
        return;
    }
    FFH_REQUIRE(P->family == FFH_FAMILY_H1, FFH_ERR_INVALID,
                "assembly cotangents are built for the tetrahedral H1 and Nedelec families");
    const int T = P->ncell;
    double kappa = 1.0;
    const double *dk = scalar_coefficients(P, coeff_kind, coeff, kappa);
    ensure_h1_tables(P, qdeg);
    const int64_t nnz = P->full.m.nnz;
    DBuf<double> vb, xb, kb, mb;
    vb.upload(values_bar, (size_t)nnz, s);
    xb.alloc(3 * (size_t)P->nvert);
    FFH_HIP(hipMemsetAsync(xb.p, 0, sizeof(double) * 3 * (size_t)P->nvert, s));
    kb.alloc(std::max(T, 1));
    mb.alloc(std::max(T, 1));
    const int nl = P->n_local;
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        if (cnt == 0) continue;
        FFH_NL_DISPATCH(nl, {
            hipLaunchKernelGGL((k_h1_elem_vjp<NL>), dim3((cnt + kBlockA - 1) / kBlockA),
                               dim3(kBlockA), 0, s, p0, cnt, T, P->d_cells_c.p,
                               P->d_vertices.p, P->d_perm.p, dk, kappa, mass,
                               P->d_tab64.p, P->d_slots.p, vb.p, xb.p, kb.p, mb.p,
                               P->d_err.p);
        });
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "tetrahedral H1 assembly cotangent");
    std::vector<double> hk(T), hm(T);
    kb.download(hk.data(), T, s);
    mb.download(hm.data(), T, s);
    xb.download(vertices_bar, 3 * (size_t)P->nvert, s);
    FFH_HIP(hipStreamSynchronize(s));
    // colour position -> cell, then sums in cell order
    std::vector<double> ck(T), cm(T);
    for (int pos = 0; pos < T; ++pos) {
        ck[P->perm[pos]] = hk[pos];
        cm[P->perm[pos]] = hm[pos];
    }
    double ksum = 0.0, msum = 0.0;
    for (int e = 0; e < T; ++e) {
        ksum += ck[e];
        msum += cm[e];
    }
    if (coeff_kind == FFH_COEFF_CELL)
        std::copy(ck.begin(), ck.end(), coeff_bar);
    else
        coeff_bar[0] = ksum;
    *mass_bar = msum;
}

// Complex curl-curl PML operator on the Nedelec pattern:
// assemble_tetra_nedelec_pml_csc (src/assembly/assembly_tetra_nedelec_3d.f90:
// 1613-1685, element :826-887) for kind 0 and
// assemble_tetra_nedelec_curvilinear_pml_csc (:1372-1444, element :1092-1150)
// for kind 1.  The element is the weighted curl-mass element with per-cell
// complex symmetric tensors, curl K and value -k^2 M, from the stretch:
// cartesian (cartesian_curl_curl_pml_coefficients,
// src/boundary/cartesian_helmholtz_pml.f90:23-41, 94-108) m_i = s_j s_k / s_i,
// K = diag(1 / m_i), M = diag(m_i); curvilinear
// (curvilinear_curl_curl_pml_coefficients, curvilinear_helmholtz_pml.f90:
// 113-131, 202-216) K = S^T S / det S, M = det S S^-1 S^-T.  The element is
// linear in (K, M), so the real and imaginary parts are two real tensor
// assemblies on the class tables.  stretch: complex (re, im) pairs, 3 per
// cell (kind 0) or 3 x 3 column-major per cell (kind 1), in cell order;
// values: nnz (re, im) pairs in the pattern's order (host).
void problem_nedelec_pml(Problem *P, int kind, const double *stretch, double k0,
                         double *values)
{
    using cd = std::complex<double>;
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0 && P->family == FFH_FAMILY_NEDELEC, FFH_ERR_INVALID,
                "the PML operator needs a Nedelec pattern");
    FFH_REQUIRE(kind == 0 || kind == 1, FFH_ERR_INVALID, "PML kind is 0 or 1");
    FFH_REQUIRE(k0 > 0.0, FFH_ERR_INVALID, "wave number must be positive");
    const int T = P->ncell;
    std::vector<double> kt[2], mt[2];
    for (int c = 0; c < 2; ++c) {
        kt[c].assign((size_t)9 * T, 0.0);
        mt[c].assign((size_t)9 * T, 0.0);
    }
    const double k2 = k0 * k0;
    for (int e = 0; e < T; ++e) {
        cd K[9], M[9];
        if (kind == 0) {
            cd sv[3];
            for (int i = 0; i < 3; ++i) {
                sv[i] = cd(stretch[6 * (size_t)e + 2 * i], stretch[6 * (size_t)e + 2 * i + 1]);
                FFH_REQUIRE(std::abs(sv[i]) > DBL_MIN, FFH_ERR_INVALID, "zero PML stretch");
            }
            for (int i = 0; i < 9; ++i) K[i] = M[i] = 0.0;
            for (int i = 0; i < 3; ++i) {
                const cd m = sv[(i + 1) % 3] * sv[(i + 2) % 3] / sv[i];
                M[4 * i] = m;
                K[4 * i] = 1.0 / m;
            }
        } else {
            cd S[9];
            double mx = 0.0;
            for (int i = 0; i < 9; ++i) {
                S[i] = cd(stretch[18 * (size_t)e + 2 * i], stretch[18 * (size_t)e + 2 * i + 1]);
                mx = std::max(mx, std::abs(S[i]));
            }
#define Sx(r, c) S[(r) + 3 * (c)]
            cd cof[9]; // cofactor matrix, cof(r, c)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3,
                              c2 = (c + 2) % 3;
                    cof[r + 3 * c] = Sx(r1, c1) * Sx(r2, c2) - Sx(r1, c2) * Sx(r2, c1);
                }
            const cd det = Sx(0, 0) * cof[0] + Sx(0, 1) * cof[3] + Sx(0, 2) * cof[6];
#undef Sx
            const double sc = std::max(1.0, mx);
            FFH_REQUIRE(std::isfinite(det.real()) && std::isfinite(det.imag()) &&
                            std::abs(det) > 64.0 * DBL_EPSILON * sc * sc * sc,
                        FFH_ERR_INVALID, "singular PML stretch");
            cd inv[9]; // S^-1 = cof^T / det
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) inv[r + 3 * c] = cof[c + 3 * r] / det;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    cd a = 0.0, b = 0.0;
                    for (int t = 0; t < 3; ++t) {
                        a += S[t + 3 * r] * S[t + 3 * c];     // (S^T S)_rc
                        b += inv[r + 3 * t] * inv[c + 3 * t]; // (S^-1 S^-T)_rc
                    }
                    K[r + 3 * c] = a / det;
                    M[r + 3 * c] = det * b;
                }
        }
        for (int i = 0; i < 9; ++i) {
            const cd m = -k2 * M[i];
            kt[0][9 * (size_t)e + i] = K[i].real();
            kt[1][9 * (size_t)e + i] = K[i].imag();
            mt[0][9 * (size_t)e + i] = m.real();
            mt[1][9 * (size_t)e + i] = m.imag();
        }
    }
    ensure_ned_tables(P, true);
    const int nl = P->n_local;
    const int64_t nnz = P->full.m.nnz;
    P->d_gterm.alloc((size_t)12 * T);
    DBuf<double> dk, dm, vals[2];
    std::vector<double> part[2];
    const int gy = (nl * nl + 255) / 256;
    for (int c = 0; c < 2; ++c) {
        dk.upload(kt[c], s);
        dm.upload(mt[c], s);
        hipLaunchKernelGGL(k_ned_terms, dim3((T + kBlockA - 1) / kBlockA), dim3(kBlockA), 0, s,
                           T, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, nullptr, dk.p, 1.0,
                           0.0, P->d_gterm.p, P->d_err.p, dm.p);
        vals[c].alloc(std::max<int64_t>(nnz, 1));
        for (int col = 0; col < P->n_colors; ++col) {
            const int p0 = P->color_start[col], cnt = P->color_start[col + 1] - p0;
            if (cnt == 0) continue;
            hipLaunchKernelGGL((k_ned_assemble<double, double>),
                               dim3((cnt + kNedCells - 1) / kNedCells, gy), dim3(256), 0, s, p0,
                               cnt, T, nl, P->d_cls_c.p, P->d_gterm.p, P->d_ned64.p,
                               P->d_slots_cm.p, vals[c].p);
        }
        FFH_HIP(hipGetLastError());
        check_err(P, "Nedelec PML assembly");
        part[c].resize(nnz);
        vals[c].download(part[c].data(), nnz, s);
        FFH_HIP(hipStreamSynchronize(s)); // dk / dm are reused for the imaginary part
    }
    for (int64_t k = 0; k < nnz; ++k) {
        values[2 * k] = part[0][k];
        values[2 * k + 1] = part[1][k];
    }
}

// Tangent of the complex PML operator (assemble_tetra_nedelec_pml_csc_jvp,
// assembly_tetra_nedelec_3d.f90:1687-1768; curvilinear :1446-1521) along
// (vertices_dot, stretch_dot, k0_dot).  The operator is linear in its twelve
// per-cell terms c (curl: J^T K J / det; value: det J^-1 M' J^-T with
// M' = -k0^2 M), so the tangent is the same class-table assembly applied to
// the terms' tangents dc, formed per cell in complex arithmetic on the host:
//   dK, dM from the stretch tangent (cartesian: dm_i = m_i (ds_j/s_j +
//   ds_k/s_k - ds_i/s_i), dK_ii = -dm_i / m_i^2; curvilinear: dK = (dS^T S +
//   S^T dS)/det S - K ddet/det, dM = ddet S^-1 S^-T + det (dI S^-T + S^-1 dI^T),
//   dI = -S^-1 dS S^-1, ddet = det tr(S^-1 dS)),
//   dM' = -k0^2 dM - 2 k0 dk0 M, and the geometry tangent from dJ (the
//   vertex tangents of the cell's edges), ddetJ = det tr(J^-1 dJ),
//   d(J^-1) = -J^-1 dJ J^-1.
void problem_nedelec_pml_jvp(Problem *P, int kind, const double *stretch,
                             const double *stretch_dot, double k0, double k0_dot,
                             const double *vertices_dot, double *values_dot)
{
    using cd = std::complex<double>;
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0 && P->family == FFH_FAMILY_NEDELEC, FFH_ERR_INVALID,
                "the PML operator needs a Nedelec pattern");
    FFH_REQUIRE(kind == 0 || kind == 1, FFH_ERR_INVALID, "PML kind is 0 or 1");
    FFH_REQUIRE(k0 > 0.0 && std::isfinite(k0_dot), FFH_ERR_INVALID,
                "wave number must be positive");
    const int T = P->ncell;
    const int per = kind == 0 ? 3 : 9;
    std::vector<double> gt[2];
    gt[0].assign((size_t)12 * T, 0.0);
    gt[1].assign((size_t)12 * T, 0.0);
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    auto mul = [](const cd *A, const cd *B, cd *C) { // column-major 3x3
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                cd a = 0.0;
                for (int t = 0; t < 3; ++t) a += A[r + 3 * t] * B[t + 3 * c];
                C[r + 3 * c] = a;
            }
    };
    auto tr = [](const cd *A, cd *B) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[r + 3 * c] = A[c + 3 * r];
    };
    auto inv3 = [](const cd *A, cd *I, cd &det) {
        cd cof[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                cof[r + 3 * c] = A[r1 + 3 * c1] * A[r2 + 3 * c2] - A[r1 + 3 * c2] * A[r2 + 3 * c1];
            }
        det = A[0] * cof[0] + A[3] * cof[3] + A[6] * cof[6];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) I[r + 3 * c] = cof[c + 3 * r] / det;
    };
    for (int pos = 0; pos < T; ++pos) {
        const int e = P->perm[pos];
        // coefficient tensors and their tangents
        cd K[9], M[9], dK[9], dM[9];
        for (int i = 0; i < 9; ++i) K[i] = M[i] = dK[i] = dM[i] = 0.0;
        const double *st = stretch + 2 * (size_t)per * e;
        const double *sd = stretch_dot + 2 * (size_t)per * e;
        if (kind == 0) {
            cd sv[3], sdv[3];
            for (int i = 0; i < 3; ++i) {
                sv[i] = cd(st[2 * i], st[2 * i + 1]);
                sdv[i] = cd(sd[2 * i], sd[2 * i + 1]);
                FFH_REQUIRE(std::abs(sv[i]) > DBL_MIN, FFH_ERR_INVALID, "zero PML stretch");
            }
            for (int i = 0; i < 3; ++i) {
                const int j = (i + 1) % 3, k = (i + 2) % 3;
                const cd m = sv[j] * sv[k] / sv[i];
                const cd dm = m * (sdv[j] / sv[j] + sdv[k] / sv[k] - sdv[i] / sv[i]);
                M[4 * i] = m;
                dM[4 * i] = dm;
                K[4 * i] = 1.0 / m;
                dK[4 * i] = -dm / (m * m);
            }
        } else {
            cd S[9], dS[9], Si[9], det;
            double mx = 0.0;
            for (int i = 0; i < 9; ++i) {
                S[i] = cd(st[2 * i], st[2 * i + 1]);
                dS[i] = cd(sd[2 * i], sd[2 * i + 1]);
                mx = std::max(mx, std::abs(S[i]));
            }
            inv3(S, Si, det);
            const double sc = std::max(1.0, mx);
            FFH_REQUIRE(std::isfinite(det.real()) && std::isfinite(det.imag()) &&
                            std::abs(det) > 64.0 * DBL_EPSILON * sc * sc * sc,
                        FFH_ERR_INVALID, "singular PML stretch");
            cd St[9], dSt[9], StS[9], a1[9], a2[9], SidS[9], dI[9], dIt[9], Sit[9], W[9];
            tr(S, St);
            tr(dS, dSt);
            mul(St, S, StS);
            mul(Si, dS, SidS);
            const cd ddet = det * (SidS[0] + SidS[4] + SidS[8]);
            mul(SidS, Si, dI); // S^-1 dS S^-1
            for (int i = 0; i < 9; ++i) dI[i] = -dI[i];
            tr(dI, dIt);
            tr(Si, Sit);
            mul(dSt, S, a1);
            mul(St, dS, a2);
            for (int i = 0; i < 9; ++i) {
                K[i] = StS[i] / det;
                dK[i] = (a1[i] + a2[i]) / det - K[i] * ddet / det;
            }
            mul(Si, Sit, W);
            mul(dI, Sit, a1);
            mul(Si, dIt, a2);
            for (int i = 0; i < 9; ++i) {
                M[i] = det * W[i];
                dM[i] = ddet * W[i] + det * (a1[i] + a2[i]);
            }
        }
        cd Mp[9], dMp[9];
        for (int i = 0; i < 9; ++i) {
            Mp[i] = -k0 * k0 * M[i];
            dMp[i] = -k0 * k0 * dM[i] - 2.0 * k0 * k0_dot * M[i];
        }
        // geometry and its tangent
        const int *cv = P->cells.data() + 4 * (size_t)e;
        cd J[9], dJ[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                const size_t a = 3 * (size_t)(cv[c + 1] - 1) + r, o = 3 * (size_t)(cv[0] - 1) + r;
                J[r + 3 * c] = P->vertices[a] - P->vertices[o];
                dJ[r + 3 * c] = vertices_dot ? vertices_dot[a] - vertices_dot[o] : 0.0;
            }
        cd Ji[9], det;
        inv3(J, Ji, det);
        FFH_REQUIRE(det.real() > 0.0, FFH_ERR_INVALID, "degenerate or inverted cell");
        cd JidJ[9], dJi[9], Jt[9], dJt[9], Jit[9], dJit[9];
        mul(Ji, dJ, JidJ);
        const cd ddet = det * (JidJ[0] + JidJ[4] + JidJ[8]);
        mul(JidJ, Ji, dJi);
        for (int i = 0; i < 9; ++i) dJi[i] = -dJi[i];
        tr(J, Jt);
        tr(dJ, dJt);
        tr(Ji, Jit);
        tr(dJi, dJit);
        // curl terms: J^T K J / det
        cd KJ[9], C0[9], t1[9], t2[9], t3[9], tmp[9];
        mul(K, J, KJ);
        mul(Jt, KJ, C0);
        mul(dJt, KJ, t1);
        mul(K, dJ, tmp);
        mul(Jt, tmp, t2);
        mul(dK, J, tmp);
        mul(Jt, tmp, t3);
        // value terms: det J^-1 M' J^-T
        cd MJ[9], V0[9], u1[9], u2[9], u3[9];
        mul(Mp, Jit, MJ);
        mul(Ji, MJ, V0);
        mul(dJi, MJ, u1);
        mul(dMp, Jit, tmp);
        mul(Ji, tmp, u2);
        mul(Mp, dJit, tmp);
        mul(Ji, tmp, u3);
        for (int q = 0; q < 6; ++q) {
            const int i = pa[q] + 3 * pb[q];
            const cd dc = (t1[i] + t2[i] + t3[i]) / det - C0[i] * ddet / (det * det);
            const cd dv = ddet * V0[i] + det * (u1[i] + u2[i] + u3[i]);
            gt[0][(size_t)q * T + pos] = dc.real();
            gt[1][(size_t)q * T + pos] = dc.imag();
            gt[0][(size_t)(6 + q) * T + pos] = dv.real();
            gt[1][(size_t)(6 + q) * T + pos] = dv.imag();
        }
    }
    ensure_ned_tables(P, true);
    const int nl = P->n_local;
    const int64_t nnz = P->full.m.nnz;
    const int gy = (nl * nl + 255) / 256;
    DBuf<double> vals[2];
    std::vector<double> part[2];
    for (int c = 0; c < 2; ++c) {
        P->d_gterm.upload(gt[c], s);
        vals[c].alloc(std::max<int64_t>(nnz, 1));
        for (int col = 0; col < P->n_colors; ++col) {
            const int p0 = P->color_start[col], cnt = P->color_start[col + 1] - p0;
            if (cnt == 0) continue;
            hipLaunchKernelGGL((k_ned_assemble<double, double>),
                               dim3((cnt + kNedCells - 1) / kNedCells, gy), dim3(256), 0, s, p0,
                               cnt, T, nl, P->d_cls_c.p, P->d_gterm.p, P->d_ned64.p,
                               P->d_slots_cm.p, vals[c].p);
        }
        FFH_HIP(hipGetLastError());
        part[c].resize(nnz);
        vals[c].download(part[c].data(), nnz, s);
        FFH_HIP(hipStreamSynchronize(s)); // d_gterm is rewritten for the imaginary part
    }
    for (int64_t k = 0; k < nnz; ++k) {
        values_dot[2 * k] = part[0][k];
        values_dot[2 * k + 1] = part[1][k];
    }
}

// Cotangent of the complex PML operator (assemble_tetra_nedelec_pml_csc_vjp,
// assembly_tetra_nedelec_3d.f90:1770-1868; curvilinear :1523-1611) in the
// reference's convention: for matrix cotangent seeds w, the cotangent of an
// input x is sum w conj(dA/dx) (its real part for the real vertices and wave
// number; the stretch cotangent is complex).  The device contracts the seeds
// (real and imaginary parts) with the class tables into the twelve per-cell
// term cotangents; the host chains them per cell to J, K, M' = -k0^2 M and
// on to the vertices, the wave number and the stretch (the stretch part
// through the coefficient tangents of problem_nedelec_pml_jvp, one unit
// direction per stretch entry).
void problem_nedelec_pml_vjp(Problem *P, int kind, const double *stretch, double k0,
                             const double *values_bar, double *vertices_bar,
                             double *stretch_bar, double *k0_bar)
{
    using cd = std::complex<double>;
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0 && P->family == FFH_FAMILY_NEDELEC, FFH_ERR_INVALID,
                "the PML operator needs a Nedelec pattern");
    FFH_REQUIRE(kind == 0 || kind == 1, FFH_ERR_INVALID, "PML kind is 0 or 1");
    FFH_REQUIRE(k0 > 0.0, FFH_ERR_INVALID, "wave number must be positive");
    const int T = P->ncell, nl = P->n_local;
    const int64_t nnz = P->full.m.nnz;
    const int per = kind == 0 ? 3 : 9;
    ensure_ned_tables(P, true);
    std::vector<double> gb[2];
    {
        std::vector<double> h(std::max<int64_t>(nnz, 1));
        DBuf<double> vb, gbar;
        gbar.alloc((size_t)12 * T);
        for (int c = 0; c < 2; ++c) {
            for (int64_t k = 0; k < nnz; ++k) h[k] = values_bar[2 * k + c];
            vb.upload(h.data(), h.size(), s);
            hipLaunchKernelGGL(k_ned_term_bar, dim3((T + 3) / 4), dim3(256), 0, s, T, T, nl,
                               P->d_cls_c.p, P->d_ned64.p, P->d_slots_cm.p, vb.p, gbar.p);
            FFH_HIP(hipGetLastError());
            gb[c].resize((size_t)12 * T);
            gbar.download(gb[c].data(), gb[c].size(), s);
            FFH_HIP(hipStreamSynchronize(s)); // vb is rewritten for the imaginary part
        }
    }
    auto mul = [](const cd *A, const cd *B, cd *C) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                cd a = 0.0;
                for (int t = 0; t < 3; ++t) a += A[r + 3 * t] * B[t + 3 * c];
                C[r + 3 * c] = a;
            }
    };
    auto tr = [](const cd *A, cd *B) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[r + 3 * c] = A[c + 3 * r];
    };
    auto cj = [](const cd *A, cd *B) {
        for (int i = 0; i < 9; ++i) B[i] = std::conj(A[i]);
    };
    auto inv3 = [](const cd *A, cd *I, cd &det) {
        cd cof[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                cof[r + 3 * c] = A[r1 + 3 * c1] * A[r2 + 3 * c2] - A[r1 + 3 * c2] * A[r2 + 3 * c1];
            }
        det = A[0] * cof[0] + A[3] * cof[3] + A[6] * cof[6];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) I[r + 3 * c] = cof[c + 3 * r] / det;
    };
    // K, M of a stretch, and their tangents along dS (same formulas as the JVP)
    auto coeffs = [&](const cd *sv, const cd *dsv, cd *K, cd *M, cd *dK, cd *dM) {
        for (int i = 0; i < 9; ++i) K[i] = M[i] = dK[i] = dM[i] = 0.0;
        if (kind == 0) {
            for (int i = 0; i < 3; ++i) {
                const int j = (i + 1) % 3, k = (i + 2) % 3;
                const cd m = sv[j] * sv[k] / sv[i];
                const cd dm = m * (dsv[j] / sv[j] + dsv[k] / sv[k] - dsv[i] / sv[i]);
                M[4 * i] = m;
                dM[4 * i] = dm;
                K[4 * i] = 1.0 / m;
                dK[4 * i] = -dm / (m * m);
            }
            return;
        }
        cd Si[9], det, St[9], dSt[9], StS[9], a1[9], a2[9], SidS[9], dI[9], dIt[9], Sit[9], W[9];
        inv3(sv, Si, det);
        tr(sv, St);
        tr(dsv, dSt);
        mul(St, sv, StS);
        mul(Si, dsv, SidS);
        const cd ddet = det * (SidS[0] + SidS[4] + SidS[8]);
        mul(SidS, Si, dI);
        for (int i = 0; i < 9; ++i) dI[i] = -dI[i];
        tr(dI, dIt);
        tr(Si, Sit);
        mul(dSt, sv, a1);
        mul(St, dsv, a2);
        for (int i = 0; i < 9; ++i) {
            K[i] = StS[i] / det;
            dK[i] = (a1[i] + a2[i]) / det - K[i] * ddet / det;
        }
        mul(Si, Sit, W);
        mul(dI, Sit, a1);
        mul(Si, dIt, a2);
        for (int i = 0; i < 9; ++i) {
            M[i] = det * W[i];
            dM[i] = ddet * W[i] + det * (a1[i] + a2[i]);
        }
    };
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    std::fill(vertices_bar, vertices_bar + 3 * (size_t)P->nvert, 0.0);
    std::fill(stretch_bar, stretch_bar + 2 * (size_t)per * T, 0.0);
    double kbar = 0.0;
    for (int pos = 0; pos < T; ++pos) {
        const int e = P->perm[pos];
        const double *st = stretch + 2 * (size_t)per * e;
        cd sv[9], zero[9];
        for (int i = 0; i < per; ++i) {
            sv[i] = cd(st[2 * i], st[2 * i + 1]);
            zero[i] = 0.0;
        }
        if (kind == 0) {
            for (int i = 0; i < 3; ++i)
                FFH_REQUIRE(std::abs(sv[i]) > DBL_MIN, FFH_ERR_INVALID, "zero PML stretch");
        } else {
            cd Si[9], d;
            inv3(sv, Si, d);
            double mx = 0.0;
            for (int i = 0; i < 9; ++i) mx = std::max(mx, std::abs(sv[i]));
            const double sc = std::max(1.0, mx);
            FFH_REQUIRE(std::isfinite(d.real()) && std::isfinite(d.imag()) &&
                            std::abs(d) > 64.0 * DBL_EPSILON * sc * sc * sc,
                        FFH_ERR_INVALID, "singular PML stretch");
        }
        cd K[9], M[9], dK[9], dM[9];
        coeffs(sv, zero, K, M, dK, dM);
        cd Mp[9];
        for (int i = 0; i < 9; ++i) Mp[i] = -k0 * k0 * M[i];
        // term cotangents as (upper-triangle) 3x3 matrices
        cd Cb[9], Vb[9];
        for (int i = 0; i < 9; ++i) Cb[i] = Vb[i] = 0.0;
        for (int q = 0; q < 6; ++q) {
            const int i = pa[q] + 3 * pb[q];
            Cb[i] = cd(gb[0][(size_t)q * T + pos], gb[1][(size_t)q * T + pos]);
            Vb[i] = cd(gb[0][(size_t)(6 + q) * T + pos], gb[1][(size_t)(6 + q) * T + pos]);
        }
        const int *cv = P->cells.data() + 4 * (size_t)e;
        cd J[9], Ji[9], det;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r)
                J[r + 3 * c] = P->vertices[3 * (size_t)(cv[c + 1] - 1) + r] -
                               P->vertices[3 * (size_t)(cv[0] - 1) + r];
        inv3(J, Ji, det);
        FFH_REQUIRE(det.real() > 0.0, FFH_ERR_INVALID, "degenerate or inverted cell");
        cd Jt[9], Jit[9], tmp[9], tmp2[9];
        tr(J, Jt);
        tr(Ji, Jit);
        // K_bar = J Cb J^T / det ; M'_bar = det J^-T Vb J^-1
        cd Kb[9], Mpb[9];
        mul(J, Cb, tmp);
        mul(tmp, Jt, Kb);
        for (int i = 0; i < 9; ++i) Kb[i] /= det;
        mul(Jit, Vb, tmp);
        mul(tmp, Ji, Mpb);
        for (int i = 0; i < 9; ++i) Mpb[i] *= det;
        // J_bar from the curl terms: Re[(K* J Cb^T + K*^T J Cb)/det - tr(Cb^T C*) J^-T]
        cd Kc[9], Kct[9], Cbt[9], C0[9], Cc[9], Jb[9];
        cj(K, Kc);
        tr(Kc, Kct);
        tr(Cb, Cbt);
        mul(K, J, tmp);
        mul(Jt, tmp, C0);
        for (int i = 0; i < 9; ++i) C0[i] /= det;
        cj(C0, Cc);
        cd trc = 0.0;
        for (int i = 0; i < 9; ++i) trc += Cb[i] * Cc[i];
        mul(Kc, J, tmp);
        mul(tmp, Cbt, Jb);
        mul(Kct, J, tmp);
        mul(tmp, Cb, tmp2);
        for (int i = 0; i < 9; ++i) Jb[i] = (Jb[i] + tmp2[i]) / det - trc * Jit[i];
        // J_bar from the value terms: Re[det tr(Vb^T W) J^-T - J^-T D J^-T],
        // W = J^-1 M'* J^-T, D = det (Vb J^-1 M'*^T + Vb^T J^-1 M'*)
        cd Mc[9], Mct[9], W[9], Vbt[9], D[9];
        cj(Mp, Mc);
        tr(Mc, Mct);
        mul(Mc, Jit, tmp);
        mul(Ji, tmp, W);
        cd trv = 0.0;
        for (int i = 0; i < 9; ++i) trv += Vb[i] * W[i];
        tr(Vb, Vbt);
        mul(Vb, Ji, tmp);
        mul(tmp, Mct, D);
        mul(Vbt, Ji, tmp);
        mul(tmp, Mc, tmp2);
        for (int i = 0; i < 9; ++i) D[i] = det * (D[i] + tmp2[i]);
        mul(Jit, D, tmp);
        mul(tmp, Jit, tmp2);
        for (int i = 0; i < 9; ++i) Jb[i] += det * trv * Jit[i] - tmp2[i];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                const double g = Jb[r + 3 * c].real();
                vertices_bar[3 * (size_t)(cv[c + 1] - 1) + r] += g;
                vertices_bar[3 * (size_t)(cv[0] - 1) + r] -= g;
            }
        // wave number and the M cotangent
        cd Mb[9];
        for (int i = 0; i < 9; ++i) {
            kbar += (Mpb[i] * std::conj(-2.0 * k0 * M[i])).real();
            Mb[i] = -k0 * k0 * Mpb[i];
        }
        // stretch: sum over entries of K_bar conj(dK/dS_pq) + M_bar conj(dM/dS_pq)
        double *sb = stretch_bar + 2 * (size_t)per * e;
        for (int q = 0; q < per; ++q) {
            cd dir[9];
            for (int i = 0; i < per; ++i) dir[i] = i == q ? 1.0 : 0.0;
            cd K1[9], M1[9], dK1[9], dM1[9];
            coeffs(sv, dir, K1, M1, dK1, dM1);
            cd acc = 0.0;
            for (int i = 0; i < 9; ++i) acc += Kb[i] * std::conj(dK1[i]) + Mb[i] * std::conj(dM1[i]);
            sb[2 * q] += acc.real();
            sb[2 * q + 1] += acc.imag();
        }
    }
    *k0_bar = kbar;
}

// Nedelec vector load into rhs_full: tables once per problem, then g per
// cell, the oriented element vectors and a colour-by-colour scatter.
static void nedelec_load(Problem *P, int kind, const double *params, int64_t np)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    const int T = P->ncell, G = P->n_global, nl = P->n_local;
    if (!P->d_ned_load_tab.p) {
        std::vector<double> tab, pts;
        nedelec_load_tables(P->order, P->class_list, tab, pts);
        P->d_ned_load_tab.upload(tab, s);
        P->d_ned_load_pts.upload(pts, s);
        P->ned_load_nq = (int)(pts.size() / 4);
    }
    const int nq = P->ned_load_nq;
    double f[3] = {0.0, 0.0, 0.0};
    const double *samples = nullptr;
    if (kind == FFH_SOURCE_CONSTANT) {
        FFH_REQUIRE(params && np >= 3, FFH_ERR_INVALID, "constant vector load needs params[0..2]");
        for (int i = 0; i < 3; ++i) f[i] = params[i];
    } else if (kind == FFH_SOURCE_SAMPLES) {
        FFH_REQUIRE(params && np == 3ll * nq * T, FFH_ERR_INVALID,
                    "samples must be (3, nq, ncell) at the Nedelec load points");
        P->d_samples.upload(params, (size_t)np, s);
        samples = P->d_samples.p;
    } else {
        throw Error{FFH_ERR_INVALID, "Nedelec loads take a constant vector or samples"};
    }
    P->d_ned_g.alloc((size_t)T * nq * 3);
    P->d_elem.alloc(std::max<size_t>(P->d_elem.n, (size_t)nl * T));
    FFH_HIP(hipMemsetAsync(P->rhs_full.p, 0, sizeof(double) * G, s));
    hipLaunchKernelGGL(k_ned_load_g, dim3((T + kBlockA - 1) / kBlockA), dim3(kBlockA), 0, s, T, nq,
                       P->d_cells_c.p, P->d_vertices.p, P->d_perm.p, P->d_ned_load_pts.p, kind,
                       f[0], f[1], f[2], samples, P->d_ned_g.p, P->d_err.p);
    const int64_t work = (int64_t)T * nl;
    hipLaunchKernelGGL(k_ned_load_elem, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, T,
                       nl, 3 * nq, P->d_cls_c.p, P->d_ned_g.p, P->d_ned_load_tab.p, P->d_elem.p);
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        const int64_t w = (int64_t)cnt * nl;
        if (w == 0) continue;
        hipLaunchKernelGGL(k_scatter_vec_colour, dim3((int)((w + 255) / 256)), dim3(256), 0, s, p0,
                           cnt, T, nl, P->d_elem.p, P->d_gdofs_c.p, P->rhs_full.p);
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "Nedelec vector load");
    P->loaded = true;
    P->red_dirty = true;
}

void problem_load(Problem *P, int qdeg, int kind, const double *params, int64_t np)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
    const int T = P->ncell, G = P->n_global;
    if (kind == FFH_SOURCE_DOF_VECTOR) {
        FFH_REQUIRE(params && np == G, FFH_ERR_INVALID, "rhs vector must have n_global entries");
        P->rhs_full.upload(params, G, s);
        P->loaded = true;
        P->red_dirty = true;
        return;
    }
    if (P->family == FFH_FAMILY_P1_2D) {
        FFH_REQUIRE(kind == FFH_SOURCE_CONSTANT && params && np >= 1, FFH_ERR_INVALID,
                    "the 2D P1 load takes a constant coefficient");
        FFH_HIP(hipMemsetAsync(P->rhs_full.p, 0, sizeof(double) * G, s));
        for (int c = 0; c < P->n_colors; ++c) {
            const int p0 = P->color_start[c], p1 = P->color_start[c + 1];
            const int grid = (p1 - p0 + kBlockA - 1) / kBlockA;
            if (grid == 0) continue;
            hipLaunchKernelGGL(k_p1_2d_load, dim3(grid), dim3(kBlockA), 0, s, p0, p1,
                               P->d_cells_c.p, P->d_vertices.p, params[0], P->rhs_full.p,
                               P->d_err.p);
        }
        FFH_HIP(hipGetLastError());
        check_err(P, "2D P1 load");
        P->loaded = true;
        P->red_dirty = true;
        return;
    }
    if (P->family == FFH_FAMILY_NEDELEC) {
        nedelec_load(P, kind, params, np);
        return;
    }
    FFH_REQUIRE(P->family == FFH_FAMILY_H1, FFH_ERR_INVALID,
                "closed-form loads are implemented for the H1, Nedelec and 2D P1 families");
    FFH_REQUIRE(qdeg >= 0, FFH_ERR_INVALID, "quadrature degree must be >= 0");
    if (P->load_qdeg != qdeg) {
        std::vector<double> tab;
        int nq = 0;
        h1_load_table(P->order, qdeg, tab, nq);
        P->d_load_tab.upload(tab, s);
        P->load_qdeg = qdeg;
        P->load_nq = nq;
    }
    double p[4] = {0, 0, 0, 0};
    const double *samples = nullptr;
    if (kind == FFH_SOURCE_CONSTANT) {
        FFH_REQUIRE(params && np >= 1, FFH_ERR_INVALID, "constant load needs params[0]");
        p[0] = params[0];
    } else if (kind == FFH_SOURCE_SIN_PRODUCT) {
        FFH_REQUIRE(params && np >= 4, FFH_ERR_INVALID, "sin-product load needs 4 params");
        for (int i = 0; i < 4; ++i) p[i] = params[i];
    } else if (kind == FFH_SOURCE_SAMPLES) {
        FFH_REQUIRE(params && np == (int64_t)P->load_nq * T, FFH_ERR_INVALID,
                    "samples must be (nq, ncell)");
        P->d_samples.upload(params, (size_t)np, s);
        samples = P->d_samples.p;
    } else {
        throw Error{FFH_ERR_INVALID, "unknown source kind"};
    }
    FFH_HIP(hipMemsetAsync(P->rhs_full.p, 0, sizeof(double) * G, s));
    const int nl = P->n_local;
    P->d_elem.alloc(std::max<size_t>(P->d_elem.n, (size_t)nl * T));
    FFH_NL_DISPATCH(nl, {
        hipLaunchKernelGGL((k_h1_load<NL>), dim3((T + kBlockA - 1) / kBlockA), dim3(kBlockA),
                           0, s, T, P->d_cells_c.p, P->d_vertices.p, P->d_perm.p,
                           P->d_load_tab.p, P->load_nq, kind, p[0], p[1], p[2], p[3], samples,
                           P->d_elem.p, P->d_err.p);
    });
    for (int c = 0; c < P->n_colors; ++c) {
        const int p0 = P->color_start[c], cnt = P->color_start[c + 1] - p0;
        const int64_t work = (int64_t)cnt * nl;
        if (work == 0) continue;
        hipLaunchKernelGGL(k_scatter_vec_colour, dim3((int)((work + 255) / 256)), dim3(256), 0, s,
                           p0, cnt, T, nl, P->d_elem.p, P->d_gdofs_c.p, P->rhs_full.p);
    }
    FFH_HIP(hipGetLastError());
    check_err(P, "tetrahedral H1 load");
    P->loaded = true;
    P->red_dirty = true;
}

// Dirichlet: symbolic reduction (host, once) + numeric (device)
// Boundary faces of a tetrahedral mesh on the device: a face contained in
// exactly one cell (build_dirichlet_data,
// src/solvers/fortfem_tetra_lagrange_solver_3d.f90:587-597, and
// build_zero_tangential_mask, fortfem_tetra_nedelec_solver_3d.f90:442-462).
// Vertex -> cell lists by a counting sort, then one thread per (cell, face)
// looks among the cells around the face's smallest vertex for another one
// holding the other two vertices.  bf[4 e + o] = 1 for the boundary face of
// cell e (reference order) opposite its local vertex o.
__global__ void k_vertex_count(int T, const int4 *__restrict__ cells_c, int *__restrict__ cnt)
{
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= T) return;
    const int4 c = cells_c[pos];
    atomicAdd(&cnt[c.x], 1);
    atomicAdd(&cnt[c.y], 1);
    atomicAdd(&cnt[c.z], 1);
    atomicAdd(&cnt[c.w], 1);
}

__global__ void k_vertex_fill(int T, const int4 *__restrict__ cells_c, int *__restrict__ cursor,
                              int *__restrict__ list)
{
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= T) return;
    const int4 c = cells_c[pos];
    list[atomicAdd(&cursor[c.x], 1)] = pos;
    list[atomicAdd(&cursor[c.y], 1)] = pos;
    list[atomicAdd(&cursor[c.z], 1)] = pos;
    list[atomicAdd(&cursor[c.w], 1)] = pos;
}

__global__ void k_boundary_faces(int T, const int4 *__restrict__ cells_c,
                                 const int *__restrict__ perm, const int *__restrict__ start,
                                 const int *__restrict__ list, unsigned char *__restrict__ bf)
{
    const int64_t t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= 4ll * T) return;
    const int pos = (int)(t >> 2), o = (int)(t & 3);
    const int4 c = cells_c[pos];
    const int v[4] = {c.x, c.y, c.z, c.w};
    int f[3], k = 0;
    for (int i = 0; i < 4; ++i)
        if (i != o) f[k++] = v[i];
    int a = min(f[0], min(f[1], f[2]));
    int b = a == f[0] ? f[1] : f[0];
    int d = (a == f[0] || b == f[0]) && (a == f[1] || b == f[1]) ? f[2]
            : ((a == f[0] || b == f[0]) ? f[1] : f[0]);
    bool shared = false;
    for (int i = start[a]; i < start[a + 1] && !shared; ++i) {
        const int q = list[i];
        if (q == pos) continue;
        const int4 cq = cells_c[q];
        const bool hb = cq.x == b || cq.y == b || cq.z == b || cq.w == b;
        const bool hd = cq.x == d || cq.y == d || cq.z == d || cq.w == d;
        shared = hb && hd;
    }
    bf[4 * (size_t)perm[pos] + o] = shared ? 0 : 1;
}

static void boundary_faces(Problem *P, std::vector<unsigned char> &bf)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    const int T = P->ncell, nv = P->nvert;
    const int4 *cells = reinterpret_cast<const int4 *>(P->d_cells_c.p);
    DBuf<int> cnt, start, list;
    DBuf<unsigned char> dbf;
    cnt.alloc(nv);
    FFH_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * nv, s));
    const int g1 = (T + 255) / 256;
    hipLaunchKernelGGL(k_vertex_count, dim3(g1), dim3(256), 0, s, T, cells, cnt.p);
    FFH_HIP(hipGetLastError());
    std::vector<int> h(nv), hs(nv + 1);
    cnt.download(h.data(), nv, s);
    FFH_HIP(hipStreamSynchronize(s));
    hs[0] = 0;
    for (int i = 0; i < nv; ++i) hs[i + 1] = hs[i] + h[i];
    start.upload(hs, s);
    cnt.upload(hs.data(), nv, s); // fill cursors
    list.alloc(std::max<size_t>(1, (size_t)4 * T));
    hipLaunchKernelGGL(k_vertex_fill, dim3(g1), dim3(256), 0, s, T, cells, cnt.p, list.p);
    dbf.alloc((size_t)4 * T);
    hipLaunchKernelGGL(k_boundary_faces, dim3((unsigned)((4ll * T + 255) / 256)), dim3(256), 0, s,
                       T, cells, P->d_perm.p, start.p, list.p, dbf.p);
    FFH_HIP(hipGetLastError());
    bf.resize((size_t)4 * T);
    dbf.download(bf.data(), bf.size(), s);
    FFH_HIP(hipStreamSynchronize(s));
}

void problem_dirichlet(Problem *P, int mode, const double *values)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
    const int G = P->n_global;
    std::vector<int> cons;
    if (P->family == FFH_FAMILY_H1) {
        std::vector<unsigned char> bf;
        boundary_faces(P, bf);
        lagrange_boundary(P->order, P->ncell, bf, P->global_dofs, G, cons);
    } else if (P->family == FFH_FAMILY_NEDELEC) {
        // zero tangential trace (build_zero_tangential_mask,
        // src/solvers/fortfem_tetra_nedelec_solver_3d.f90:431-470)
        std::vector<unsigned char> bf, mask;
        boundary_faces(P, bf);
        nedelec_boundary_dofs(P->order, P->ncell, bf, P->global_dofs, mask);
        cons.assign(G, 0);
        for (int i = 0; i < G && i < (int)mask.size(); ++i) cons[i] = mask[i];
    } else if (P->family == FFH_FAMILY_P1_2D) {
        // boundary vertices: endpoints of edges owned by one triangle
        // (find_boundary, src/mesh/mesh_2d.f90:553-603)
        std::unordered_map<uint64_t, int> edges;
        edges.reserve((size_t)P->ncell * 3);
        for (int t = 0; t < P->ncell; ++t)
            for (int k = 0; k < 3; ++k) {
                int a = P->cells[3 * t + k], b = P->cells[3 * t + (k + 1) % 3];
                if (a > b) std::swap(a, b);
                edges[((uint64_t)a << 32) | (uint32_t)b]++;
            }
        cons.assign(G, 0);
        for (auto &kv : edges)
            if (kv.second == 1) {
                cons[(int)(kv.first >> 32) - 1] = 1;
                cons[(int)(kv.first & 0xffffffffu) - 1] = 1;
            }
    } else {
        throw Error{FFH_ERR_INVALID, "Dirichlet data not available for this family"};
    }
    P->constrained = cons;
    P->dirichlet_mode = mode;
    if (values) {
        P->d_bvals.upload(values, G, s);
    } else {
        P->d_bvals.reset();
    }
    if (mode == 1) {
        FFH_REQUIRE(P->family == FFH_FAMILY_P1_2D, FFH_ERR_INVALID,
                    "symmetric elimination (mode 1) is the 2D P1 path");
        P->have_dirichlet = true;
        P->red_dirty = true;
        return;
    }
    // free dofs ascending, reduced pattern keeps the original row order
    std::vector<int> free_dofs, free_index(G, -1);
    for (int i = 0; i < G; ++i)
        if (!cons[i]) {
            free_index[i] = (int)free_dofs.size();
            free_dofs.push_back(i);
        }
    const int nf = (int)free_dofs.size();
    std::vector<int> hcp(G + 1), hri(P->full.m.nnz);
    P->full.m.col_ptr.download(hcp.data(), G + 1, s);
    P->full.m.row_idx.download(hri.data(), hri.size(), s);
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<int> cnt(nf, 0), lcnt(nf, 0);
    parallel_for(nf, [&](int64_t lo, int64_t hi) {
        for (int64_t f = lo; f < hi; ++f) {
            int j = free_dofs[f], a = 0, l = 0;
            for (int k = hcp[j]; k < hcp[j + 1]; ++k) {
                if (cons[hri[k]]) ++l;
                else ++a;
            }
            cnt[f] = a;
            lcnt[f] = l;
        }
    });
    std::vector<int> rcp(nf + 1, 0), lptr(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
        rcp[f + 1] = rcp[f] + cnt[f];
        lptr[f + 1] = lptr[f] + lcnt[f];
    }
    std::vector<int> rri(rcp[nf]), orig(rcp[nf]), lk(lptr[nf]), lc(lptr[nf]);
    parallel_for(nf, [&](int64_t lo, int64_t hi) {
        for (int64_t f = lo; f < hi; ++f) {
            int j = free_dofs[f], o = rcp[f], l = lptr[f];
            for (int k = hcp[j]; k < hcp[j + 1]; ++k) {
                int r = hri[k];
                if (cons[r]) {
                    // A(j, r) = A(r, j) (exactly symmetric assembly)
                    lk[l] = k;
                    lc[l] = r;
                    ++l;
                } else {
                    rri[o] = free_index[r];
                    orig[o] = k;
                    ++o;
                }
            }
        }
    });
    DevMatrix &R = P->reduced.m;
    R.n = nf;
    R.nnz = rcp[nf];
    R.col_ptr.upload(rcp, s);
    R.row_idx.upload(rri, s);
    R.val.alloc(R.nnz);
    R.symmetric = true;
    R.values_changed();
    build_rowblocks(ctx, R, rcp);
    P->reduced.ctx = ctx;
    P->reduced.borrowed = true;
    invalidate_graphs(P->reduced.ws);
    P->d_red_orig.upload(orig, s);
    P->d_free_dofs.upload(free_dofs, s);
    P->d_lift_ptr.upload(lptr, s);
    P->d_lift_k.upload(lk, s);
    P->d_lift_c.upload(lc, s);
    P->rhs_red.alloc(nf);
    P->free_dofs = free_dofs;
    P->have_dirichlet = true;
    P->red_dirty = true;
    FFH_HIP(hipStreamSynchronize(s));
}

// numeric reduction of the current full matrix / rhs
void problem_reduce(Problem *P)
{
    if (!P->have_dirichlet || !P->red_dirty) return;
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    ensure_val64(ctx, P->full.m);
    if (P->dirichlet_mode == 1) {
        // P1_2D symmetric elimination + drop |a| <= 1e-14 (sparse_from_dense,
        // src/solvers/fortfem_sparse_matrix.f90:16-44)
        DevMatrix &A = P->full.m;
        const int n = A.n;
        DBuf<int> bnd;
        bnd.upload(P->constrained, s);
        DBuf<double> vals;
        vals.alloc(A.nnz);
        FFH_HIP(hipMemcpyAsync(vals.p, A.val.p, sizeof(double) * A.nnz,
                               hipMemcpyDeviceToDevice, s));
        P->rhs_red.alloc(n);
        double g = 0.0;
        if (P->d_bvals.p) {
            // constant boundary value of solve_compiled_p1_problem (bc%value)
            std::vector<double> hv(n);
            P->d_bvals.download(hv.data(), n, s);
            FFH_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < n; ++i)
                if (P->constrained[i]) {
                    g = hv[i];
                    break;
                }
        }
        hipLaunchKernelGGL(k_eliminate, dim3(grid1(ctx, n)), dim3(256), 0, s, n,
                           A.col_ptr.p, A.row_idx.p, vals.p, bnd.p, g, P->rhs_full.p,
                           P->rhs_red.p);
        hipLaunchKernelGGL(k_eliminate_vals, dim3(grid1(ctx, n)), dim3(256), 0, s, n,
                           A.col_ptr.p, A.row_idx.p, vals.p, bnd.p);
        FFH_HIP(hipGetLastError());
        // drop pattern (value dependent): compacted on the host side of the
        // symbolic stage from the device values
        std::vector<int> cp(n + 1), ri(A.nnz);
        std::vector<double> hv(A.nnz);
        A.col_ptr.download(cp.data(), n + 1, s);
        A.row_idx.download(ri.data(), A.nnz, s);
        vals.download(hv.data(), A.nnz, s);
        FFH_HIP(hipStreamSynchronize(s));
        std::vector<int> rcp(n + 1, 0), rri, orig;
        for (int j = 0; j < n; ++j) {
            for (int k = cp[j]; k < cp[j + 1]; ++k)
                if (std::fabs(hv[k]) > 1.0e-14) {
                    rri.push_back(ri[k]);
                    orig.push_back(k);
                }
            rcp[j + 1] = (int)rri.size();
        }
        DevMatrix &R = P->reduced.m;
        R.n = n;
        R.nnz = (int64_t)rri.size();
        R.col_ptr.upload(rcp, s);
        R.row_idx.upload(rri, s);
        R.val.alloc(R.nnz);
        R.symmetric = true;
        R.values_changed();
        build_rowblocks(ctx, R, rcp);
        P->reduced.ctx = ctx;
        P->reduced.borrowed = true;
        invalidate_graphs(P->reduced.ws);
        P->d_red_orig.upload(orig, s);
        hipLaunchKernelGGL(k_gather_vals, dim3(grid1(ctx, R.nnz)), dim3(256), 0, s, R.nnz,
                           P->d_red_orig.p, vals.p, R.val.p);
        FFH_HIP(hipGetLastError());
        FFH_HIP(hipStreamSynchronize(s));
        P->red_dirty = false;
        return;
    }
    DevMatrix &R = P->reduced.m;
    hipLaunchKernelGGL(k_gather_vals, dim3(grid1(ctx, R.nnz)), dim3(256), 0, s, R.nnz,
                       P->d_red_orig.p, P->full.m.val.p, R.val.p);
    hipLaunchKernelGGL(k_reduce_rhs, dim3(grid1(ctx, R.n)), dim3(256), 0, s, R.n,
                       P->d_free_dofs.p, P->rhs_full.p, P->d_lift_ptr.p, P->d_lift_k.p,
                       P->d_lift_c.p, P->full.m.val.p, P->d_bvals.p, P->rhs_red.p);
    FFH_HIP(hipGetLastError());
    R.values_changed();
    P->red_dirty = false;
}

void problem_solve(Problem *P, const ffh_solver_options &o, double *solution,
                   ffh_solver_stats &st)
{
    Context *ctx = P->ctx;
    hipStream_t s = ctx->stream;
    FFH_REQUIRE(P->assembled, FFH_ERR_INVALID, "assemble before solving");
    const int G = P->n_global;
    Matrix *M;
    const double *b;
    if (P->have_dirichlet) {
        problem_reduce(P);
        M = &P->reduced;
        b = P->rhs_red.p;
    } else {
        M = &P->full;
        b = P->rhs_full.p;
    }
    const int n = M->m.n;
    P->x_red.alloc(std::max(n, 1));
    FFH_HIP(hipMemsetAsync(P->x_red.p, 0, sizeof(double) * n, s));
    if (o.precision == FFH_MIXED)
        solve_mixed(ctx, *M, b, P->x_red.p, o, st);
    else
        solve_fp64(ctx, *M, b, P->x_red.p, o, st);
    P->x_full.alloc(G);
    if (P->have_dirichlet && P->dirichlet_mode == 0) {
        hipLaunchKernelGGL(k_scatter_solution, dim3(grid1(ctx, G)), dim3(256), 0, s, G,
                           (int)P->free_dofs.size(), P->d_free_dofs.p, P->x_red.p,
                           P->d_bvals.p, P->x_full.p);
        hipLaunchKernelGGL(k_scatter_free, dim3(grid1(ctx, n)), dim3(256), 0, s, n,
                           P->d_free_dofs.p, P->x_red.p, P->x_full.p);
        FFH_HIP(hipGetLastError());
    } else {
        FFH_HIP(hipMemcpyAsync(P->x_full.p, P->x_red.p, sizeof(double) * G,
                               hipMemcpyDeviceToDevice, s));
    }
    if (solution) P->x_full.download(solution, G, s);
    FFH_HIP(hipStreamSynchronize(s));
}

} // namespace ffh
