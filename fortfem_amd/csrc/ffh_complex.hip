// Complex sparse operators resident in HBM and their Krylov solve: the
// device replacement of the reference's retained complex factorisation
// (fortfem_complex_factor_csc / fortfem_complex_solve,
// src/bindings/fortfem_capi_solver.f90:10-124), whose operators here are the
// complex-symmetric PML curl-curl matrices (A = A^T, not Hermitian).
//
// Solve: Jacobi-preconditioned COCG (conjugate orthogonal conjugate
// gradient, van der Vorst and Melissen 1990), CG with the unconjugated
// bilinear form x^T y in place of the inner product, which needs A = A^T:
//   r = b - A x, z = D r, p = z, rho = r^T z;
//   q = A p, alpha = rho / p^T q, x += alpha p, r -= alpha q,
//   z = D r, rho' = r^T z, beta = rho' / rho, p = z + beta p,
// stopped when ||r||_2 <= tol ||b||_2 (the 2-norm with conjugation).
// Per iteration: the product with the p^T q partials, a one-block scalar
// step, the update with the r^T z and r^H r partials, a one-block scalar
// step and the direction; the host reads the control block every 16
// iterations.
#include <cmath>
#include <vector>

#include "ffh_internal.hpp"

#include "ffh_krylov.hpp"

namespace ffh {

struct ZMatrix {
    Context *ctx = nullptr;
    int n = 0;
    int64_t nnz = 0;
    DBuf<int> ptr, idx;        // CSR of the operator (rows gathered)
    DBuf<double2> val, dinv;   // complex values, Jacobi inverse diagonal
    DBuf<double2> x, r, z, p, q, b;
    DBuf<double> part;         // block partials (4 doubles per block)
    DBuf<double> ctl;          // control block
};

namespace {

constexpr int kZBlock = 256;
constexpr int kZLanes = 8; // lanes per row in the product

// control block layout (doubles)
enum {
    Z_RHO_RE, Z_RHO_IM, Z_ALPHA_RE, Z_ALPHA_IM, Z_BETA_RE, Z_BETA_IM, Z_RR, Z_BB, Z_TOL2,
    Z_ITER, Z_DONE, Z_CONV, Z_N
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], double *out)
{
    __shared__ double sh[NQ][kZBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double t = v[q];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) sh[q][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double t = 0.0;
        for (int i = 0; i < kZBlock / 64; ++i) t += sh[threadIdx.x][i];
        out[(size_t)blockIdx.x * 4 + threadIdx.x] = t;
    }
}

// q = A p (kZLanes lanes per row) and partials of p^T q
__global__ __launch_bounds__(kZBlock) void k_zspmv(int n, const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const double2 *__restrict__ val,
                                                   const double2 *__restrict__ p,
                                                   double2 *__restrict__ q,
                                                   double *__restrict__ part,
                                                   const double *__restrict__ ctl)
{
    if (ctl[Z_DONE] != 0.0) return;
    const int lane = threadIdx.x % kZLanes;
    const int rows = gridDim.x * (kZBlock / kZLanes);
    double acc[2] = {0.0, 0.0};
    for (int row = blockIdx.x * (kZBlock / kZLanes) + threadIdx.x / kZLanes; row < n;
         row += rows) {
        double2 s = make_double2(0.0, 0.0);
        for (int k = ptr[row] + lane; k < ptr[row + 1]; k += kZLanes) {
            const double2 a = val[k], v = p[idx[k]];
            s.x += a.x * v.x - a.y * v.y;
            s.y += a.x * v.y + a.y * v.x;
        }
        for (int off = kZLanes / 2; off > 0; off >>= 1) {
            s.x += __shfl_xor(s.x, off, kZLanes);
            s.y += __shfl_xor(s.y, off, kZLanes);
        }
        if (lane == 0) {
            q[row] = s;
            const double2 t = cmul(p[row], s);
            acc[0] += t.x;
            acc[1] += t.y;
        }
    }
    block_partials<2>(acc, part);
}

// one block: alpha = rho / p^T q (a zero p^T q stops the solve)
__global__ __launch_bounds__(kZBlock) void k_zalpha(int nb, const double *__restrict__ part,
                                                    double *__restrict__ ctl)
{
    if (ctl[Z_DONE] != 0.0) return;
    double v[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += kZBlock) {
        v[0] += part[4 * (size_t)i];
        v[1] += part[4 * (size_t)i + 1];
    }
    __shared__ double tot[2];
    __shared__ double sh[2][kZBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < 2; ++c) {
        double t = v[c];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) sh[c][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < 2; ++c) {
            tot[c] = 0.0;
            for (int i = 0; i < kZBlock / 64; ++i) tot[c] += sh[c][i];
        }
        const double2 mu = make_double2(tot[0], tot[1]);
        if (!(mu.x * mu.x + mu.y * mu.y > 0.0) || !isfinite(mu.x) || !isfinite(mu.y)) {
            ctl[Z_DONE] = 1.0; // breakdown
        } else {
            const double2 a = cdiv(make_double2(ctl[Z_RHO_RE], ctl[Z_RHO_IM]), mu);
            ctl[Z_ALPHA_RE] = a.x;
            ctl[Z_ALPHA_IM] = a.y;
        }
    }
}

// x += alpha p ; r -= alpha q ; z = D r ; partials of r^T z and r^H r
__global__ __launch_bounds__(kZBlock) void k_zupdate(int n, double2 *__restrict__ x,
                                                     double2 *__restrict__ r,
                                                     double2 *__restrict__ z,
                                                     const double2 *__restrict__ p,
                                                     const double2 *__restrict__ q,
                                                     const double2 *__restrict__ dinv,
                                                     double *__restrict__ part,
                                                     const double *__restrict__ ctl)
{
    if (ctl[Z_DONE] != 0.0) return;
    const double2 a = make_double2(ctl[Z_ALPHA_RE], ctl[Z_ALPHA_IM]);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kZBlock + threadIdx.x; i < n; i += gridDim.x * kZBlock) {
        const double2 ap = cmul(a, p[i]), aq = cmul(a, q[i]);
        double2 xi = x[i], ri = r[i];
        xi.x += ap.x;
        xi.y += ap.y;
        ri.x -= aq.x;
        ri.y -= aq.y;
        const double2 zi = cmul(dinv[i], ri);
        x[i] = xi;
        r[i] = ri;
        z[i] = zi;
        const double2 t = cmul(ri, zi);
        acc[0] += t.x;
        acc[1] += t.y;
        acc[2] += ri.x * ri.x + ri.y * ri.y;
    }
    block_partials<3>(acc, part);
}

// one block: rho' and ||r||^2 -> stop test, beta, iteration count
__global__ __launch_bounds__(kZBlock) void k_zbeta(int nb, const double *__restrict__ part,
                                                   double *__restrict__ ctl, int max_iter,
                                                   int first)
{
    if (ctl[Z_DONE] != 0.0) return;
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += kZBlock)
        for (int c = 0; c < 3; ++c) v[c] += part[4 * (size_t)i + c];
    __shared__ double sh[3][kZBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < 3; ++c) {
        double t = v[c];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) sh[c][w] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[3];
    for (int c = 0; c < 3; ++c) {
        tot[c] = 0.0;
        for (int i = 0; i < kZBlock / 64; ++i) tot[c] += sh[c][i];
    }
    const double2 rho = make_double2(tot[0], tot[1]);
    const double2 old = make_double2(ctl[Z_RHO_RE], ctl[Z_RHO_IM]);
    if (first) {
        ctl[Z_BB] = tot[2]; // ||b - A x0||^2 with x0 = 0
        ctl[Z_TOL2] = ctl[Z_TOL2] * tot[2];
    } else {
        ctl[Z_ITER] += 1.0;
    }
    ctl[Z_RR] = tot[2];
    if (tot[2] <= ctl[Z_TOL2]) {
        ctl[Z_CONV] = 1.0;
        ctl[Z_DONE] = 1.0;
        return;
    }
    if (ctl[Z_ITER] >= max_iter || !(rho.x * rho.x + rho.y * rho.y > 0.0)) {
        ctl[Z_DONE] = 1.0;
        return;
    }
    const double2 beta = first ? make_double2(0.0, 0.0) : cdiv(rho, old);
    ctl[Z_BETA_RE] = beta.x;
    ctl[Z_BETA_IM] = beta.y;
    ctl[Z_RHO_RE] = rho.x;
    ctl[Z_RHO_IM] = rho.y;
}

// p = z + beta p
__global__ __launch_bounds__(kZBlock) void k_zdir(int n, const double2 *__restrict__ z,
                                                  double2 *__restrict__ p,
                                                  const double *__restrict__ ctl)
{
    if (ctl[Z_DONE] != 0.0) return;
    const double2 b = make_double2(ctl[Z_BETA_RE], ctl[Z_BETA_IM]);
    for (int i = blockIdx.x * kZBlock + threadIdx.x; i < n; i += gridDim.x * kZBlock) {
        const double2 t = cmul(b, p[i]);
        p[i] = make_double2(z[i].x + t.x, z[i].y + t.y);
    }
}

// r = b (x0 = 0), z = D r, x = 0, partials of r^T z and r^H r
__global__ __launch_bounds__(kZBlock) void k_zinit(int n, const double2 *__restrict__ b,
                                                   double2 *__restrict__ x,
                                                   double2 *__restrict__ r,
                                                   double2 *__restrict__ z,
                                                   double2 *__restrict__ p,
                                                   const double2 *__restrict__ dinv,
                                                   double *__restrict__ part)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kZBlock + threadIdx.x; i < n; i += gridDim.x * kZBlock) {
        const double2 ri = b[i], zi = cmul(dinv[i], ri);
        x[i] = make_double2(0.0, 0.0);
        r[i] = ri;
        z[i] = zi;
        p[i] = make_double2(0.0, 0.0);
        const double2 t = cmul(ri, zi);
        acc[0] += t.x;
        acc[1] += t.y;
        acc[2] += ri.x * ri.x + ri.y * ri.y;
    }
    block_partials<3>(acc, part);
}

} // namespace



ZMatrix *zmatrix_create(Context *ctx, int n, int64_t nnz, int base, const int *col_ptr,
                        const int *row_idx, const double *values)
{
    FFH_REQUIRE(n >= 1 && nnz >= 1 && col_ptr && row_idx && values && (base == 0 || base == 1),
                FFH_ERR_INVALID, "invalid complex CSC");
    {
        // CSC -> CSR (the rows the product gathers), zero-based; Jacobi from
        // the diagonal (1 / a_ii, 1 where a_ii is zero)
        FFH_REQUIRE(col_ptr[0] == base && col_ptr[n] - base == nnz, FFH_ERR_INVALID,
                    "col_ptr does not match nnz");
        std::vector<int> rp(n + 1, 0), ci(nnz);
        std::vector<double2> rv(nnz), dg(n, make_double2(1.0, 0.0));
        for (int j = 0; j < n; ++j) {
            FFH_REQUIRE(col_ptr[j + 1] >= col_ptr[j], FFH_ERR_INVALID, "col_ptr not monotone");
            for (int k = col_ptr[j] - base; k < col_ptr[j + 1] - base; ++k) {
                const int i = row_idx[k] - base;
                FFH_REQUIRE(i >= 0 && i < n, FFH_ERR_INVALID, "row index out of range");
                rp[i + 1]++;
            }
        }
        for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int k = col_ptr[j] - base; k < col_ptr[j + 1] - base; ++k) {
                const int i = row_idx[k] - base, o = fill[i]++;
                ci[o] = j;
                rv[o] = make_double2(values[2 * k], values[2 * k + 1]);
                if (i == j) {
                    const double d = rv[o].x * rv[o].x + rv[o].y * rv[o].y;
                    if (d > 0.0) dg[i] = make_double2(rv[o].x / d, -rv[o].y / d);
                }
            }
        auto *z = new ZMatrix;
        z->ctx = ctx;
        z->n = n;
        z->nnz = nnz;
        hipStream_t s = ctx->stream;
        z->ptr.upload(rp, s);
        z->idx.upload(ci, s);
        z->val.upload(rv, s);
        z->dinv.upload(dg, s);
        FFH_HIP(hipStreamSynchronize(s));
        return z;
    }
}

void zmatrix_destroy(ZMatrix *m) { delete m; }

int zmatrix_size(const ZMatrix *m) { return m->n; }

void zmatrix_solve(ZMatrix *m, const double *b, double *x, double tolerance, int max_iterations,
                   int *iterations, double *relative_residual, int *converged)
{
    FFH_REQUIRE(b && x && tolerance > 0.0 && max_iterations >= 0, FFH_ERR_INVALID,
                "invalid complex solve arguments");
    {
        Context *ctx = m->ctx;
        hipStream_t s = ctx->stream;
        const int n = m->n;
        m->x.alloc(n), m->r.alloc(n), m->z.alloc(n), m->p.alloc(n), m->q.alloc(n),
            m->b.alloc(n);
        const int gv = (int)std::min<int64_t>((n + kZBlock - 1) / kZBlock, ctx->cu_count * 4);
        const int gs = (int)std::min<int64_t>(
            ((int64_t)n * kZLanes + kZBlock - 1) / kZBlock, ctx->cu_count * 4);
        const int nb = std::max(gv, gs);
        m->part.alloc(4 * (size_t)nb);
        m->ctl.alloc(Z_N);
        std::vector<double> c0(Z_N, 0.0);
        c0[Z_TOL2] = tolerance * tolerance;
        m->ctl.upload(c0, s);
        m->b.upload(reinterpret_cast<const double2 *>(b), n, s);
        hipLaunchKernelGGL(k_zinit, dim3(gv), dim3(kZBlock), 0, s, n, m->b.p, m->x.p, m->r.p,
                           m->z.p, m->p.p, m->dinv.p, m->part.p);
        hipLaunchKernelGGL(k_zbeta, dim3(1), dim3(kZBlock), 0, s, gv, m->part.p, m->ctl.p,
                           max_iterations, 1);
        hipLaunchKernelGGL(k_zdir, dim3(gv), dim3(kZBlock), 0, s, n, m->z.p, m->p.p, m->ctl.p);
        std::vector<double> hc(Z_N);
        for (;;) {
            for (int it = 0; it < 16; ++it) {
                hipLaunchKernelGGL(k_zspmv, dim3(gs), dim3(kZBlock), 0, s, n, m->ptr.p,
                                   m->idx.p, m->val.p, m->p.p, m->q.p, m->part.p, m->ctl.p);
                hipLaunchKernelGGL(k_zalpha, dim3(1), dim3(kZBlock), 0, s, gs, m->part.p,
                                   m->ctl.p);
                hipLaunchKernelGGL(k_zupdate, dim3(gv), dim3(kZBlock), 0, s, n, m->x.p, m->r.p,
                                   m->z.p, m->p.p, m->q.p, m->dinv.p, m->part.p, m->ctl.p);
                hipLaunchKernelGGL(k_zbeta, dim3(1), dim3(kZBlock), 0, s, gv, m->part.p,
                                   m->ctl.p, max_iterations, 0);
                hipLaunchKernelGGL(k_zdir, dim3(gv), dim3(kZBlock), 0, s, n, m->z.p, m->p.p,
                                   m->ctl.p);
            }
            FFH_HIP(hipGetLastError());
            m->ctl.download(hc.data(), Z_N, s);
            FFH_HIP(hipStreamSynchronize(s));
            if (hc[Z_DONE] != 0.0) break;
        }
        m->x.download(reinterpret_cast<double2 *>(x), n, s);
        FFH_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = (int)hc[Z_ITER];
        if (relative_residual)
            *relative_residual = hc[Z_BB] > 0.0 ? std::sqrt(hc[Z_RR] / hc[Z_BB]) : 0.0;
        if (converged) *converged = hc[Z_CONV] != 0.0;
    }
}

} // namespace ffh
