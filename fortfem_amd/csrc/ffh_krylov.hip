// ffh_krylov.hip -- SpMV and the Jacobi-PCG loop of solve_sparse on gfx950.
//
// Reference loop: src/solvers/advanced_solvers.f90 pcg_solve_operator
// (:831-857), initialize_pcg_state (:859-878), pcg_iteration_loop (:880-925),
// cg_iteration_loop (:752-797), compute_solver_tolerance (:1063-1073),
// finalize_solver_status (:1075-1086), Jacobi diagonal (:1124-1155).
//
// Design: one iteration is three launches on the context stream
//   (1) Ap = A p fused with the p.Ap reduction -> alpha     (HBM: matrix)
//   (2) x += alpha p, r -= alpha Ap, r.r and r.(D r)        -> beta, stop test
//   (3) p = D r + beta p
// Every reduction is per-block partials; the next kernel's blocks each re-sum
// the previous kernel's partials in block order (no atomics, deterministic)
// and derive alpha / beta / the stop test identically.  Scalars that a kernel
// both reads and updates (r.z) live in two slots indexed by iteration parity,
// so the one block that writes never races the blocks that read.  The host
// polls the control block once per chunk of iterations; a chunk may be
// replayed from a HIP graph.  Once the control block says done, later
// launches of the chunk return immediately, so the stopping iteration is
// exactly the reference's.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "ffh_internal.hpp"
#include "ffh_krylov.hpp"

namespace ffh {

constexpr int kBlock = 256;       // elementwise and reduction kernels
constexpr int kStreamBlock = 256; // CSR-stream SpMV
constexpr int kStreamNnz = 2048;  // nonzeros staged in LDS per row block

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over a block of BS threads; the result is valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double sh[BS / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; ++i) t += sh[i];
    return t;
}

// Total of partials[0 .. nb) in block order, broadcast to every thread.
template <int BS>
__device__ __forceinline__ double sum_partials(const double *__restrict__ partials, int nb)
{
    __shared__ double tot;
    double t = 0.0;
    for (int i = threadIdx.x; i < nb; i += BS) t += partials[i];
    t = block_sum<BS>(t);
    if (threadIdx.x == 0) tot = t;
    __syncthreads();
    const double r = tot;
    __syncthreads();
    return r;
}

__device__ __forceinline__ bool is_breakdown(double pap, double rz, int mode)
{
    if (mode == 1) return !(fabs(pap) > 1.0e-14 * fabs(rz));
    return fabs(pap) < 1.0e-14;
}

// ------------------------------------------------------------ CSR stream --
// A row block is a run of consecutive rows whose nonzeros fit in LDS.  The
// block streams the run's values and column indices with fully coalesced
// loads (each lane takes every BS-th nonzero), multiplies by the gathered x
// into LDS, then one thread per row adds its row's products in column order.
// A row longer than the LDS stage is summed by the whole block.
// Epilogue E receives (row, A x row value) and accumulates up to two partial
// sums.
template <typename MT, typename VT, typename E>
__device__ __forceinline__ void csr_stream(const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val,
                                           const VT *__restrict__ x,
                                           const int *__restrict__ rowblocks, int nrb,
                                           E &epi, double &d0, double &d1)
{
    constexpr int BS = kStreamBlock, NZ = kStreamNnz, U = NZ / BS;
    __shared__ VT prod[NZ];
    for (int blk = blockIdx.x; blk < nrb; blk += gridDim.x) {
        const int r0 = rowblocks[blk], r1 = rowblocks[blk + 1];
        const int k0 = ptr[r0], k1 = ptr[r1];
        const int cnt = k1 - k0;
        if (cnt <= NZ) {
            int c[U];
            MT v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = threadIdx.x + u * BS;
                if (k < cnt) {
                    c[u] = __builtin_nontemporal_load(idx + k0 + k);
                    v[u] = __builtin_nontemporal_load(val + k0 + k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = threadIdx.x + u * BS;
                if (k < cnt) prod[k] = (VT)v[u] * x[c[u]];
            }
            __syncthreads();
            const int nrows = r1 - r0;
            if (2 * nrows > BS) {
                for (int r = r0 + threadIdx.x; r < r1; r += BS) {
                    const int b = ptr[r] - k0, e = ptr[r + 1] - k0;
                    VT s = VT(0);
                    for (int k = b; k < e; ++k) s += prod[k];
                    epi(r, s, d0, d1);
                }
            } else {
                // gs lanes per row (a power of two, 2..64, gs * nrows <= BS):
                // every row of the block is summed in one pass
                int gs = 2;
                while (gs < 64 && 2 * gs * nrows <= BS) gs *= 2;
                const int lane = threadIdx.x & (gs - 1);
                const int r = r0 + threadIdx.x / gs;
                VT s = VT(0);
                if (r < r1) {
                    const int b = ptr[r] - k0, e = ptr[r + 1] - k0;
                    for (int k = b + lane; k < e; k += gs) s += prod[k];
                }
                for (int off = gs >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, gs);
                if (lane == 0 && r < r1) epi(r, s, d0, d1);
            }
            __syncthreads();
        } else {
            // one long row (r1 == r0 + 1)
            VT s = VT(0);
            for (int k = k0 + threadIdx.x; k < k1; k += BS) s += (VT)val[k] * x[idx[k]];
            double t = block_sum<BS>((double)s);
            __shared__ double tot;
            if (threadIdx.x == 0) tot = t;
            __syncthreads();
            if (threadIdx.x == 0) epi(r0, (VT)tot, d0, d1);
            __syncthreads();
        }
    }
}

// Vector CSR: a group of G lanes per row (grid-stride over rows); each lane
// issues two independent (index, value) loads before its gathers, then the
// group reduces by shuffles.  Used when it beats the stream kernel for a
// matrix (chosen by the autotuner below).
template <int G, typename MT, typename VT, typename E>
__device__ __forceinline__ void csr_vector(int n, const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val,
                                           const VT *__restrict__ x, E &epi, double &d0,
                                           double &d1)
{
    constexpr int RPB = kStreamBlock / G;
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * RPB;
    for (int row = blockIdx.x * RPB + threadIdx.x / G; row < n; row += ngrp) {
        const int b = ptr[row], e = ptr[row + 1];
        VT s = VT(0);
        int k = b + lane;
        for (; k + G < e; k += 2 * G) {
            const int c0 = __builtin_nontemporal_load(idx + k);
            const int c1 = __builtin_nontemporal_load(idx + k + G);
            const MT v0 = __builtin_nontemporal_load(val + k);
            const MT v1 = __builtin_nontemporal_load(val + k + G);
            s += (VT)v0 * x[c0];
            s += (VT)v1 * x[c1];
        }
        if (k < e) s += (VT)val[k] * x[idx[k]];
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
        if (lane == 0) epi(row, s, d0, d1);
    }
}

// SELL-64: one wave per slice of 64 rows, one lane per row.  Every index
// and value load of a wave reads 64 consecutive words; each lane adds its
// row's products in column order (the CSR order).  Entries go in steps of
// eight: the sixteen index/value loads of a step are issued together, then
// the eight gathers, so a wave has many loads in flight.  Here ptr is the
// slice offset array, idx/val the SELL entries and nrb the slice count.
template <typename MT, typename VT, typename E>
__device__ __forceinline__ void sell_rows(int n, const int *__restrict__ off,
                                          const int *__restrict__ sidx,
                                          const MT *__restrict__ sval,
                                          const VT *__restrict__ x, int nslice, E &epi,
                                          double &d0, double &d1)
{
    constexpr int WPB = kStreamBlock / 64, U = 8;
    const int lane = threadIdx.x & 63;
    for (int sl = blockIdx.x * WPB + (int)(threadIdx.x >> 6); sl < nslice;
         sl += gridDim.x * WPB) {
        const int o = off[sl];
        const int w = (off[sl + 1] - o) >> 6;
        const int *ip = sidx + o + lane;
        const MT *vp = sval + o + lane;
        VT s = VT(0);
        int j = 0;
        for (; j + U <= w; j += U) {
            int c[U];
            MT v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = __builtin_nontemporal_load(ip + 64 * u);
                v[u] = __builtin_nontemporal_load(vp + 64 * u);
            }
            VT xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) s += (VT)v[u] * xv[u];
            ip += 64 * U;
            vp += 64 * U;
        }
        if (j < w) {
            // the last w - j (< U) entries, loads first
            int c[U];
            MT v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) {
                    c[u] = __builtin_nontemporal_load(ip + 64 * u);
                    v[u] = __builtin_nontemporal_load(vp + 64 * u);
                }
            VT xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) s += (VT)v[u] * xv[u];
        }
        const int row = sl * 64 + lane;
        if (row < n) epi(row, s, d0, d1);
    }
}

// Row engine of a launch: V = 0 stream, V = 1 SELL-64, otherwise vector CSR
// with G = V lanes.
template <int V, typename MT, typename VT, typename E>
__device__ __forceinline__ void row_engine(int n, const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val,
                                           const VT *__restrict__ x,
                                           const int *__restrict__ rowblocks, int nrb, E &epi,
                                           double &d0, double &d1)
{
    if constexpr (V == 0)
        csr_stream<MT, VT>(ptr, idx, val, x, rowblocks, nrb, epi, d0, d1);
    else if constexpr (V == 1)
        sell_rows<MT, VT>(n, ptr, idx, val, x, nrb, epi, d0, d1);
    else
        csr_vector<V, MT, VT>(n, ptr, idx, val, x, epi, d0, d1);
}

template <typename VT>
struct EpiStore {
    VT *y;
    __device__ void operator()(int r, VT s, double &, double &) const { y[r] = s; }
};

template <typename VT>
struct EpiAp { // ap = A p ; p.Ap
    const VT *p;
    VT *ap;
    __device__ void operator()(int r, VT s, double &d0, double &) const
    {
        ap[r] = s;
        d0 += (double)p[r] * (double)s;
    }
};

template <typename VT>
struct EpiInit { // r = b - A x ; z = D r ; p = z ; r.r, r.z
    const VT *b, *dinv;
    VT *r, *p;
    __device__ void operator()(int i, VT s, double &d0, double &d1) const
    {
        VT ri = b[i] - s;
        VT zi = dinv[i] * ri;
        r[i] = ri;
        p[i] = zi;
        d0 += (double)ri * (double)ri;
        d1 += (double)ri * (double)zi;
    }
};

struct EpiResidual { // r = b - A x ; r.r
    const double *b;
    double *r;
    __device__ void operator()(int i, double s, double &d0, double &) const
    {
        double ri = b[i] - s;
        r[i] = ri;
        d0 += ri * ri;
    }
};

// --------------------------------------------------------------- kernels --
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_spmv(int n, const int *__restrict__ ptr,
                                                       const int *__restrict__ idx,
                                                       const MT *__restrict__ val,
                                                       const VT *__restrict__ x,
                                                       VT *__restrict__ y,
                                                       const int *__restrict__ rowblocks,
                                                       int nrb)
{
    EpiStore<VT> epi{y};
    double d0 = 0.0, d1 = 0.0;
    row_engine<V, MT, VT>(n, ptr, idx, val, x, rowblocks, nrb, epi, d0, d1);
}

// r = b - A x ; z = D r ; p = z ; partials of r.r and r.z
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_pcg_init(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx, const MT *__restrict__ val,
    const VT *__restrict__ x, const VT *__restrict__ b, VT *__restrict__ r,
    VT *__restrict__ p, const VT *__restrict__ dinv, const int *__restrict__ rowblocks,
    int nrb, double *__restrict__ partials, int stride)
{
    EpiInit<VT> epi{b, dinv, r, p};
    double d0 = 0.0, d1 = 0.0;
    row_engine<V, MT, VT>(n, ptr, idx, val, x, rowblocks, nrb, epi, d0, d1);
    const double s0 = block_sum<kStreamBlock>(d0);
    const double s1 = block_sum<kStreamBlock>(d1);
    if (threadIdx.x == 0) {
        partials[stride + blockIdx.x] = s0;
        partials[2 * stride + blockIdx.x] = s1;
    }
}

// initial norm, tolerance (compute_solver_tolerance, :1063-1073), r.z
__global__ __launch_bounds__(kBlock) void k_pcg_init_finish(const double *__restrict__ partials,
                                                            int stride, int nb, KrylovCtl *ctl)
{
    const double trr = sum_partials<kBlock>(partials + stride, nb);
    const double trz = sum_partials<kBlock>(partials + 2 * stride, nb);
    if (threadIdx.x == 0) {
        const double init = sqrt(trr);
        ctl->rr = trr;
        ctl->initial_norm = init;
        ctl->residual_norm = init;
        ctl->tol = ctl->tol_type == 1 ? ctl->tolerance : ctl->tolerance * init;
        ctl->rzs[0] = trz;
        ctl->iter = 0;
        ctl->done = ctl->max_iter <= 0 ? 1 : 0;
    }
}

// Ap = A p ; partials of p.Ap
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_pcg_spmv(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx, const MT *__restrict__ val,
    const VT *__restrict__ p, VT *__restrict__ ap, const int *__restrict__ rowblocks,
    int nrb, double *__restrict__ partials, const KrylovCtl *ctl)
{
    if (ctl->done) return;
    EpiAp<VT> epi{p, ap};
    double d0 = 0.0, d1 = 0.0;
    row_engine<V, MT, VT>(n, ptr, idx, val, p, rowblocks, nrb, epi, d0, d1);
    const double s0 = block_sum<kStreamBlock>(d0);
    if (threadIdx.x == 0) partials[blockIdx.x] = s0;
}

// p.Ap -> alpha (or breakdown); x += alpha p ; r -= alpha Ap ; partials of
// r.r and r.(D r)
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_update(
    int n, VT *__restrict__ x, VT *__restrict__ r, const VT *__restrict__ p,
    const VT *__restrict__ ap, const VT *__restrict__ dinv,
    double *__restrict__ partials, int stride, int nb_spmv, int parity, KrylovCtl *ctl)
{
    if (ctl->done) return;
    const double pap = sum_partials<kBlock>(partials, nb_spmv);
    const double rz_old = ctl->rzs[parity];
    if (is_breakdown(pap, rz_old, ctl->breakdown_mode)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->pap = pap;
            ctl->done = 1;
            ctl->breakdown_hit = 1;
        }
        return;
    }
    const double alpha_d = rz_old / pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->pap = pap;
        ctl->alpha = alpha_d;
    }
    const VT alpha = (VT)alpha_d;
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = x[i] + alpha * p[i];
        VT ri = r[i] - alpha * ap[i];
        r[i] = ri;
        VT zi = dinv[i] * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    const double s0 = block_sum<kBlock>(rr);
    const double s1 = block_sum<kBlock>(rz);
    if (threadIdx.x == 0) {
        partials[stride + blockIdx.x] = s0;
        partials[2 * stride + blockIdx.x] = s1;
    }
}

// stop test on ||r|| ; beta = rz_new / rz_old ; p = D r + beta p
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, VT *__restrict__ p,
                                                    const VT *__restrict__ r,
                                                    const VT *__restrict__ dinv,
                                                    const double *__restrict__ partials,
                                                    int stride, int nb_update, int parity,
                                                    KrylovCtl *ctl)
{
    if (ctl->done) return;
    const double trr = sum_partials<kBlock>(partials + stride, nb_update);
    const double trz = sum_partials<kBlock>(partials + 2 * stride, nb_update);
    const double rn = sqrt(trr);
    const bool conv = rn <= ctl->tol;
    const double rz_old = ctl->rzs[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int it = ctl->iter + 1;
        ctl->iter = it;
        ctl->rr = trr;
        ctl->residual_norm = rn;
        if (conv) {
            ctl->converged = 1;
            ctl->done = 1;
        } else {
            ctl->rzs[parity ^ 1] = trz;
            ctl->beta = trz / rz_old;
            if (it >= ctl->max_iter) ctl->done = 1;
        }
    }
    if (conv) return;
    const VT beta = (VT)(trz / rz_old);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        p[i] = dinv[i] * r[i] + beta * p[i];
}

// Jacobi diagonal (build_sparse_preconditioner 'jacobi', :1130-1144):
// first entry of row i with column i; 1/a_ii if |a_ii| > 1e-14 else 1.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const double *__restrict__ val,
                                                   int use, VT *__restrict__ dinv)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double a = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (idx[k] == i) {
                a = val[k];
                break;
            }
        dinv[i] = (use && fabs(a) > 1.0e-14) ? (VT)(1.0 / a) : (VT)1;
    }
}

// r = b - A x (FP64) and partials of ||r||^2 -- the true residual of the
// refinement loop and of the final FP64 report.
template <int V>
__global__ __launch_bounds__(kStreamBlock) void k_residual(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    const int *__restrict__ rowblocks, int nrb, double *__restrict__ partials)
{
    EpiResidual epi{b, r};
    double d0 = 0.0, d1 = 0.0;
    row_engine<V, double, double>(n, ptr, idx, val, x, rowblocks, nrb, epi, d0, d1);
    const double s0 = block_sum<kStreamBlock>(d0);
    if (threadIdx.x == 0) partials[blockIdx.x] = s0;
}

__global__ __launch_bounds__(kBlock) void k_residual_finish(const double *__restrict__ partials,
                                                            int nb, KrylovCtl *ctl)
{
    const double t = sum_partials<kBlock>(partials, nb);
    if (threadIdx.x == 0) ctl->rr = t;
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ a,
                                                   float *__restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

// SELL-64 construction: the longest row of every slice, then the entries
// (padding: zero value on the row's own column, source position -1).
__global__ __launch_bounds__(kBlock) void k_sell_width(int n, const int *__restrict__ ptr,
                                                       int nslice, int *__restrict__ width)
{
    const int sl = blockIdx.x * kBlock + threadIdx.x;
    if (sl >= nslice) return;
    const int r0 = sl * 64, r1 = min(n, r0 + 64);
    int w = 0;
    for (int r = r0; r < r1; ++r) w = max(w, ptr[r + 1] - ptr[r]);
    width[sl] = w;
}

__global__ __launch_bounds__(kBlock) void k_sell_fill(int n, const int *__restrict__ ptr,
                                                      const int *__restrict__ idx, int nslice,
                                                      const int *__restrict__ off,
                                                      int *__restrict__ sidx,
                                                      int *__restrict__ src)
{
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (t >= (int64_t)nslice * 64) return;
    const int row = (int)t, sl = row >> 6, lane = row & 63;
    const int o = off[sl], w = (off[sl + 1] - o) >> 6;
    const int b = row < n ? ptr[row] : 0;
    const int len = row < n ? ptr[row + 1] - b : 0;
    const int pad = row < n ? row : 0;
    for (int j = 0; j < w; ++j) {
        const int e = o + 64 * j + lane;
        if (j < len) {
            sidx[e] = idx[b + j];
            src[e] = b + j;
        } else {
            sidx[e] = pad;
            src[e] = -1;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_values(int64_t total, const int *__restrict__ src,
                                                        const T *__restrict__ val,
                                                        T *__restrict__ sval)
{
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kBlock) {
        const int k = src[e];
        sval[e] = k >= 0 ? val[k] : T(0);
    }
}

// ---------------------------------------------------------------- driver --
static int grid_cap(const Context *ctx) { return ctx->cu_count * 4; }

static int grid_for(const Context *ctx, int64_t items, int per_block)
{
    int64_t need = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, grid_cap(ctx)));
}

// SpMV variants: engine (0 = CSR stream, G = vector CSR with G lanes per row)
// and a grid multiplier over the base cap of 4 blocks per CU.
struct SpmvVariant {
    int engine, gmul;
};
static const SpmvVariant kVariants[] = {{0, 1}, {0, 2}, {0, 4}, {4, 2}, {8, 1},
                                        {8, 2}, {16, 2}, {32, 2}, {32, 4}, {1, 1},
                                        {1, 2}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static int variant_grid(const Context *ctx, const DevMatrix &A, int vi)
{
    const SpmvVariant v = kVariants[vi];
    const int64_t cap = (int64_t)grid_cap(ctx) * v.gmul;
    int64_t need;
    if (v.engine == 0)
        need = A.nrb;
    else if (v.engine == 1)
        need = ((int64_t)A.nslice + kStreamBlock / 64 - 1) / (kStreamBlock / 64);
    else
        need = ((int64_t)A.n * v.engine + kStreamBlock - 1) / kStreamBlock;
    return (int)std::max<int64_t>(1, std::min(need, cap));
}

// Gather arrays of a launch: the CSR (CSC read as CSR) arrays with the row
// blocks of the stream engine, or the SELL-64 arrays with the slice count.
template <typename MT>
struct GArgs {
    const int *ptr, *idx;
    const MT *val;
    const int *rb;
    int nrb;
};

template <typename MT>
static GArgs<MT> gargs(const DevMatrix &A, int vi)
{
    if (kVariants[vi].engine == 1) {
        if constexpr (sizeof(MT) == 8)
            return {A.sell_off.p, A.sell_idx.p, A.sell_val.p, nullptr, A.nslice};
        else
            return {A.sell_off.p, A.sell_idx.p, A.sell_val32.p, nullptr, A.nslice};
    }
    if constexpr (sizeof(MT) == 8)
        return {A.gptr(), A.gidx(), A.gval(), A.rowblocks.p, A.nrb};
    else
        return {A.gptr(), A.gidx(), A.gval32(), A.rowblocks.p, A.nrb};
}

#define FFH_ENGINE(eng, ...)                                                  \
    switch (eng) {                                                           \
    case 0: { constexpr int VV = 0; __VA_ARGS__; } break;                     \
    case 1: { constexpr int VV = 1; __VA_ARGS__; } break;                     \
    case 4: { constexpr int VV = 4; __VA_ARGS__; } break;                     \
    case 8: { constexpr int VV = 8; __VA_ARGS__; } break;                     \
    case 16: { constexpr int VV = 16; __VA_ARGS__; } break;                   \
    default: { constexpr int VV = 32; __VA_ARGS__; } break;                   \
    }

// Row blocks of the CSR-stream kernels: consecutive rows whose nonzeros fit
// the LDS stage; a longer row gets a block of its own.
void build_rowblocks(Context *ctx, DevMatrix &A, const std::vector<int> &ptr)
{
    std::vector<int> rb;
    rb.reserve(ptr.size() / 16 + 2);
    const int n = (int)ptr.size() - 1;
    int r = 0;
    rb.push_back(0);
    while (r < n) {
        int r1 = r + 1;
        while (r1 < n && ptr[r1 + 1] - ptr[r] <= kStreamNnz) ++r1;
        rb.push_back(r1);
        r = r1;
    }
    if (n == 0) rb.push_back(0);
    A.nrb = (int)rb.size() - 1;
    A.rowblocks.upload(rb, ctx->stream);
    A.variant[0] = A.variant[1] = -1;
    A.sell_built = A.sell_ok = false;
    A.sell64_valid = A.sell32_valid = false;
    FFH_HIP(hipStreamSynchronize(ctx->stream));
}

Workspace::~Workspace()
{
    free_solo(solo);
    for (auto &g : graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (host_ctl) (void)hipHostFree(host_ctl);
}

static int partial_stride(const Context *ctx) { return grid_cap(ctx) * 4; }

void ensure_workspace(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int n = M.m.n;
    if (w.n != n) invalidate_graphs(w);
    w.n = n;
    w.r.alloc(n), w.p.alloc(n), w.ap.alloc(n), w.x.alloc(n), w.b.alloc(n),
        w.dinv.alloc(n);
    if (!w.partials.p) {
        w.partials.alloc(3 * (size_t)partial_stride(ctx));
        FFH_HIP(hipMemsetAsync(w.partials.p, 0, sizeof(double) * w.partials.n, ctx->stream));
    }
    w.ctl.alloc(1);
    if (!w.host_ctl)
        FFH_HIP(hipHostMalloc((void **)&w.host_ctl, sizeof(KrylovCtl),
                              hipHostMallocDefault));
}

static void ensure_workspace32(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int n = M.m.n;
    w.r32.alloc(n), w.p32.alloc(n), w.ap32.alloc(n), w.x32.alloc(n),
        w.dinv32.alloc(n);
    (void)ctx;
}

void ensure_val32(Context *ctx, DevMatrix &A)
{
    if (A.val32_valid) return;
    A.val32.alloc(A.nnz);
    int grid = grid_for(ctx, A.nnz, kBlock);
    hipLaunchKernelGGL(k_to_f32, dim3(grid), dim3(kBlock), 0, ctx->stream, A.nnz,
                       A.val.p, A.val32.p);
    if (!A.symmetric) {
        A.t_val32.alloc(A.nnz);
        hipLaunchKernelGGL(k_to_f32, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           A.nnz, A.t_val.p, A.t_val32.p);
    }
    FFH_HIP(hipGetLastError());
    A.val32_valid = true;
}

// SELL-64 structure of the gather arrays.  Not built (sell_ok false, the
// SELL variants fall back to the stream engine) when the padding would more
// than triple the stored entries or the entry count exceeds int32.
static void ensure_sell_structure(Context *ctx, DevMatrix &A)
{
    if (A.sell_built) return;
    A.sell_built = true;
    A.sell_ok = false;
    const int n = A.n;
    if (n == 0 || A.nnz == 0) return;
    const int nslice = (n + 63) / 64;
    hipStream_t s = ctx->stream;
    DBuf<int> width;
    width.alloc(nslice);
    hipLaunchKernelGGL(k_sell_width, dim3((nslice + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                       A.gptr(), nslice, width.p);
    FFH_HIP(hipGetLastError());
    std::vector<int> w(nslice);
    width.download(w.data(), nslice, s);
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<int> off(nslice + 1, 0);
    int64_t total = 0;
    for (int i = 0; i < nslice; ++i) {
        total += 64ll * w[i];
        if (total > INT32_MAX) return;
        off[i + 1] = (int)total;
    }
    if (total > 3 * A.nnz + (1 << 20)) return;
    A.sell_off.upload(off, s);
    A.sell_idx.alloc(std::max<int64_t>(total, 1));
    A.sell_src.alloc(std::max<int64_t>(total, 1));
    hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)(((int64_t)nslice * 64 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, n, A.gptr(), A.gidx(), nslice, A.sell_off.p,
                       A.sell_idx.p, A.sell_src.p);
    FFH_HIP(hipGetLastError());
    FFH_HIP(hipStreamSynchronize(s));
    A.nslice = nslice;
    A.sell_total = total;
    A.sell_ok = true;
    A.sell64_valid = A.sell32_valid = false;
}

// SELL values of one precision, refreshed after the values change.
static void ensure_sell_values(Context *ctx, DevMatrix &A, bool f32)
{
    ensure_sell_structure(ctx, A);
    if (!A.sell_ok) return;
    const int grid = grid_for(ctx, A.sell_total, kBlock);
    if (f32) {
        if (A.sell32_valid) return;
        ensure_val32(ctx, A);
        A.sell_val32.alloc(A.sell_total);
        hipLaunchKernelGGL(k_sell_values<float>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           A.sell_total, A.sell_src.p, A.gval32(), A.sell_val32.p);
        A.sell32_valid = true;
    } else {
        if (A.sell64_valid) return;
        ensure_val64(ctx, A);
        A.sell_val.alloc(A.sell_total);
        hipLaunchKernelGGL(k_sell_values<double>, dim3(grid), dim3(kBlock), 0, ctx->stream,
                           A.sell_total, A.sell_src.p, A.gval(), A.sell_val.p);
        A.sell64_valid = true;
    }
    FFH_HIP(hipGetLastError());
}

void ensure_sell(Context *ctx, DevMatrix &A, bool f32) { ensure_sell_values(ctx, A, f32); }

bool sell_structure(Context *ctx, DevMatrix &A)
{
    ensure_sell_structure(ctx, A);
    return A.sell_ok;
}

__global__ __launch_bounds__(kBlock) void k_sell_values2(int64_t total, const int *__restrict__ src,
                                                         const double *__restrict__ val,
                                                         double *__restrict__ o64,
                                                         float *__restrict__ o32)
{
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kBlock) {
        const int k = src[e];
        const double v = k >= 0 ? val[k] : 0.0;
        if (o64) o64[e] = v;
        if (o32) o32[e] = (float)v;
    }
}

void sell_values(Context *ctx, const DevMatrix &A, const double *val, double *out64, float *out32)
{
    FFH_REQUIRE(A.sell_ok, FFH_ERR_INVALID, "matrix has no SELL-64 copy");
    hipLaunchKernelGGL(k_sell_values2, dim3(grid_for(ctx, A.sell_total, kBlock)),
                       dim3(kBlock), 0, ctx->stream, A.sell_total, A.sell_src.p, val, out64, out32);
    FFH_HIP(hipGetLastError());
}

// The variant to launch: vi itself with its arrays current, or the stream
// engine when vi is a SELL variant and the matrix has no SELL copy.
static int ready_variant(Context *ctx, DevMatrix &A, int vi, bool f32)
{
    if (kVariants[vi].engine != 1) return vi;
    ensure_sell_values(ctx, A, f32);
    return A.sell_ok ? vi : 0;
}

// y = A x with variant vi (its arrays current: see ready_variant)
template <typename MT, typename VT>
static void launch_spmv(Context *ctx, const DevMatrix &A, const VT *x, VT *y, int vi)
{
    const int grid = variant_grid(ctx, A, vi);
    const GArgs<MT> g = gargs<MT>(A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_spmv<VV, MT, VT>), dim3(grid), dim3(kStreamBlock), 0,
                           ctx->stream, A.n, g.ptr, g.idx, g.val, x, y, g.rb, g.nrb);
    });
    FFH_HIP(hipGetLastError());
}

// Autotuner: the fastest SpMV variant for this matrix and precision, timed
// with events of its own (10 launches each after a warm-up) on the context
// stream, into scratch vectors of its own, so that no solver vector or timing
// event is touched.  FFH_SPMV_VARIANT=<index> forces a variant.
template <typename MT, typename VT>
static int tune_spmv(Context *ctx, DevMatrix &A)
{
    constexpr bool f32 = sizeof(MT) == 4;
    if (const char *env = std::getenv("FFH_SPMV_VARIANT")) {
        int v = std::atoi(env);
        if (v >= 0 && v < kNumVariants) return ready_variant(ctx, A, v, f32);
    }
    if (A.nnz == 0 || A.n == 0) return 0;
    DBuf<VT> tx, ty;
    tx.alloc(A.n);
    ty.alloc(A.n);
    FFH_HIP(hipMemsetAsync(tx.p, 0, sizeof(VT) * A.n, ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FFH_HIP(hipEventCreate(&e0));
    int best = 0;
    try {
        FFH_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        for (int vi = 0; vi < kNumVariants; ++vi) {
            if (kVariants[vi].engine == 0 && kVariants[vi].gmul > 1 &&
                A.nrb <= grid_cap(ctx) * (kVariants[vi].gmul / 2))
                continue; // same grid as the variant before it
            if (kVariants[vi].engine == 1 && ready_variant(ctx, A, vi, f32) != vi)
                continue; // no SELL copy of this matrix
            launch_spmv<MT, VT>(ctx, A, tx.p, ty.p, vi);
            FFH_HIP(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 10; ++r) launch_spmv<MT, VT>(ctx, A, tx.p, ty.p, vi);
            FFH_HIP(hipEventRecord(e1, ctx->stream));
            FFH_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            FFH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = vi;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// The variant of a matrix for a precision, tuned on first use.  Call outside
// graph capture and before a solve's timed region.
static int spmv_variant(Context *ctx, Matrix &M, bool f32)
{
    DevMatrix &A = M.m;
    int &v = A.variant[f32 ? 1 : 0];
    if (v < 0) {
        if (f32) {
            ensure_val32(ctx, A);
            v = tune_spmv<float, float>(ctx, A);
        } else {
            v = tune_spmv<double, double>(ctx, A);
        }
    }
    v = ready_variant(ctx, A, v, f32);
    return v;
}

int select_spmv_variant(Context *ctx, Matrix &M, bool f32, int request)
{
    static_assert(kNumVariants == FFH_SPMV_VARIANTS, "variant table and header disagree");
    ensure_val64(ctx, M.m);
    if (f32) ensure_val32(ctx, M.m);
    if (request >= 0) {
        request = ready_variant(ctx, M.m, request, f32);
        M.m.variant[f32 ? 1 : 0] = request;
        invalidate_graphs(M.ws);
        return request;
    }
    return spmv_variant(ctx, M, f32);
}

void spmv64(Context *ctx, DevMatrix &A, const double *x, double *y)
{
    ensure_val64(ctx, A);
    const int vi = ready_variant(ctx, A, A.variant[0] >= 0 ? A.variant[0] : 0, false);
    launch_spmv<double, double>(ctx, A, x, y, vi);
}

void spmv32(Context *ctx, DevMatrix &A, const float *x, float *y)
{
    ensure_val32(ctx, A);
    const int vi = ready_variant(ctx, A, A.variant[1] >= 0 ? A.variant[1] : 0, true);
    launch_spmv<float, float>(ctx, A, x, y, vi);
}

// Launch `iters` PCG iterations (3 kernels each) of precision VT.
template <typename MT, typename VT>
static void launch_iterations(Context *ctx, Matrix &M, int iters, VT *x, VT *r,
                              VT *p, VT *ap, const VT *dinv, int vi, int prof_slot)
{
    const DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int n = A.n;
    const int gridv = grid_for(ctx, n, kBlock);
    const int grids = variant_grid(ctx, A, vi);
    const int stride = partial_stride(ctx);
    const GArgs<MT> g = gargs<MT>(A, vi);
    for (int it = 0; it < iters; ++it) {
        const int parity = it & 1; // chunks are even, so this is the global parity
        if (prof_slot >= 0) ctx->prof_mark(prof_slot, true);
        FFH_ENGINE(kVariants[vi].engine, {
            hipLaunchKernelGGL((k_pcg_spmv<VV, MT, VT>), dim3(grids), dim3(kStreamBlock), 0,
                               ctx->stream, n, g.ptr, g.idx, g.val, p, ap, g.rb, g.nrb,
                               w.partials.p, w.ctl.p);
        });
        if (prof_slot >= 0) ctx->prof_mark(prof_slot, false);
        hipLaunchKernelGGL((k_pcg_update<VT>), dim3(gridv), dim3(kBlock), 0, ctx->stream,
                           n, x, r, p, ap, dinv, w.partials.p, stride, grids, parity,
                           w.ctl.p);
        hipLaunchKernelGGL((k_pcg_dir<VT>), dim3(gridv), dim3(kBlock), 0, ctx->stream, n,
                           p, r, dinv, w.partials.p, stride, gridv, parity, w.ctl.p);
    }
    FFH_HIP(hipGetLastError());
}

// One PCG solve on device vectors b, x (x = initial guess, overwritten).
template <typename MT, typename VT>
static void pcg_device(Context *ctx, Matrix &M, const VT *b, VT *x,
                       VT *r, VT *p, VT *ap, VT *dinv, int method, int precond,
                       double tolerance, int tol_type, int max_iter,
                       int breakdown_mode, int check_interval, int use_graph,
                       int slot, KrylovCtl &out)
{
    DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int n = A.n;
    // Jacobi diagonal ('none' and method cg use the identity)
    hipLaunchKernelGGL((k_jacobi<VT>), dim3(grid_for(ctx, n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, n, A.gptr(), A.gidx(), A.gval(),
                       (method == 1 && precond == 1) ? 1 : 0, dinv);
    KrylovCtl init{};
    init.max_iter = max_iter;
    init.breakdown_mode = breakdown_mode;
    init.tol_type = tol_type;
    init.tolerance = tolerance;
    *w.host_ctl = init;
    FFH_HIP(hipMemcpyAsync(w.ctl.p, w.host_ctl, sizeof(KrylovCtl),
                           hipMemcpyHostToDevice, ctx->stream));
    const int vi = spmv_variant(ctx, M, sizeof(VT) == 4);
    const int gi = variant_grid(ctx, A, vi);
    const GArgs<MT> g = gargs<MT>(A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_pcg_init<VV, MT, VT>), dim3(gi), dim3(kStreamBlock), 0,
                           ctx->stream, n, g.ptr, g.idx, g.val, x, b, r, p, dinv, g.rb, g.nrb,
                           w.partials.p, partial_stride(ctx));
    });
    hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kBlock), 0, ctx->stream, w.partials.p,
                       partial_stride(ctx), gi, w.ctl.p);
    FFH_HIP(hipGetLastError());
    // iteration chunks are even so that the parity slots line up
    int chunk = check_interval > 0 ? std::max(2, (check_interval + 1) & ~1) : 16;
    const bool profiling = ctx->profiling;
    if (profiling) use_graph = 0;
    const int gslot = sizeof(VT) == 8 ? 0 : 1;
    if (use_graph && (w.graph[gslot] == nullptr || w.graph_chunk[gslot] != chunk * 64 + vi)) {
        if (w.graph[gslot]) {
            (void)hipGraphExecDestroy(w.graph[gslot]);
            w.graph[gslot] = nullptr;
        }
        hipGraph_t g;
        FFH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        launch_iterations<MT, VT>(ctx, M, chunk, x, r, p, ap, dinv, vi, -1);
        FFH_HIP(hipStreamEndCapture(ctx->stream, &g));
        FFH_HIP(hipGraphInstantiate(&w.graph[gslot], g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        w.graph_chunk[gslot] = chunk * 64 + vi;
    }
    // graphs bake in the vector pointers; they are the workspace's own
    for (;;) {
        if (use_graph)
            FFH_HIP(hipGraphLaunch(w.graph[gslot], ctx->stream));
        else
            launch_iterations<MT, VT>(ctx, M, chunk, x, r, p, ap, dinv, vi,
                                      profiling ? slot : -1);
        FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl),
                               hipMemcpyDeviceToHost, ctx->stream));
        FFH_HIP(hipStreamSynchronize(ctx->stream));
        if (w.host_ctl->done) break;
        if (!use_graph && chunk < 64) chunk *= 2;
    }
    out = *w.host_ctl;
}

static void launch_residual(Context *ctx, Matrix &M)
{
    DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int vi = spmv_variant(ctx, M, false);
    const int gr = variant_grid(ctx, A, vi);
    const GArgs<double> g = gargs<double>(A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_residual<VV>), dim3(gr), dim3(kStreamBlock), 0, ctx->stream, A.n,
                           g.ptr, g.idx, g.val, w.x.p, w.b.p, w.r.p, g.rb, g.nrb,
                           w.partials.p);
    });
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kBlock), 0, ctx->stream, w.partials.p,
                       gr, w.ctl.p);
    FFH_HIP(hipGetLastError());
}

// Public FP64 solve on device vectors.
void solve_fp64(Context *ctx, Matrix &M, const double *b, double *x,
                const ffh_solver_options &o, ffh_solver_stats &st)
{
    ensure_workspace(ctx, M);
    ensure_val64(ctx, M.m);
    spmv_variant(ctx, M, false);
    Workspace &w = M.ws;
    KrylovCtl out{};
    if (ctx->profiling) ctx->prof_begin();
    FFH_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    // graphs capture the workspace vectors; copy b/x in and out
    FFH_HIP(hipMemcpyAsync(w.b.p, b, sizeof(double) * M.m.n,
                           hipMemcpyDeviceToDevice, ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.x.p, x, sizeof(double) * M.m.n,
                           hipMemcpyDeviceToDevice, ctx->stream));
    pcg_device<double, double>(ctx, M, w.b.p, w.x.p, w.r.p, w.p.p,
                               w.ap.p, w.dinv.p, o.method, o.preconditioner,
                               o.tolerance, o.tolerance_type, o.max_iterations,
                               o.breakdown_mode, o.check_interval, o.use_graph, 0,
                               out);
    FFH_HIP(hipMemcpyAsync(x, w.x.p, sizeof(double) * M.m.n,
                           hipMemcpyDeviceToDevice, ctx->stream));
    launch_residual(ctx, M); // true residual in FP64
    FFH_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl),
                           hipMemcpyDeviceToHost, ctx->stream));
    FFH_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    FFH_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (ctx->profiling) ctx->prof_end();
    st.converged = out.converged;
    // finalize_solver_status: converged also if the last residual meets tol
    if (!st.converged && out.residual_norm <= out.tol) st.converged = 1;
    st.iterations = out.iter;
    st.final_residual = out.residual_norm;
    st.solve_time = ms * 1e-3;
    st.refinements = 0;
    st.true_residual = std::sqrt(w.host_ctl->rr);
    st.initial_residual = out.initial_norm;
}

double bench_spmv(Context *ctx, Matrix &M, int precision, int reps, double *bytes)
{
    ensure_workspace(ctx, M);
    ensure_val64(ctx, M.m);
    DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int n = A.n;
    const bool f32 = precision == FFH_FP32;
    if (f32) {
        ensure_workspace32(ctx, M);
        ensure_val32(ctx, A);
        FFH_HIP(hipMemsetAsync(w.p32.p, 0, sizeof(float) * n, ctx->stream));
    } else {
        FFH_HIP(hipMemsetAsync(w.p.p, 0, sizeof(double) * n, ctx->stream));
    }
    const int sv = f32 ? 4 : 8;
    // algorithmic bytes: values + column indices + row pointers + x + y
    *bytes = (double)A.nnz * (sv + 4) + 4.0 * (n + 1) + (double)n * 2 * sv;
    spmv_variant(ctx, M, f32);
    auto once = [&]() {
        if (f32) spmv32(ctx, A, w.p32.p, w.ap32.p);
        else spmv64(ctx, A, w.p.p, w.ap.p);
    };
    once();
    FFH_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int i = 0; i < reps; ++i) once();
    FFH_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    FFH_HIP(hipEventSynchronize(ctx->ev_stop));
    float ms = 0.0f;
    FFH_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    return (double)ms / std::max(1, reps);
}

void invalidate_graphs(Workspace &w)
{
    for (auto &g : w.graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

} // namespace ffh
