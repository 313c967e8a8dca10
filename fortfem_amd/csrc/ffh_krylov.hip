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

// Row engine of a launch: V = 0 stream, otherwise vector CSR with G = V lanes.
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

__global__ __launch_bounds__(kBlock) void k_scale_f32(int n, const double *__restrict__ r,
                                                      double s, float *__restrict__ b32,
                                                      float *__restrict__ x32)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        b32[i] = (float)(r[i] * s);
        x32[i] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_refine(int n, double *__restrict__ x,
                                                   const float *__restrict__ d32,
                                                   double s)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        x[i] = x[i] + s * (double)d32[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ a,
                                                   float *__restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
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
                                        {8, 2}, {16, 2}, {32, 2}, {32, 4}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static int variant_grid(const Context *ctx, const DevMatrix &A, int vi)
{
    const SpmvVariant v = kVariants[vi];
    const int64_t cap = (int64_t)grid_cap(ctx) * v.gmul;
    const int64_t need = v.engine == 0 ? A.nrb : ((int64_t)A.n * v.engine + kStreamBlock - 1) / kStreamBlock;
    return (int)std::max<int64_t>(1, std::min(need, cap));
}

#define FFH_ENGINE(eng, ...)                                                  \
    switch (eng) {                                                           \
    case 0: { constexpr int VV = 0; __VA_ARGS__; } break;                     \
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
    FFH_HIP(hipStreamSynchronize(ctx->stream));
}

Workspace::~Workspace()
{
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

template <typename MT, typename VT>
static void launch_spmv(Context *ctx, const DevMatrix &A, const MT *val, const VT *x, VT *y,
                        int vi)
{
    const int grid = variant_grid(ctx, A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_spmv<VV, MT, VT>), dim3(grid), dim3(kStreamBlock), 0,
                           ctx->stream, A.n, A.gptr(), A.gidx(), val, x, y, A.rowblocks.p,
                           A.nrb);
    });
    FFH_HIP(hipGetLastError());
}

// Autotuner: the fastest SpMV variant for this matrix and precision, timed
// with events of its own (10 launches each after a warm-up) on the context
// stream, into scratch vectors of its own, so that no solver vector or timing
// event is touched.  FFH_SPMV_VARIANT=<index> forces a variant.
template <typename MT, typename VT>
static int tune_spmv(Context *ctx, DevMatrix &A, const MT *val)
{
    if (const char *env = std::getenv("FFH_SPMV_VARIANT")) {
        int v = std::atoi(env);
        if (v >= 0 && v < kNumVariants) return v;
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
            launch_spmv<MT, VT>(ctx, A, val, tx.p, ty.p, vi);
            FFH_HIP(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 10; ++r) launch_spmv<MT, VT>(ctx, A, val, tx.p, ty.p, vi);
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
            v = tune_spmv<float, float>(ctx, A, A.gval32());
        } else {
            v = tune_spmv<double, double>(ctx, A, A.gval());
        }
    }
    return v;
}

int select_spmv_variant(Context *ctx, Matrix &M, bool f32, int request)
{
    static_assert(kNumVariants == FFH_SPMV_VARIANTS, "variant table and header disagree");
    ensure_val64(ctx, M.m);
    if (f32) ensure_val32(ctx, M.m);
    if (request >= 0) {
        M.m.variant[f32 ? 1 : 0] = request;
        invalidate_graphs(M.ws);
        return request;
    }
    return spmv_variant(ctx, M, f32);
}

void spmv64(Context *ctx, const DevMatrix &A, const double *x, double *y)
{
    const int vi = A.variant[0] >= 0 ? A.variant[0] : 0;
    launch_spmv<double, double>(ctx, A, A.gval(), x, y, vi);
}

void spmv32(Context *ctx, DevMatrix &A, const float *x, float *y)
{
    ensure_val32(ctx, A);
    const int vi = A.variant[1] >= 0 ? A.variant[1] : 0;
    launch_spmv<float, float>(ctx, A, A.gval32(), x, y, vi);
}

// Launch `iters` PCG iterations (3 kernels each) of precision VT.
template <typename MT, typename VT>
static void launch_iterations(Context *ctx, Matrix &M, int iters, VT *x, VT *r,
                              VT *p, VT *ap, const VT *dinv, const MT *val, int vi,
                              int prof_slot)
{
    const DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int n = A.n;
    const int gridv = grid_for(ctx, n, kBlock);
    const int grids = variant_grid(ctx, A, vi);
    const int stride = partial_stride(ctx);
    for (int it = 0; it < iters; ++it) {
        const int parity = it & 1; // chunks are even, so this is the global parity
        if (prof_slot >= 0) ctx->prof_mark(prof_slot, true);
        FFH_ENGINE(kVariants[vi].engine, {
            hipLaunchKernelGGL((k_pcg_spmv<VV, MT, VT>), dim3(grids), dim3(kStreamBlock), 0,
                               ctx->stream, n, A.gptr(), A.gidx(), val, p, ap, A.rowblocks.p,
                               A.nrb, w.partials.p, w.ctl.p);
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
static void pcg_device(Context *ctx, Matrix &M, const MT *val, const VT *b, VT *x,
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
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_pcg_init<VV, MT, VT>), dim3(gi), dim3(kStreamBlock), 0,
                           ctx->stream, n, A.gptr(), A.gidx(), val, x, b, r, p, dinv,
                           A.rowblocks.p, A.nrb, w.partials.p, partial_stride(ctx));
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
        launch_iterations<MT, VT>(ctx, M, chunk, x, r, p, ap, dinv, val, vi, -1);
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
            launch_iterations<MT, VT>(ctx, M, chunk, x, r, p, ap, dinv, val, vi,
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
    const int g = variant_grid(ctx, A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_residual<VV>), dim3(g), dim3(kStreamBlock), 0, ctx->stream, A.n,
                           A.gptr(), A.gidx(), A.gval(), w.x.p, w.b.p, w.r.p, A.rowblocks.p,
                           A.nrb, w.partials.p);
    });
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kBlock), 0, ctx->stream, w.partials.p,
                       g, w.ctl.p);
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
    pcg_device<double, double>(ctx, M, M.m.gval(), w.b.p, w.x.p, w.r.p, w.p.p,
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

// Mixed precision: FP32 matrix, vectors and SpMV accumulation in the inner
// Jacobi-PCG (reductions in FP64), inside an FP64 iterative-refinement loop
// whose residual b - A x is recomputed in FP64 with the FP64 matrix.  Stops
// when ||b - A x|| <= tol * ||b - A x0|| (relative) or <= tol (absolute).
void solve_mixed(Context *ctx, Matrix &M, const double *b, double *x,
                 const ffh_solver_options &o, ffh_solver_stats &st)
{
    ensure_workspace(ctx, M);
    ensure_workspace32(ctx, M);
    DevMatrix &A = M.m;
    ensure_val64(ctx, A);
    ensure_val32(ctx, A);
    spmv_variant(ctx, M, false);
    spmv_variant(ctx, M, true);
    Workspace &w = M.ws;
    const int n = A.n;
    if (ctx->profiling) ctx->prof_begin();
    FFH_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.b.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.x.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    auto true_residual = [&]() -> double {
        if (ctx->profiling) ctx->prof_mark(0, true);
        launch_residual(ctx, M);
        if (ctx->profiling) ctx->prof_mark(0, false);
        FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl),
                               hipMemcpyDeviceToHost, ctx->stream));
        FFH_HIP(hipStreamSynchronize(ctx->stream));
        return std::sqrt(w.host_ctl->rr);
    };
    // the control block counters must start at zero
    KrylovCtl zero{};
    *w.host_ctl = zero;
    FFH_HIP(hipMemcpyAsync(w.ctl.p, w.host_ctl, sizeof(KrylovCtl),
                           hipMemcpyHostToDevice, ctx->stream));
    double rnorm = true_residual();
    const double r0 = rnorm;
    const double tol = o.tolerance_type == 1 ? o.tolerance : o.tolerance * r0;
    int total_iters = 0, outer = 0, converged = rnorm <= tol;
    const int max_ref = o.max_refinements > 0 ? o.max_refinements : 50;
    const double inner_tol = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-4;
    const int gridv = grid_for(ctx, n, kBlock);
    while (!converged && outer < max_ref && total_iters < o.max_iterations) {
        const double s = rnorm > 0 ? 1.0 / rnorm : 1.0;
        // inner right-hand side r / ||r|| in FP32, held in ap32 until the
        // init kernel has consumed it; the correction accumulates in x32
        hipLaunchKernelGGL(k_scale_f32, dim3(gridv), dim3(kBlock), 0, ctx->stream, n,
                           w.r.p, s, w.ap32.p, w.x32.p);
        KrylovCtl out{};
        pcg_device<float, float>(ctx, M, A.gval32(), w.ap32.p, w.x32.p, w.r32.p, w.p32.p,
                                 w.ap32.p, w.dinv32.p, o.method, o.preconditioner,
                                 inner_tol, 0, o.max_iterations - total_iters, 1,
                                 o.check_interval, o.use_graph, 1, out);
        total_iters += out.iter;
        hipLaunchKernelGGL(k_refine, dim3(gridv), dim3(kBlock), 0, ctx->stream, n,
                           w.x.p, w.x32.p, rnorm);
        FFH_HIP(hipGetLastError());
        ++outer;
        const double prev = rnorm;
        rnorm = true_residual();
        converged = rnorm <= tol;
        if (out.iter == 0 || !(rnorm < prev)) break; // stagnation
    }
    FFH_HIP(hipMemcpyAsync(x, w.x.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    FFH_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    FFH_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    FFH_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (ctx->profiling) ctx->prof_end();
    st.converged = converged;
    st.iterations = total_iters;
    st.final_residual = rnorm;
    st.solve_time = ms * 1e-3;
    st.refinements = outer;
    st.true_residual = rnorm;
    st.initial_residual = r0;
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
