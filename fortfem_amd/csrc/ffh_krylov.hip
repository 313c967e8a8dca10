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
// Every reduction is block partials + a last-arriving-block finish in a fixed
// order, so results are deterministic run to run.  All scalars live in a
// device control block; the host polls it once per chunk of iterations and a
// chunk may be replayed from a HIP graph.  Once the control block says done,
// later launches of the chunk return immediately, so the stopping iteration
// is exactly the reference's.
#include <algorithm>
#include <cmath>

#include "ffh_internal.hpp"
#include "ffh_krylov.hpp"

namespace ffh {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double sh[kBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += sh[i];
    return t;
}

// Arrive at a reduction; true in every thread of the last block to arrive.
__device__ __forceinline__ bool last_block(unsigned int *counter)
{
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int prev = __hip_atomic_fetch_add(
            counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) __threadfence();
    return is_last;
}

// Sum of partials[0 .. gridDim.x) in a fixed order; valid in thread 0.
__device__ __forceinline__ double sum_partials(const double *partials,
                                               int stride, int offset)
{
    double t = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
        t += partials[(size_t)i * stride + offset];
    return block_sum(t);
}

__device__ __forceinline__ bool is_breakdown(double pap, double rz, int mode)
{
    if (mode == 1) return !(fabs(pap) > 1.0e-14 * fabs(rz));
    return fabs(pap) < 1.0e-14;
}

// Row gather of one row by a group of G lanes: sum_k a_k x[col_k].
template <int G, typename MT, typename VT>
__device__ __forceinline__ VT row_dot(int row, int lane, const int *__restrict__ ptr,
                                      const int *__restrict__ idx,
                                      const MT *__restrict__ val,
                                      const VT *__restrict__ x)
{
    const int b = ptr[row], e = ptr[row + 1];
    VT s = VT(0);
    for (int k = b + lane; k < e; k += G)
        s += (VT)val[k] * x[__builtin_nontemporal_load(idx + k)];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    return s;
}

// --------------------------------------------------------------- kernels --
template <int G, typename MT, typename VT>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int *__restrict__ ptr,
                                                 const int *__restrict__ idx,
                                                 const MT *__restrict__ val,
                                                 const VT *__restrict__ x,
                                                 VT *__restrict__ y)
{
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * (kBlock / G);
    for (int row = blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n;
         row += ngrp) {
        VT s = row_dot<G>(row, lane, ptr, idx, val, x);
        if (lane == 0) y[row] = s;
    }
}

// r = b - A x ; z = D r ; p = z ; partial r.r, r.z -> initial state
template <int G, typename MT, typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_init(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const MT *__restrict__ val, const VT *__restrict__ x,
    const VT *__restrict__ b, VT *__restrict__ r, VT *__restrict__ p,
    const VT *__restrict__ dinv, double *__restrict__ partials, KrylovCtl *ctl)
{
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * (kBlock / G);
    double rr = 0.0, rz = 0.0;
    for (int row = blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n;
         row += ngrp) {
        VT ax = row_dot<G>(row, lane, ptr, idx, val, x);
        if (lane == 0) {
            VT ri = b[row] - ax;
            VT zi = dinv[row] * ri;
            r[row] = ri;
            p[row] = zi;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)zi;
        }
    }
    double s0 = block_sum(rr);
    double s1 = block_sum(rz);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s0;
        partials[2 * blockIdx.x + 1] = s1;
    }
    if (last_block(&ctl->counter[0])) {
        double trr = sum_partials(partials, 2, 0);
        double trz = sum_partials(partials, 2, 1);
        if (threadIdx.x == 0) {
            double init = sqrt(trr);
            ctl->rr = trr;
            ctl->initial_norm = init;
            ctl->residual_norm = init;
            ctl->tol = ctl->tol_type == 1 ? ctl->tolerance : ctl->tolerance * init;
            ctl->rz_old = trz;
            ctl->iter = 0;
            ctl->done = ctl->max_iter <= 0 ? 1 : 0;
            __hip_atomic_store(&ctl->counter[0], 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Ap = A p ; p.Ap -> alpha (or breakdown)
template <int G, typename MT, typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const MT *__restrict__ val, const VT *__restrict__ p, VT *__restrict__ ap,
    double *__restrict__ partials, KrylovCtl *ctl)
{
    if (ctl->done) return;
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * (kBlock / G);
    double dot = 0.0;
    for (int row = blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n;
         row += ngrp) {
        VT y = row_dot<G>(row, lane, ptr, idx, val, p);
        if (lane == 0) {
            ap[row] = y;
            dot += (double)p[row] * (double)y;
        }
    }
    double s = block_sum(dot);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = s;
    if (last_block(&ctl->counter[1])) {
        double pap = sum_partials(partials, 2, 0);
        if (threadIdx.x == 0) {
            ctl->pap = pap;
            if (is_breakdown(pap, ctl->rz_old, ctl->breakdown_mode)) {
                ctl->done = 1;
                ctl->breakdown_hit = 1;
            } else {
                ctl->alpha = ctl->rz_old / pap;
            }
            __hip_atomic_store(&ctl->counter[1], 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// x += alpha p ; r -= alpha Ap ; r.r, r.(D r) -> stop test, beta
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_update(
    int n, VT *__restrict__ x, VT *__restrict__ r, const VT *__restrict__ p,
    const VT *__restrict__ ap, const VT *__restrict__ dinv,
    double *__restrict__ partials, KrylovCtl *ctl)
{
    if (ctl->done) return;
    const VT alpha = (VT)ctl->alpha;
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] = x[i] + alpha * p[i];
        VT ri = r[i] - alpha * ap[i];
        r[i] = ri;
        VT zi = dinv[i] * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    double s0 = block_sum(rr);
    double s1 = block_sum(rz);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s0;
        partials[2 * blockIdx.x + 1] = s1;
    }
    if (last_block(&ctl->counter[2])) {
        double trr = sum_partials(partials, 2, 0);
        double trz = sum_partials(partials, 2, 1);
        if (threadIdx.x == 0) {
            ctl->iter += 1;
            ctl->rr = trr;
            double rn = sqrt(trr);
            ctl->residual_norm = rn;
            if (rn <= ctl->tol) {
                ctl->converged = 1;
                ctl->done = 1;
            } else {
                ctl->beta = trz / ctl->rz_old;
                ctl->rz_old = trz;
                ctl->rz = trz;
                if (ctl->iter >= ctl->max_iter) ctl->done = 1;
            }
            __hip_atomic_store(&ctl->counter[2], 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// p = D r + beta p
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, VT *__restrict__ p,
                                                    const VT *__restrict__ r,
                                                    const VT *__restrict__ dinv,
                                                    const KrylovCtl *ctl)
{
    if (ctl->done) return;
    const VT beta = (VT)ctl->beta;
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

// r = b - A x (FP64) with ||r||^2 into ctl->rr -- true residual of the
// refinement loop.
template <int G>
__global__ __launch_bounds__(kBlock) void k_residual(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    double *__restrict__ partials, KrylovCtl *ctl)
{
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * (kBlock / G);
    double rr = 0.0;
    for (int row = blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n;
         row += ngrp) {
        double ax = row_dot<G>(row, lane, ptr, idx, val, x);
        if (lane == 0) {
            double ri = b[row] - ax;
            r[row] = ri;
            rr += ri * ri;
        }
    }
    double s = block_sum(rr);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = s;
    if (last_block(&ctl->counter[3])) {
        double t = sum_partials(partials, 2, 0);
        if (threadIdx.x == 0) {
            ctl->rr = t;
            __hip_atomic_store(&ctl->counter[3], 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
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
static int pick_group(double avg_row)
{
    if (avg_row <= 6) return 4;
    if (avg_row <= 20) return 8;
    if (avg_row <= 48) return 16;
    if (avg_row <= 128) return 32;
    return 64;
}

static int grid_for(const Context *ctx, int64_t items, int per_block)
{
    int64_t need = (items + per_block - 1) / per_block;
    int64_t cap = (int64_t)ctx->cu_count * 8;
    return (int)std::max<int64_t>(1, std::min(need, cap));
}

#define FFH_GROUP_DISPATCH(G, ...)                                            \
    switch (G) {                                                             \
    case 4: { constexpr int GG = 4; __VA_ARGS__; } break;                    \
    case 8: { constexpr int GG = 8; __VA_ARGS__; } break;                    \
    case 16: { constexpr int GG = 16; __VA_ARGS__; } break;                  \
    case 32: { constexpr int GG = 32; __VA_ARGS__; } break;                  \
    default: { constexpr int GG = 64; __VA_ARGS__; } break;                  \
    }

Workspace::~Workspace()
{
    for (auto &g : graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (host_ctl) (void)hipHostFree(host_ctl);
}

void ensure_workspace(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int n = M.m.n;
    if (w.n != n) {
        for (auto &g : w.graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }
    w.n = n;
    w.r.alloc(n), w.p.alloc(n), w.ap.alloc(n), w.x.alloc(n), w.b.alloc(n),
        w.dinv.alloc(n);
    w.partials.alloc(2 * (size_t)ctx->cu_count * 8 + 64);
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

int spmv_group(const DevMatrix &A)
{
    return pick_group(A.n ? (double)A.nnz / A.n : 1.0);
}

void spmv64(Context *ctx, const DevMatrix &A, const double *x, double *y)
{
    const int G = spmv_group(A);
    FFH_GROUP_DISPATCH(G, {
        int grid = grid_for(ctx, A.n, kBlock / GG);
        hipLaunchKernelGGL((k_spmv<GG, double, double>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, A.n, A.gptr(), A.gidx(), A.gval(), x, y);
    });
    FFH_HIP(hipGetLastError());
}

void spmv32(Context *ctx, DevMatrix &A, const float *x, float *y)
{
    ensure_val32(ctx, A);
    const int G = spmv_group(A);
    FFH_GROUP_DISPATCH(G, {
        int grid = grid_for(ctx, A.n, kBlock / GG);
        hipLaunchKernelGGL((k_spmv<GG, float, float>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, A.n, A.gptr(), A.gidx(), A.gval32(), x, y);
    });
    FFH_HIP(hipGetLastError());
}

// Launch `iters` PCG iterations (3 kernels each) of precision VT.
template <typename MT, typename VT>
static void launch_iterations(Context *ctx, Matrix &M, int iters, VT *x, VT *r,
                              VT *p, VT *ap, const VT *dinv, const MT *val,
                              int prof_slot)
{
    const DevMatrix &A = M.m;
    Workspace &w = M.ws;
    const int n = A.n;
    const int G = spmv_group(A);
    const int gridv = grid_for(ctx, n, kBlock);
    FFH_GROUP_DISPATCH(G, {
        const int grids = grid_for(ctx, n, kBlock / GG);
        for (int it = 0; it < iters; ++it) {
            if (prof_slot >= 0) ctx->prof_mark(prof_slot, true);
            hipLaunchKernelGGL((k_pcg_spmv<GG, MT, VT>), dim3(grids), dim3(kBlock), 0,
                               ctx->stream, n, A.gptr(), A.gidx(), val, p, ap,
                               w.partials.p, w.ctl.p);
            if (prof_slot >= 0) ctx->prof_mark(prof_slot, false);
            hipLaunchKernelGGL((k_pcg_update<VT>), dim3(gridv), dim3(kBlock), 0,
                               ctx->stream, n, x, r, p, ap, dinv, w.partials.p,
                               w.ctl.p);
            hipLaunchKernelGGL((k_pcg_dir<VT>), dim3(gridv), dim3(kBlock), 0,
                               ctx->stream, n, p, r, dinv, w.ctl.p);
        }
    });
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
    {
        int grid = grid_for(ctx, n, kBlock);
        hipLaunchKernelGGL((k_jacobi<VT>), dim3(grid), dim3(kBlock), 0, ctx->stream, n,
                           A.gptr(), A.gidx(), A.gval(),
                           (method == 1 && precond == 1) ? 1 : 0, dinv);
    }
    KrylovCtl init{};
    init.max_iter = max_iter;
    init.breakdown_mode = breakdown_mode;
    init.tol_type = tol_type;
    init.tolerance = tolerance;
    *w.host_ctl = init;
    FFH_HIP(hipMemcpyAsync(w.ctl.p, w.host_ctl, sizeof(KrylovCtl),
                           hipMemcpyHostToDevice, ctx->stream));
    const int G = spmv_group(A);
    FFH_GROUP_DISPATCH(G, {
        int grid = grid_for(ctx, n, kBlock / GG);
        hipLaunchKernelGGL((k_pcg_init<GG, MT, VT>), dim3(grid), dim3(kBlock), 0,
                           ctx->stream, n, A.gptr(), A.gidx(), val, x, b, r, p,
                           dinv, w.partials.p, w.ctl.p);
    });
    FFH_HIP(hipGetLastError());
    int chunk = check_interval > 0 ? check_interval : 16;
    const bool profiling = ctx->profiling;
    if (profiling) use_graph = 0;
    const int gslot = sizeof(VT) == 8 ? 0 : 1;
    if (use_graph && (w.graph[gslot] == nullptr || w.graph_chunk[gslot] != chunk)) {
        if (w.graph[gslot]) {
            (void)hipGraphExecDestroy(w.graph[gslot]);
            w.graph[gslot] = nullptr;
        }
        hipGraph_t g;
        FFH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        launch_iterations<MT, VT>(ctx, M, chunk, x, r, p, ap, dinv, val, -1);
        FFH_HIP(hipStreamEndCapture(ctx->stream, &g));
        FFH_HIP(hipGraphInstantiate(&w.graph[gslot], g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        w.graph_chunk[gslot] = chunk;
    }
    // graphs bake in the vector pointers; they are the workspace's own
    int done_iters = 0;
    for (;;) {
        int todo = chunk;
        if (use_graph)
            FFH_HIP(hipGraphLaunch(w.graph[gslot], ctx->stream));
        else
            launch_iterations<MT, VT>(ctx, M, todo, x, r, p, ap, dinv, val,
                                      profiling ? slot : -1);
        done_iters += todo;
        FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl),
                               hipMemcpyDeviceToHost, ctx->stream));
        FFH_HIP(hipStreamSynchronize(ctx->stream));
        if (w.host_ctl->done) break;
        // grow the chunk towards the expected remaining work
        if (!use_graph && chunk < 64) chunk *= 2;
        (void)done_iters;
    }
    out = *w.host_ctl;
}

// Public FP64 solve on device vectors.
void solve_fp64(Context *ctx, Matrix &M, const double *b, double *x,
                const ffh_solver_options &o, ffh_solver_stats &st)
{
    ensure_workspace(ctx, M);
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
    // true residual in FP64
    {
        const int G = spmv_group(M.m);
        FFH_GROUP_DISPATCH(G, {
            int grid = grid_for(ctx, M.m.n, kBlock / GG);
            hipLaunchKernelGGL((k_residual<GG>), dim3(grid), dim3(kBlock), 0,
                               ctx->stream, M.m.n, M.m.gptr(), M.m.gidx(), M.m.gval(),
                               w.x.p, w.b.p, w.r.p, w.partials.p, w.ctl.p);
        });
        FFH_HIP(hipGetLastError());
    }
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
    ensure_val32(ctx, A);
    Workspace &w = M.ws;
    const int n = A.n;
    if (ctx->profiling) ctx->prof_begin();
    FFH_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.b.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    FFH_HIP(hipMemcpyAsync(w.x.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           ctx->stream));
    const int G = spmv_group(A);
    auto true_residual = [&]() -> double {
        if (ctx->profiling) ctx->prof_mark(0, true);
        FFH_GROUP_DISPATCH(G, {
            int grid = grid_for(ctx, n, kBlock / GG);
            hipLaunchKernelGGL((k_residual<GG>), dim3(grid), dim3(kBlock), 0,
                               ctx->stream, n, A.gptr(), A.gidx(), A.gval(), w.x.p,
                               w.b.p, w.r.p, w.partials.p, w.ctl.p);
        });
        if (ctx->profiling) ctx->prof_mark(0, false);
        FFH_HIP(hipGetLastError());
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
    int gridv = grid_for(ctx, n, kBlock);
    while (!converged && outer < max_ref && total_iters < o.max_iterations) {
        const double s = rnorm > 0 ? 1.0 / rnorm : 1.0;
        hipLaunchKernelGGL(k_scale_f32, dim3(gridv), dim3(kBlock), 0, ctx->stream, n,
                           w.r.p, s, w.r32.p + 0, w.x32.p);
        // inner solve A d = r / ||r|| in FP32 (b32 held in ap32 scratch)
        FFH_HIP(hipMemcpyAsync(w.ap32.p, w.r32.p, sizeof(float) * n,
                               hipMemcpyDeviceToDevice, ctx->stream));
        KrylovCtl out{};
        // pcg_device reads b (ap32 copy) only in the init kernel, before any
        // write to ap32 by the iteration kernels
        float *b32 = w.ap32.p;
        pcg_device<float, float>(ctx, M, A.gval32(), b32, w.x32.p, w.r32.p, w.p32.p,
                                 w.ap32.p, w.dinv32.p, o.method, o.preconditioner,
                                 inner_tol, 0, o.max_iterations - total_iters, 1,
                                 o.check_interval, o.use_graph, 1, out);
        total_iters += out.iter;
        hipLaunchKernelGGL(k_refine, dim3(gridv), dim3(kBlock), 0, ctx->stream, n,
                           w.x.p, w.x32.p, rnorm);
        FFH_HIP(hipGetLastError());
        ++outer;
        double prev = rnorm;
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
