// ffh_batch.hip -- served solves: many resident systems with several
// right-hand sides each, advanced together by one launch per PCG step.
//
// Reference loop: src/solvers/advanced_solvers.f90 pcg_iteration_loop
// (:880-925), compute_solver_tolerance (:1063-1073), Jacobi diagonal
// (:1124-1155).  Each (system, right-hand side) runs its own Jacobi-PCG with
// x0 = 0 and the relative stop ||r|| <= tol ||b||.
//
// Design: the active systems form one block-diagonal operator.  A work table
// lists row chunks (kRowsPerItem rows of one system); every PCG launch covers
// the whole table, and the right-hand sides of a system are interleaved
// (row-major n x R, R in {1, 2, 4, 8}) so one pass over the matrix serves all
// of them.  Dot products are per-chunk partials, summed per system in chunk
// order by one block per system (deterministic: a system's result does not
// depend on what else is in the batch).  The inner iteration runs in FP32
// (matrix, r, p, Ap) with FP64 x and FP64 reductions; at the end of every
// chunk of iterations the residual is recomputed in FP64 from the FP64
// matrix (reliable update: r is replaced, p kept), which decides convergence.
// Systems join and leave only at chunk boundaries; admission is by the
// device-memory footprint of the request.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>

#include "ffh_internal.hpp"
#include "ffh_krylov.hpp"

namespace ffh {

constexpr int kRowsPerItem = 256;
constexpr int kMaxRhs = 8;

struct BatchSys {
    const int *ptr, *idx;
    const float *val32;
    const double *val64;
    const double *dinv;
    const double *b; // n x R
    double *x;       // n x R
    float *r, *p, *ap;       // FP32 iteration
    double *r64, *p64, *ap64; // FP64 iteration (prec = 1)
    int n, R, w0, w1, slot, prec;
};

// per (system slot, right-hand side)
enum BatchState : int { kActive = 0, kConverged = 1, kStopped = 2, kUnused = 3, kCandidate = 4,
                        kFresh = 5 };
struct BatchScal {
    double rz, alpha, beta, rr, bb, tol2, true_rr;
    int iters, state, max_iter, restart;
};

// ---------------------------------------------------------------- kernels --
// Block-wide total of per-thread values (256 threads), valid in thread 0.
__device__ __forceinline__ double bsum256(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

// Sum of the per-thread value of each right-hand side r = t % R over the
// block, in thread order; result for r in thread r.
__device__ __forceinline__ double rhs_sum(double v, int R, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < R)
        for (int j = threadIdx.x; j < 256; j += R) t += sh[j];
    __syncthreads();
    return t;
}

// Ap = A p over one chunk (8 lanes per row, R interleaved right-hand sides),
// partials of p.Ap
template <int R, typename MT, typename VT, typename AT>
__device__ __forceinline__ void chunk_spmv(const int *__restrict__ ptr, const int *__restrict__ idx,
                                           const MT *__restrict__ val, const VT *__restrict__ xin,
                                           int row0, int row1, AT *acc_out, double &dot,
                                           const VT *__restrict__ dotv)
{
    const int lane = threadIdx.x & 7, grp = threadIdx.x >> 3;
    for (int row = row0 + grp; row < row1; row += 32) {
        AT acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = AT(0);
        const int b = ptr[row], e = ptr[row + 1];
        for (int k = b + lane; k < e; k += 8) {
            const int c = idx[k];
            const AT v = (AT)val[k];
            const VT *xc = xin + (size_t)c * R;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += v * (AT)xc[r];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            for (int off = 4; off > 0; off >>= 1) acc[r] += __shfl_xor(acc[r], off, 8);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane == r) {
                acc_out[(size_t)row * R + r] = acc[r];
                if (dotv) dot += (double)dotv[(size_t)row * R + r] * (double)acc[r];
            }
    }
}

template <typename VT>
__device__ __forceinline__ double spmv_body(const BatchSys &s, const VT *val, const VT *p, VT *ap,
                                            int row0, int row1)
{
    double dot = 0.0;
    switch (s.R) {
    case 1: chunk_spmv<1, VT, VT, VT>(s.ptr, s.idx, val, p, row0, row1, ap, dot, p); break;
    case 2: chunk_spmv<2, VT, VT, VT>(s.ptr, s.idx, val, p, row0, row1, ap, dot, p); break;
    case 4: chunk_spmv<4, VT, VT, VT>(s.ptr, s.idx, val, p, row0, row1, ap, dot, p); break;
    default: chunk_spmv<8, VT, VT, VT>(s.ptr, s.idx, val, p, row0, row1, ap, dot, p); break;
    }
    return dot;
}

__global__ __launch_bounds__(256) void k_b_spmv(const int2 *__restrict__ work,
                                                const BatchSys *__restrict__ sys,
                                                const BatchScal *__restrict__ sc,
                                                double *__restrict__ part)
{
    __shared__ double sh[256];
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    const int row0 = wk.y, row1 = min(wk.y + kRowsPerItem, s.n);
    const double dot = s.prec ? spmv_body<double>(s, s.val64, s.p64, s.ap64, row0, row1)
                              : spmv_body<float>(s, s.val32, s.p, s.ap, row0, row1);
    // lane r of each 8-lane group holds right-hand side r
    sh[threadIdx.x] = dot;
    __syncthreads();
    if (threadIdx.x < s.R) {
        double t = 0.0;
        for (int g = 0; g < 32; ++g) t += sh[g * 8 + threadIdx.x];
        part[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = t;
    }
}

// per system: p.Ap -> alpha (one block per system)
__global__ __launch_bounds__(256) void k_b_alpha(const BatchSys *__restrict__ sys,
                                                 BatchScal *__restrict__ sc,
                                                 const double *__restrict__ part)
{
    __shared__ double sh[4];
    const BatchSys s = sys[blockIdx.x];
    for (int r = 0; r < s.R; ++r) {
        double t = 0.0;
        for (int w = s.w0 + threadIdx.x; w < s.w1; w += 256) t += part[(size_t)w * kMaxRhs + r];
        t = bsum256(t, sh);
        if (threadIdx.x == 0) {
            BatchScal &q = sc[s.slot * kMaxRhs + r];
            if (q.state == kActive) {
                // relative breakdown test; the FP64 refresh then decides
                if (!(fabs(t) > 1.0e-14 * fabs(q.rz))) {
                    q.state = kCandidate;
                    q.restart = 2; // breakdown: restart the direction after the refresh
                    q.alpha = 0.0;
                } else {
                    q.alpha = q.rz / t;
                }
            } else {
                q.alpha = 0.0;
            }
        }
    }
}

template <typename VT>
__device__ __forceinline__ void update_body(const BatchSys &s, VT *r, const VT *p, const VT *ap,
                                            double a, int row0, int row1, double &crr,
                                            double &crz)
{
    const int R = s.R;
    for (int e = threadIdx.x; e < (row1 - row0) * R; e += 256) {
        const size_t i = (size_t)row0 * R + e;
        const int row = row0 + e / R;
        s.x[i] += a * (double)p[i];
        const VT ri = r[i] - (VT)a * ap[i];
        r[i] = ri;
        const double zi = s.dinv[row] * (double)ri;
        crr += (double)ri * (double)ri;
        crz += (double)ri * zi;
    }
}

// x += alpha p, r -= alpha Ap, partials of r.r and r.(D r)
__global__ __launch_bounds__(256) void k_b_update(const int2 *__restrict__ work,
                                                  const BatchSys *__restrict__ sys,
                                                  const BatchScal *__restrict__ sc,
                                                  double *__restrict__ part_rr,
                                                  double *__restrict__ part_rz)
{
    __shared__ double sh[256];
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    const int R = s.R;
    const int row0 = wk.y, row1 = min(wk.y + kRowsPerItem, s.n);
    const int rr_ = threadIdx.x % R;
    const BatchScal q = sc[s.slot * kMaxRhs + rr_];
    double crr = 0.0, crz = 0.0;
    if (q.state == kActive) {
        if (s.prec)
            update_body<double>(s, s.r64, s.p64, s.ap64, q.alpha, row0, row1, crr, crz);
        else
            update_body<float>(s, s.r, s.p, s.ap, q.alpha, row0, row1, crr, crz);
    }
    const double trr = rhs_sum(crr, R, sh);
    const double trz = rhs_sum(crz, R, sh);
    if (threadIdx.x < R) {
        part_rr[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = trr;
        part_rz[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = trz;
    }
}

// per system: beta, stop tests
__global__ __launch_bounds__(256) void k_b_beta(const BatchSys *__restrict__ sys,
                                                BatchScal *__restrict__ sc,
                                                const double *__restrict__ part_rr,
                                                const double *__restrict__ part_rz)
{
    __shared__ double sh[4];
    const BatchSys s = sys[blockIdx.x];
    for (int r = 0; r < s.R; ++r) {
        double trr = 0.0, trz = 0.0;
        for (int w = s.w0 + threadIdx.x; w < s.w1; w += 256) {
            trr += part_rr[(size_t)w * kMaxRhs + r];
            trz += part_rz[(size_t)w * kMaxRhs + r];
        }
        trr = bsum256(trr, sh);
        trz = bsum256(trz, sh);
        if (threadIdx.x == 0) {
            BatchScal &q = sc[s.slot * kMaxRhs + r];
            if (q.state == kActive) {
                q.iters += 1;
                q.rr = trr;
                q.beta = q.rz != 0.0 ? trz / q.rz : 0.0;
                q.rz = trz;
                // FP64 iteration: the reference's stop on the recurrence
                // residual (pcg_iteration_loop :880-925), finalised at the
                // refresh (which records the true residual).  FP32 iteration:
                // convergence is decided on the FP64 residual at the refresh.
                if (s.prec && trr <= q.tol2) {
                    q.state = kCandidate;
                    q.restart = 3;
                } else if (q.iters >= q.max_iter) {
                    q.state = kStopped;
                }
            }
        }
    }
}

// p = D r + beta p
template <typename VT>
__device__ __forceinline__ void dir_body(const BatchSys &s, const VT *r, VT *p, double beta,
                                         bool restart, int row0, int row1)
{
    const int R = s.R;
    const VT bt = (VT)beta;
    for (int e = threadIdx.x; e < (row1 - row0) * R; e += 256) {
        const size_t i = (size_t)row0 * R + e;
        const int row = row0 + e / R;
        const VT z = (VT)(s.dinv[row] * (double)r[i]);
        // a restart never reads the old direction (it may hold anything)
        p[i] = restart ? z : z + bt * p[i];
    }
}

// (restart_only: p = D r where the refresh asked for a restart)
__global__ __launch_bounds__(256) void k_b_dir(const int2 *__restrict__ work,
                                               const BatchSys *__restrict__ sys,
                                               const BatchScal *__restrict__ sc, int restart_only)
{
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    const int R = s.R;
    const int row0 = wk.y, row1 = min(wk.y + kRowsPerItem, s.n);
    const int r = threadIdx.x % R;
    const BatchScal q = sc[s.slot * kMaxRhs + r];
    if (restart_only ? (q.restart == 0 || q.state != kActive) : q.state != kActive) return;
    if (s.prec)
        dir_body<double>(s, s.r64, s.p64, q.beta, restart_only != 0, row0, row1);
    else
        dir_body<float>(s, s.r, s.p, q.beta, restart_only != 0, row0, row1);
}

// FP64 residual b - A x (reliable update): r <- residual (FP32 copy, or
// FP64 for a system iterating in FP64),
// partials of its r.r and r.(D r).  Fresh systems start here with x = 0.
__global__ __launch_bounds__(256) void k_b_resid(const int2 *__restrict__ work,
                                                 const BatchSys *__restrict__ sys,
                                                 const BatchScal *__restrict__ sc,
                                                 double *__restrict__ res_scratch,
                                                 double *__restrict__ part_rr,
                                                 double *__restrict__ part_rz,
                                                 double *__restrict__ part_bb,
                                                 double *__restrict__ part_dd)
{
    __shared__ double sh[256];
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    const int R = s.R;
    const int row0 = wk.y, row1 = min(wk.y + kRowsPerItem, s.n);
    // A x in FP64 into the system's own slice of the scratch (n x R)
    double *ax = res_scratch + (size_t)s.w0 * kRowsPerItem * kMaxRhs;
    double dummy = 0.0;
    switch (R) {
    case 1: chunk_spmv<1, double, double, double>(s.ptr, s.idx, s.val64, s.x, row0, row1, ax, dummy, (const double *)nullptr); break;
    case 2: chunk_spmv<2, double, double, double>(s.ptr, s.idx, s.val64, s.x, row0, row1, ax, dummy, (const double *)nullptr); break;
    case 4: chunk_spmv<4, double, double, double>(s.ptr, s.idx, s.val64, s.x, row0, row1, ax, dummy, (const double *)nullptr); break;
    default: chunk_spmv<8, double, double, double>(s.ptr, s.idx, s.val64, s.x, row0, row1, ax, dummy, (const double *)nullptr); break;
    }
    __syncthreads();
    const int r = threadIdx.x % R;
    const BatchScal q = sc[s.slot * kMaxRhs + r];
    double crr = 0.0, crz = 0.0, cbb = 0.0, cdd = 0.0;
    const bool live = q.state == kActive || q.state == kCandidate || q.state == kFresh;
    for (int e = threadIdx.x; e < (row1 - row0) * R; e += 256) {
        const size_t i = (size_t)row0 * R + e;
        const int row = row0 + e / R;
        const double bi = s.b[i];
        if (q.state == kFresh) s.x[i] = 0.0;
        const double ri = q.state == kFresh ? bi : bi - ax[i];
        if (live) {
            // recurrence vs FP64 residual; an FP64 recurrence is kept as it
            // is (the reference's iteration), only started from b
            const double drift = (s.prec ? s.r64[i] : (double)s.r[i]) - ri;
            cdd += drift * drift;
            if (!s.prec)
                s.r[i] = (float)ri;
            else if (q.state == kFresh || q.restart == 2)
                s.r64[i] = ri;
        }
        crr += ri * ri;
        crz += ri * s.dinv[row] * ri;
        cbb += bi * bi;
    }
    const double trr = rhs_sum(crr, R, sh);
    const double trz = rhs_sum(crz, R, sh);
    const double tbb = rhs_sum(cbb, R, sh);
    const double tdd = rhs_sum(cdd, R, sh);
    if (threadIdx.x < R) {
        part_rr[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = trr;
        part_rz[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = trz;
        part_bb[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = tbb;
        part_dd[(size_t)blockIdx.x * kMaxRhs + threadIdx.x] = tdd;
    }
}

// per system: true residual -> converged, or continue (reliable update keeps
// p; a breakdown, or a fresh system, restarts p = D r)
__global__ __launch_bounds__(256) void k_b_refresh(const BatchSys *__restrict__ sys,
                                                   BatchScal *__restrict__ sc,
                                                   const double *__restrict__ part_rr,
                                                   const double *__restrict__ part_rz,
                                                   const double *__restrict__ part_bb,
                                                   const double *__restrict__ part_dd)
{
    __shared__ double sh[4];
    const BatchSys s = sys[blockIdx.x];
    for (int r = 0; r < s.R; ++r) {
        double trr = 0.0, trz = 0.0, tbb = 0.0, tdd = 0.0;
        for (int w = s.w0 + threadIdx.x; w < s.w1; w += 256) {
            trr += part_rr[(size_t)w * kMaxRhs + r];
            trz += part_rz[(size_t)w * kMaxRhs + r];
            tbb += part_bb[(size_t)w * kMaxRhs + r];
            tdd += part_dd[(size_t)w * kMaxRhs + r];
        }
        trr = bsum256(trr, sh);
        trz = bsum256(trz, sh);
        tbb = bsum256(tbb, sh);
        tdd = bsum256(tdd, sh);
        if (threadIdx.x == 0) {
            BatchScal &q = sc[s.slot * kMaxRhs + r];
            const bool broke = q.restart == 2;
            const bool done_by_recurrence = q.state == kCandidate && q.restart == 3;
            q.restart = 0;
            if (q.state == kFresh) {
                q.bb = tbb;
                q.tol2 = q.tol2 * tbb; // tol2 holds tol^2 until here; ||r0|| = ||b|| (x0 = 0)
                q.iters = 0;
                q.state = tbb == 0.0 ? kConverged : kActive;
                q.rz = trz;
                q.rr = trr;
                q.true_rr = trr;
                q.restart = q.state == kActive;
            } else if (q.state == kActive || q.state == kCandidate) {
                q.true_rr = trr;
                q.rr = trr;
                if (trr <= q.tol2 || done_by_recurrence) {
                    q.state = kConverged;
                } else if (q.iters >= q.max_iter) {
                    q.state = kStopped;
                } else if (s.prec) {
                    // FP64 iteration: the recurrence continues untouched
                    // (after a breakdown it restarts from the FP64 residual)
                    q.state = kActive;
                    if (broke) {
                        q.restart = 1;
                        q.rz = trz;
                    }
                } else {
                    // FP32 iteration, reliable update: r is now the FP64
                    // residual.  The direction is kept while the recurrence
                    // tracked it (drift under 1 % of the residual); after a
                    // breakdown or a larger drift CG restarts from r.
                    q.restart = broke || tdd > 1.0e-4 * trr;
                    q.state = kActive;
                    q.rz = trz;
                }
            }
        }
    }
}

// Jacobi diagonal in FP64 (|a_ii| > 1e-14, else 1: advanced_solvers.f90:1124-1155)
__global__ void k_b_jacobi(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                           const double *__restrict__ val, double *__restrict__ dinv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
        if (idx[k] == i) d = val[k];
    dinv[i] = fabs(d) > 1.0e-14 ? 1.0 / d : 1.0;
}

__global__ void k_b_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// b (n x nrhs, column-major host layout) -> n x R interleaved, zero padding
__global__ void k_b_interleave(int n, int nrhs, int R, const double *__restrict__ src,
                               double *__restrict__ dst)
{
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= (int64_t)n * R) return;
    const int row = (int)(i / R), r = (int)(i % R);
    dst[i] = r < nrhs ? src[(size_t)r * n + row] : 0.0;
}

// ----------------------------------------------------------------- server --
struct ServedRequest {
    int64_t ticket = 0;
    Problem *sym = nullptr;
    int n = 0, nrhs = 0, R = 1;
    int64_t nnz = 0;
    std::vector<double> coeff, rhs;
    double mass = 0.0, tol = 1e-10;
    int max_iter = 100000;
    size_t bytes = 0;
    DBuf<float> val32, r, p, ap;
    DBuf<double> val64, dinv, b, x, r64, p64, ap64;
    int family = 0;
    int prec = 0;                 // 0 FP32 iteration, 1 FP64 iteration
    std::vector<double> best_rr;  // best FP64 residual seen, per right-hand side
    std::vector<int> stall;       // refreshes without halving it
    int slot = -1;
    bool done = false;
    std::vector<double> sol;
    std::vector<int> iters, converged;
    std::vector<double> resid;
};

struct ServerKey {
    int mesh, family, order;
    bool operator<(const ServerKey &o) const
    {
        return std::tie(mesh, family, order) < std::tie(o.mesh, o.family, o.order);
    }
};

} // namespace ffh

struct ffh_server {
    ffh::Context *ctx = nullptr;
    size_t budget = 0, used = 0;
    int max_sys = 64, chunk = 16;
    int policy = 0; // 0 FP32 iteration + switch on stagnation, 1 FP64, 2 FP64 for Nedelec
    int64_t switches = 0;
    std::map<ffh::ServerKey, std::unique_ptr<ffh::Problem>> cache;
    std::map<int64_t, std::unique_ptr<ffh::ServedRequest>> reqs;
    std::deque<ffh::ServedRequest *> queue;
    std::vector<ffh::ServedRequest *> active;
    std::vector<int> free_slots;
    bool changed = false;
    int64_t next_ticket = 1;
    // device batch state
    ffh::DBuf<ffh::BatchSys> d_sys;
    ffh::DBuf<int2> d_work;
    ffh::DBuf<ffh::BatchScal> d_scal;
    ffh::DBuf<double> d_p0, d_p1, d_p2, d_p3, d_scratch;
    std::vector<ffh::BatchScal> h_scal;
    int nwork = 0;
    int64_t iterations_run = 0;
};

namespace ffh {

// worst-case device bytes of a request: both value copies, b and x, and the
// FP32 and FP64 iteration vectors (an FP32 request may switch to FP64)
static size_t request_bytes(int n, int64_t nnz, int R)
{
    return (size_t)nnz * 12 + (size_t)n * 8 + (size_t)n * R * (8 + 8 + 12 + 24) + 4096;
}

static void alloc_fp64_vectors(ServedRequest *q, hipStream_t s)
{
    const size_t nr = (size_t)q->n * q->R;
    q->r64.alloc(nr);
    q->p64.alloc(nr);
    q->ap64.alloc(nr);
    FFH_HIP(hipMemsetAsync(q->r64.p, 0, sizeof(double) * nr, s));
    FFH_HIP(hipMemsetAsync(q->p64.p, 0, sizeof(double) * nr, s));
}

static int rhs_stride(int nrhs) { return nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : 8; }

Problem *server_prepare(ffh_server *S, int mesh_key, int family, int order, int dim, int nvert,
                        const double *verts, int ncell, const int *cells)
{
    ServerKey key{mesh_key, family, order};
    auto it = S->cache.find(key);
    if (it != S->cache.end()) return it->second.get();
    FFH_REQUIRE(dim == 3 && (family == FFH_FAMILY_H1 || family == FFH_FAMILY_NEDELEC),
                FFH_ERR_INVALID, "served families: tetrahedral H1 and Nedelec");
    FFH_REQUIRE(nvert > 0 && ncell > 0 && verts && cells, FFH_ERR_INVALID, "empty mesh");
    auto P = std::make_unique<Problem>();
    P->ctx = S->ctx;
    P->dim = 3;
    P->nvert = nvert;
    P->ncell = ncell;
    P->vertices.assign(verts, verts + (size_t)3 * nvert);
    P->cells.assign(cells, cells + (size_t)4 * ncell);
    P->d_vertices.upload(P->vertices, S->ctx->stream);
    problem_pattern(P.get(), family, order);
    problem_dirichlet(P.get(), 0, nullptr);
    Problem *raw = P.get();
    S->cache[key] = std::move(P);
    return raw;
}

static void admit(ffh_server *S, ServedRequest *q)
{
    Context *ctx = S->ctx;
    hipStream_t s = ctx->stream;
    Problem *P = q->sym;
    // numeric assembly on the cached pattern, then the reduced values
    const int qdeg = P->family == FFH_FAMILY_H1 ? 2 * P->order + 4 : -1;
    problem_assemble(P, qdeg, FFH_COEFF_CELL, q->coeff.data(), q->mass, FFH_FP64);
    problem_reduce(P);
    const DevMatrix &A = P->reduced.m;
    const int n = q->n;
    q->val64.alloc(std::max<int64_t>(q->nnz, 1));
    FFH_HIP(hipMemcpyAsync(q->val64.p, A.val.p, sizeof(double) * q->nnz, hipMemcpyDeviceToDevice, s));
    q->val32.alloc(std::max<int64_t>(q->nnz, 1));
    hipLaunchKernelGGL(k_b_to_f32, dim3((unsigned)((q->nnz + 255) / 256)), dim3(256), 0, s, q->nnz,
                       q->val64.p, q->val32.p);
    q->dinv.alloc(n);
    hipLaunchKernelGGL(k_b_jacobi, dim3((n + 255) / 256), dim3(256), 0, s, n, A.col_ptr.p,
                       A.row_idx.p, q->val64.p, q->dinv.p);
    const size_t nr = (size_t)n * q->R;
    DBuf<double> tmp;
    tmp.upload(q->rhs.data(), q->rhs.size(), s);
    q->b.alloc(nr);
    hipLaunchKernelGGL(k_b_interleave, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, n,
                       q->nrhs, q->R, tmp.p, q->b.p);
    q->x.alloc(nr);
    FFH_HIP(hipGetLastError());
    FFH_HIP(hipMemsetAsync(q->x.p, 0, sizeof(double) * nr, s));
    q->prec = S->policy == 1 || (S->policy == 2 && q->family == FFH_FAMILY_NEDELEC) ? 1 : 0;
    if (q->prec) {
        alloc_fp64_vectors(q, s);
    } else {
        q->r.alloc(nr);
        q->p.alloc(nr);
        q->ap.alloc(nr);
        FFH_HIP(hipMemsetAsync(q->r.p, 0, sizeof(float) * nr, s));
        FFH_HIP(hipMemsetAsync(q->p.p, 0, sizeof(float) * nr, s));
    }
    q->best_rr.assign(q->nrhs, 1e300);
    q->stall.assign(q->nrhs, 0);
    FFH_HIP(hipStreamSynchronize(s)); // tmp and the cached problem's values are reused next
    q->slot = S->free_slots.back();
    S->free_slots.pop_back();
    for (int r = 0; r < kMaxRhs; ++r) {
        BatchScal &c = S->h_scal[(size_t)q->slot * kMaxRhs + r];
        c = BatchScal{};
        c.state = r < q->nrhs ? kFresh : kUnused;
        c.max_iter = q->max_iter;
        c.tol2 = q->tol * q->tol;
    }
    FFH_HIP(hipMemcpyAsync(S->d_scal.p + (size_t)q->slot * kMaxRhs,
                           S->h_scal.data() + (size_t)q->slot * kMaxRhs,
                           sizeof(BatchScal) * kMaxRhs, hipMemcpyHostToDevice, s));
    S->used += q->bytes;
    S->active.push_back(q);
    S->changed = true;
}

static void rebuild_batch(ffh_server *S)
{
    std::vector<BatchSys> sys;
    std::vector<int2> work;
    size_t scratch = 0;
    for (ServedRequest *q : S->active) {
        const DevMatrix &A = q->sym->reduced.m;
        BatchSys b{};
        b.ptr = A.col_ptr.p;
        b.idx = A.row_idx.p;
        b.val32 = q->val32.p;
        b.val64 = q->val64.p;
        b.dinv = q->dinv.p;
        b.b = q->b.p;
        b.x = q->x.p;
        b.r = q->r.p;
        b.p = q->p.p;
        b.ap = q->ap.p;
        b.r64 = q->r64.p;
        b.p64 = q->p64.p;
        b.ap64 = q->ap64.p;
        b.prec = q->prec;
        b.n = q->n;
        b.R = q->R;
        b.slot = q->slot;
        b.w0 = (int)work.size();
        for (int row = 0; row < q->n; row += kRowsPerItem)
            work.push_back(make_int2((int)sys.size(), row));
        b.w1 = (int)work.size();
        sys.push_back(b);
    }
    // FP64 A x scratch: a system's slice starts at w0 * kRowsPerItem * kMaxRhs
    scratch = work.size() * (size_t)kRowsPerItem * kMaxRhs;
    hipStream_t s = S->ctx->stream;
    S->d_sys.upload(sys, s);
    S->d_work.upload(work, s);
    const size_t np = std::max<size_t>(1, work.size() * kMaxRhs);
    if (S->d_p0.n < np) {
        S->d_p0.alloc(np);
        S->d_p1.alloc(np);
        S->d_p2.alloc(np);
        S->d_p3.alloc(np);
    }
    if (S->d_scratch.n < scratch) S->d_scratch.alloc(std::max<size_t>(scratch, 1));
    S->nwork = (int)work.size();
    FFH_HIP(hipStreamSynchronize(s));
    S->changed = false;
}

// One step of the batch: `chunk` PCG iterations of every active system, then
// the FP64 residual refresh (which also starts systems admitted since the
// last step: they sit out the iterations of their first step).
static void run_chunk(ffh_server *S)
{
    hipStream_t s = S->ctx->stream;
    const int W = S->nwork, NS = (int)S->active.size();
    if (W == 0 || NS == 0) return;
    for (int it = 0; it < S->chunk; ++it) {
        hipLaunchKernelGGL(k_b_spmv, dim3(W), dim3(256), 0, s, S->d_work.p, S->d_sys.p, S->d_scal.p,
                           S->d_p0.p);
        hipLaunchKernelGGL(k_b_alpha, dim3(NS), dim3(256), 0, s, S->d_sys.p, S->d_scal.p, S->d_p0.p);
        hipLaunchKernelGGL(k_b_update, dim3(W), dim3(256), 0, s, S->d_work.p, S->d_sys.p, S->d_scal.p,
                           S->d_p0.p, S->d_p1.p);
        hipLaunchKernelGGL(k_b_beta, dim3(NS), dim3(256), 0, s, S->d_sys.p, S->d_scal.p, S->d_p0.p,
                           S->d_p1.p);
        hipLaunchKernelGGL(k_b_dir, dim3(W), dim3(256), 0, s, S->d_work.p, S->d_sys.p, S->d_scal.p, 0);
    }
    hipLaunchKernelGGL(k_b_resid, dim3(W), dim3(256), 0, s, S->d_work.p, S->d_sys.p, S->d_scal.p,
                       S->d_scratch.p, S->d_p0.p, S->d_p1.p, S->d_p2.p, S->d_p3.p);
    hipLaunchKernelGGL(k_b_refresh, dim3(NS), dim3(256), 0, s, S->d_sys.p, S->d_scal.p, S->d_p0.p,
                       S->d_p1.p, S->d_p2.p, S->d_p3.p);
    hipLaunchKernelGGL(k_b_dir, dim3(W), dim3(256), 0, s, S->d_work.p, S->d_sys.p, S->d_scal.p, 1);
    FFH_HIP(hipGetLastError());
    S->iterations_run += S->chunk;
}

// An FP32-iterating request whose FP64 residual has not halved over 16
// refreshes (the FP32 operator cannot resolve further progress) continues
// in FP64 from its current x: its directions restart after the next refresh.
static void watch_stagnation(ffh_server *S, ServedRequest *q)
{
    bool stuck = false;
    for (int r = 0; r < q->nrhs; ++r) {
        const BatchScal &c = S->h_scal[(size_t)q->slot * kMaxRhs + r];
        if (c.state != kActive) continue;
        if (c.true_rr < 0.25 * q->best_rr[r]) {
            q->best_rr[r] = c.true_rr;
            q->stall[r] = 0;
        } else if (++q->stall[r] >= 16) {
            stuck = true;
        }
    }
    if (!stuck) return;
    hipStream_t s = S->ctx->stream;
    alloc_fp64_vectors(q, s);
    q->prec = 1;
    for (int r = 0; r < q->nrhs; ++r) {
        BatchScal &c = S->h_scal[(size_t)q->slot * kMaxRhs + r];
        if (c.state == kActive || c.state == kCandidate) {
            c.state = kCandidate;
            c.restart = 2;
        }
    }
    FFH_HIP(hipMemcpyAsync(S->d_scal.p + (size_t)q->slot * kMaxRhs,
                           S->h_scal.data() + (size_t)q->slot * kMaxRhs, sizeof(BatchScal) * kMaxRhs,
                           hipMemcpyHostToDevice, s));
    FFH_HIP(hipStreamSynchronize(s));
    q->r.reset(), q->p.reset(), q->ap.reset();
    S->changed = true;
    ++S->switches;
}

static void retire_finished(ffh_server *S, int *finished)
{
    hipStream_t s = S->ctx->stream;
    FFH_HIP(hipMemcpyAsync(S->h_scal.data(), S->d_scal.p, sizeof(BatchScal) * S->h_scal.size(),
                           hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<ServedRequest *> keep;
    for (ServedRequest *q : S->active) {
        bool all = true;
        for (int r = 0; r < q->nrhs; ++r) {
            const int st = S->h_scal[(size_t)q->slot * kMaxRhs + r].state;
            all = all && (st == kConverged || st == kStopped);
        }
        if (!all) {
            keep.push_back(q);
            if (q->prec == 0) watch_stagnation(S, q);
            continue;
        }
        const size_t nr = (size_t)q->n * q->R;
        std::vector<double> xi(nr);
        q->x.download(xi.data(), nr, s);
        FFH_HIP(hipStreamSynchronize(s));
        q->sol.assign((size_t)q->n * q->nrhs, 0.0);
        for (int r = 0; r < q->nrhs; ++r)
            for (int i = 0; i < q->n; ++i) q->sol[(size_t)r * q->n + i] = xi[(size_t)i * q->R + r];
        q->iters.resize(q->nrhs);
        q->converged.resize(q->nrhs);
        q->resid.resize(q->nrhs);
        for (int r = 0; r < q->nrhs; ++r) {
            const BatchScal &c = S->h_scal[(size_t)q->slot * kMaxRhs + r];
            q->iters[r] = c.iters;
            q->converged[r] = c.state == kConverged;
            q->resid[r] = c.bb > 0 ? std::sqrt(c.true_rr / c.bb) : 0.0;
        }
        q->val32.reset(), q->r.reset(), q->p.reset(), q->ap.reset();
        q->val64.reset(), q->dinv.reset(), q->b.reset(), q->x.reset();
        q->r64.reset(), q->p64.reset(), q->ap64.reset();
        q->rhs.clear();
        q->rhs.shrink_to_fit();
        S->free_slots.push_back(q->slot);
        q->slot = -1;
        q->done = true;
        S->used -= q->bytes;
        S->changed = true;
        ++*finished;
    }
    S->active.swap(keep);
}

void server_step(ffh_server *S, int *n_active, int *n_queued, int *n_finished)
{
    int finished = 0;
    while (!S->queue.empty() && !S->free_slots.empty() &&
           S->used + S->queue.front()->bytes <= S->budget) {
        ServedRequest *q = S->queue.front();
        S->queue.pop_front();
        admit(S, q);
    }
    if (S->changed) rebuild_batch(S);
    if (!S->active.empty()) {
        run_chunk(S);
        retire_finished(S, &finished);
    }
    if (n_active) *n_active = (int)S->active.size();
    if (n_queued) *n_queued = (int)S->queue.size();
    if (n_finished) *n_finished = finished;
}

int64_t server_submit(ffh_server *S, int mesh_key, int family, int order, const double *coeff,
                      double mass, int nrhs, const double *rhs, double tol, int max_iter)
{
    auto it = S->cache.find(ServerKey{mesh_key, family, order});
    FFH_REQUIRE(it != S->cache.end(), FFH_ERR_INVALID, "prepare the mesh before submitting");
    Problem *P = it->second.get();
    FFH_REQUIRE(nrhs >= 1 && nrhs <= kMaxRhs, FFH_ERR_INVALID, "1..8 right-hand sides");
    FFH_REQUIRE(coeff && rhs, FFH_ERR_INVALID, "coefficients and right-hand sides required");
    FFH_REQUIRE(tol > 0 && max_iter > 0, FFH_ERR_INVALID, "tolerance and iteration cap");
    auto q = std::make_unique<ServedRequest>();
    q->ticket = S->next_ticket++;
    q->sym = P;
    q->family = family;
    q->n = P->reduced.m.n;
    q->nnz = P->reduced.m.nnz;
    q->nrhs = nrhs;
    q->R = rhs_stride(nrhs);
    q->coeff.assign(coeff, coeff + P->ncell);
    q->mass = mass;
    q->rhs.assign(rhs, rhs + (size_t)q->n * nrhs);
    q->tol = tol;
    q->max_iter = max_iter;
    q->bytes = request_bytes(q->n, q->nnz, q->R);
    FFH_REQUIRE(q->bytes <= S->budget, FFH_ERR_ADMISSION,
                "request needs more device memory than the server budget");
    const int64_t t = q->ticket;
    S->queue.push_back(q.get());
    S->reqs[t] = std::move(q);
    return t;
}

} // namespace ffh

// ---------------------------------------------------------------- C-ABI --
namespace {
template <typename F>
int server_guard(ffh::Context *ctx, F &&f)
{
    try {
        f();
        return FFH_OK;
    } catch (const ffh::Error &e) {
        if (ctx) ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        return FFH_ERR_DEVICE;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return FFH_ERR_INVALID;
    }
}
} // namespace

extern "C" {

int ffh_server_create(ffh_context *ctx, int64_t budget_bytes, int max_systems, int chunk,
                      ffh_server **out)
{
    if (!ctx) return FFH_ERR_HANDLE;
    if (!out || budget_bytes <= 0 || max_systems < 1 || max_systems > 4096 || chunk < 1)
        return FFH_ERR_INVALID;
    *out = nullptr;
    ffh_server *S = nullptr;
    int rc = server_guard(ctx, [&] {
        S = new ffh_server();
        S->ctx = ctx;
        S->budget = (size_t)budget_bytes;
        S->max_sys = max_systems;
        S->chunk = chunk;
        S->h_scal.assign((size_t)max_systems * ffh::kMaxRhs, ffh::BatchScal{});
        S->d_scal.alloc(S->h_scal.size());
        for (int i = max_systems - 1; i >= 0; --i) S->free_slots.push_back(i);
    });
    if (rc != FFH_OK) {
        delete S;
        return rc;
    }
    *out = S;
    return FFH_OK;
}

int ffh_server_free(ffh_server *S)
{
    if (!S) return FFH_ERR_HANDLE;
    (void)hipStreamSynchronize(S->ctx->stream);
    delete S;
    return FFH_OK;
}

int ffh_server_prepare(ffh_server *S, int mesh_key, int family, int order, int nvert,
                       const double *vertices, int ncell, const int *cells, int *n_free,
                       int64_t *nnz_free)
{
    if (!S) return FFH_ERR_HANDLE;
    return server_guard(S->ctx, [&] {
        ffh::Problem *P = ffh::server_prepare(S, mesh_key, family, order, 3, nvert, vertices,
                                              ncell, cells);
        if (n_free) *n_free = P->reduced.m.n;
        if (nnz_free) *nnz_free = P->reduced.m.nnz;
    });
}

int ffh_server_submit(ffh_server *S, int mesh_key, int family, int order, const double *cell_coeff,
                      double mass, int nrhs, const double *rhs, double tol, int max_iterations,
                      int64_t *ticket)
{
    if (!S) return FFH_ERR_HANDLE;
    if (!ticket) return FFH_ERR_INVALID;
    return server_guard(S->ctx, [&] {
        *ticket = ffh::server_submit(S, mesh_key, family, order, cell_coeff, mass, nrhs, rhs, tol,
                                     max_iterations);
    });
}

int ffh_server_step(ffh_server *S, int *n_active, int *n_queued, int *n_finished)
{
    if (!S) return FFH_ERR_HANDLE;
    return server_guard(S->ctx, [&] { ffh::server_step(S, n_active, n_queued, n_finished); });
}

int ffh_server_poll(ffh_server *S, int64_t ticket, double *solution, int *iterations,
                    int *converged, double *rel_residual)
{
    if (!S) return FFH_ERR_HANDLE;
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return FFH_ERR_HANDLE;
    ffh::ServedRequest *q = it->second.get();
    if (!q->done) return 1;
    if (solution) std::copy(q->sol.begin(), q->sol.end(), solution);
    for (int r = 0; r < q->nrhs; ++r) {
        if (iterations) iterations[r] = q->iters[r];
        if (converged) converged[r] = q->converged[r];
        if (rel_residual) rel_residual[r] = q->resid[r];
    }
    S->reqs.erase(it);
    return FFH_OK;
}

int ffh_server_progress(const ffh_server *S, int64_t ticket, int *state, int *iterations,
                        double *rel_residual)
{
    if (!S) return FFH_ERR_HANDLE;
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return FFH_ERR_HANDLE;
    const ffh::ServedRequest *q = it->second.get();
    for (int r = 0; r < q->nrhs; ++r) {
        int st = q->done ? (q->converged[r] ? ffh::kConverged : ffh::kStopped) : -1;
        int itn = q->done ? q->iters[r] : 0;
        double rel = q->done ? q->resid[r] : -1.0;
        if (!q->done && q->slot >= 0) {
            const ffh::BatchScal &c = S->h_scal[(size_t)q->slot * ffh::kMaxRhs + r];
            st = c.state;
            itn = c.iters;
            rel = c.bb > 0 ? std::sqrt(c.true_rr / c.bb) : 0.0;
        }
        if (state) state[r] = st;
        if (iterations) iterations[r] = itn;
        if (rel_residual) rel_residual[r] = rel;
    }
    return FFH_OK;
}

int ffh_server_stats(const ffh_server *S, int64_t *used_bytes, int64_t *iterations_run)
{
    if (!S) return FFH_ERR_HANDLE;
    if (used_bytes) *used_bytes = (int64_t)S->used;
    if (iterations_run) *iterations_run = S->iterations_run;
    return FFH_OK;
}

int ffh_server_set_precision(ffh_server *S, int policy, int64_t *switches)
{
    if (!S) return FFH_ERR_HANDLE;
    if (policy < -1 || policy > 2) return FFH_ERR_INVALID;
    if (policy >= 0) S->policy = policy;
    if (switches) *switches = S->switches;
    return FFH_OK;
}

} // extern "C"
