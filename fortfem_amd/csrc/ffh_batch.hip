// ffh_batch.hip -- served solves: many resident systems with several
// right-hand sides each, advanced together by one launch per PCG phase.
//
// Reference loop: src/solvers/advanced_solvers.f90 pcg_iteration_loop
// (:880-925), compute_solver_tolerance (:1063-1073), Jacobi diagonal
// (:1124-1155).  Each (system, right-hand side) runs its own Jacobi-PCG with
// x0 = 0 and the relative stop ||r|| <= tol ||b||.
//
// Design: the active systems form one block-diagonal operator.  Each system's
// matrix is held in SELL-64 form (slices of 64 rows, one lane per row, every
// index/value load of a wave 64 consecutive words); the structure belongs to
// the cached pattern, the values to the request.  A work table lists items of
// four slices (256 rows, one block) of one system; every launch covers the
// whole table.  The right-hand sides of a system are interleaved (row-major
// n x R, R in {1, 2, 4, 6, 8}) so one pass over the matrix serves all of them
// and each lane moves its row's R values with vector loads.  Dot products
// are per-item partials, summed per system in item order by one block per
// system (deterministic: a system's result does not depend on what else is
// in the batch).  The inner iteration runs in FP32 (matrix, r, p, Ap, D^-1)
// with FP64 x and FP64 reductions; at the end of every chunk of iterations
// the residual is recomputed in FP64 from the FP64 matrix (reliable update:
// r is replaced, p kept while the recurrence tracked the true residual),
// which decides convergence.  Systems join and leave only at chunk
// boundaries; admission is by the device-memory footprint of the request.
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

constexpr int kRowsPerItem = 256; // four SELL-64 slices, one block
constexpr int kMaxRhs = 8;

struct BatchSys {
    const int *off, *sidx, *srow; // SELL-64 structure of the cached reduced pattern
    // column tiles of that SELL copy (the row map, tile offsets and tile
    // columns in one array; 16-bit tile positions per entry), used when the
    // autotuner picked the tiled product for this system: tiled is then the
    // pattern's largest tile (columns), else 0
    const int *tpack;
    const uint16_t *tlidx;
    int tiled;
    // lanes per row of the row-split product (2, 4 or 8, dividing R), or 0
    int split;
    const float *val32;    // SELL values of the request
    const double *val64;
    const float *dinv32;
    const double *dinv64;
    const double *b; // n x R
    double *x;       // n x R
    // FP32 iteration: the steps alpha p of the chunk in FP32 (n x R), added
    // to x and cleared by k_b_fold before each refresh
    float *dx;
    float *r, *p, *ap;        // FP32 iteration
    double *r64, *p64, *ap64; // FP64 iteration (prec = 1)
    int n, R, w0, w1, slot, prec, nslice;
    int x_zero; // x is zero when the system starts (its first residual is b)
};

// per (system slot, right-hand side)
enum BatchState : int { kActive = 0, kConverged = 1, kStopped = 2, kUnused = 3, kCandidate = 4,
                        kFresh = 5 };
struct BatchScal {
    double rz, alpha, beta, rr, bb, tol2, true_rr, init_rr;
    int iters, state, max_iter, restart;
    int flags; // kAbsTol: stop at ||r|| <= tol instead of tol ||r0||
    int pad;
};
constexpr int kAbsTol = 1;
// BatchScal::restart codes: 1 restart the direction, 2 breakdown (restart
// after the refresh), 3 converged by the FP64 recurrence, kCapped stopped
// at the iteration cap (the refresh records the true residual of x),
// kRecurrence32 the FP32 recurrence residual reached the tolerance (the
// refresh decides on the FP64 residual)
constexpr int kCapped = 4;
constexpr int kRecurrence32 = 5;

// R values of one row, moved with vector accesses (R = 6: three pairs)
template <typename T, int R>
struct alignas(sizeof(T) * ((R & (R - 1)) == 0 ? R : 2)) Pack {
    T v[R];
};

template <typename T, int R>
__device__ __forceinline__ Pack<T, R> ld(const T *base, size_t row)
{
    return *reinterpret_cast<const Pack<T, R> *>(base + row * R);
}

template <typename T, int R>
__device__ __forceinline__ void st(T *base, size_t row, const Pack<T, R> &v)
{
    *reinterpret_cast<Pack<T, R> *>(base + row * R) = v;
}

// ---------------------------------------------------------------- kernels --
// Per-right-hand-side totals over a 256-thread block: wave shuffles, then the
// four waves in order; total r is written by thread r to out[r].
template <int R>
__device__ __forceinline__ void block_sums(double (&d)[R], double *sh, double *out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double v = d[r];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) sh[wv * R + r] = v;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int r = threadIdx.x;
        out[r] = sh[r] + sh[R + r] + sh[2 * R + r] + sh[3 * R + r];
    }
    __syncthreads();
}

// the R scalar states of a system, broadcast through LDS
struct ScalView {
    double alpha[kMaxRhs], beta[kMaxRhs];
    int state[kMaxRhs], restart[kMaxRhs];
    int any_active;
};

__device__ __forceinline__ void load_scal(const BatchScal *__restrict__ sc, int slot, int R,
                                          ScalView &v)
{
    if (threadIdx.x < R) {
        const BatchScal q = sc[slot * kMaxRhs + threadIdx.x];
        v.alpha[threadIdx.x] = q.alpha;
        v.beta[threadIdx.x] = q.beta;
        v.state[threadIdx.x] = q.state;
        v.restart[threadIdx.x] = q.restart;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        for (int r = 0; r < R; ++r) a |= v.state[r] == kActive;
        v.any_active = a;
    }
    __syncthreads();
}

// Ap = A p over the item's four slices (wave = slice, lane = row) and the
// item's partials of p.Ap per right-hand side.  Entries go in steps of U:
// the step's index/value loads are issued together, then its gathers of the
// R-wide rows of p.
template <int R, typename MT, typename VT>
__device__ __forceinline__ void item_spmv(const BatchSys &s, int slice0, const MT *__restrict__ val,
                                          const VT *__restrict__ p, VT *__restrict__ ap,
                                          double (&d)[R])
{
    constexpr int U = R >= 4 ? 4 : 8;
    const int lane = threadIdx.x & 63;
    const int sl = slice0 + (threadIdx.x >> 6);
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = 0.0;
    if (sl >= s.nslice) return;
    const int o = s.off[sl], w = (s.off[sl + 1] - o) >> 6;
    const int *ip = s.sidx + o + lane;
    const MT *vp = val + o + lane;
    VT acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = VT(0);
    int j = 0;
    if (w >= U) {
        // software pipeline: the next step's index/value loads are issued
        // while this step's gathers are in flight
        int c[U];
        VT v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = ldm(ip + 64 * u);
            v[u] = (VT)ldm(vp + 64 * u);
        }
        for (; j + U <= w; j += U) {
            Pack<VT, R> pc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pc[u] = ld<VT, R>(p, c[u]);
            const bool more = j + 2 * U <= w; // uniform: one slice per wave
            int cn[U];
            VT vn[U];
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    cn[u] = ldm(ip + 64 * (U + u));
                    vn[u] = (VT)ldm(vp + 64 * (U + u));
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += v[u] * pc[u].v[r];
            ip += 64 * U;
            vp += 64 * U;
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    c[u] = cn[u];
                    v[u] = vn[u];
                }
            }
        }
    }
    for (; j < w; ++j) {
        const int c = ldm(ip);
        const VT v = (VT)ldm(vp);
        const Pack<VT, R> pc = ld<VT, R>(p, c);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += v * pc.v[r];
        ip += 64;
        vp += 64;
    }
    const int row = s.srow[sl * 64 + lane];
    if (row >= 0) {
        const Pack<VT, R> own = ld<VT, R>(p, row);
        Pack<VT, R> out;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            out.v[r] = acc[r];
            d[r] = (double)own.v[r] * (double)acc[r];
        }
        st<VT, R>(ap, row, out);
    }
}

// The same product with the item's column tile staged in LDS: the item (a
// group of four slices) gathers the R-wide rows of p at its distinct
// columns once, then every lane adds its row's entries from LDS through the
// 16-bit tile positions.  Each lane sums its products in the same order as
// item_spmv, so both give the same A p to the bit.
template <int R, typename MT, typename VT>
__device__ __forceinline__ void item_spmv_tiled(const BatchSys &s, int slice0,
                                                const MT *__restrict__ val,
                                                const VT *__restrict__ p, VT *__restrict__ ap,
                                                double (&d)[R])
{
    extern __shared__ __align__(16) unsigned char b_dyn_lds[];
    VT *tile = reinterpret_cast<VT *>(b_dyn_lds);
    constexpr int U = R >= 4 ? 4 : 8;
    const int ngroup = (s.nslice + 3) >> 2;
    const int g = slice0 >> 2;
    const int *toff = s.tpack + (size_t)s.nslice * 64;
    const int *tcols = toff + ngroup + 1;
    const int c0 = toff[g], nc = toff[g + 1] - c0;
    for (int k = threadIdx.x; k < nc; k += 256) st<VT, R>(tile, k, ld<VT, R>(p, tcols[c0 + k]));
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int sl = slice0 + (threadIdx.x >> 6);
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = 0.0;
    if (sl >= s.nslice) return;
    const int o = s.off[sl], w = (s.off[sl + 1] - o) >> 6;
    const uint16_t *ip = s.tlidx + o + lane;
    const MT *vp = val + o + lane;
    VT acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = VT(0);
    int j = 0;
    for (; j + U <= w; j += U) {
        int c[U];
        VT v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = ldm(ip + 64 * u);
            v[u] = (VT)ldm(vp + 64 * u);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Pack<VT, R> pc = ld<VT, R>(tile, c[u]);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += v[u] * pc.v[r];
        }
        ip += 64 * U;
        vp += 64 * U;
    }
    for (; j < w; ++j) {
        const Pack<VT, R> pc = ld<VT, R>(tile, (int)ldm(ip));
        const VT v = (VT)ldm(vp);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += v * pc.v[r];
        ip += 64;
        vp += 64;
    }
    const int row = s.srow[sl * 64 + lane];
    if (row >= 0) {
        const Pack<VT, R> own = ld<VT, R>(p, row);
        Pack<VT, R> out;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            out.v[r] = acc[r];
            d[r] = (double)own.v[r] * (double)acc[r];
        }
        st<VT, R>(ap, row, out);
    }
}

// The same product with L lanes per row, each holding Q = R / L of the row's
// right-hand sides: a wave takes 64 / L rows of its slice per pass (L passes),
// the L lanes of a row load its index and value at one address and its
// R-wide row of p as one contiguous run, so a gather touches 64 / L rows of
// p instead of 64.  Each lane adds its products in the order of item_spmv
// (the same A p to the bit), and the products p.Ap go through LDS back to
// one lane per row, so the block sums see what item_spmv gives them.
template <int R, int L, typename MT, typename VT>
__device__ __forceinline__ void item_spmv_split(const BatchSys &s, int slice0,
                                                const MT *__restrict__ val,
                                                const VT *__restrict__ p, VT *__restrict__ ap,
                                                double (&d)[R])
{
    static_assert(R % L == 0 && 64 % L == 0 && ((R / L) & (R / L - 1)) == 0, "row split");
    extern __shared__ __align__(16) unsigned char b_dyn_lds[];
    constexpr int Q = R / L, G = 64 / L, U = 8;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int sl = slice0 + wv;
    // the wave's 64 rows x R products (4 x 64 x R doubles of dynamic LDS)
    double *dw = reinterpret_cast<double *>(b_dyn_lds) + wv * 64 * R;
    const int sub = lane % L, r0 = sub * Q;
    if (sl < s.nslice) { // uniform in the wave
        const int o = s.off[sl], w = (s.off[sl + 1] - o) >> 6;
        for (int ps = 0; ps < L; ++ps) {
            const int rl = ps * G + lane / L;
            const int *ip = s.sidx + o + rl;
            const MT *vp = val + o + rl;
            VT acc[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q] = VT(0);
            int j = 0;
            for (; j + U <= w; j += U) {
                int c[U];
                VT v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    c[u] = ldm(ip + 64 * u);
                    v[u] = (VT)ldm(vp + 64 * u);
                }
                Pack<VT, Q> pc[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    pc[u] = *reinterpret_cast<const Pack<VT, Q> *>(p + (size_t)c[u] * R + r0);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int q = 0; q < Q; ++q) acc[q] += v[u] * pc[u].v[q];
                ip += 64 * U;
                vp += 64 * U;
            }
            for (; j < w; ++j) {
                const int c = ldm(ip);
                const VT v = (VT)ldm(vp);
                const Pack<VT, Q> pc = *reinterpret_cast<const Pack<VT, Q> *>(p + (size_t)c * R + r0);
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] += v * pc.v[q];
                ip += 64;
                vp += 64;
            }
            const int row = s.srow[sl * 64 + rl];
            if (row >= 0) {
                const Pack<VT, Q> own =
                    *reinterpret_cast<const Pack<VT, Q> *>(p + (size_t)row * R + r0);
                Pack<VT, Q> out;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    out.v[q] = acc[q];
                    dw[rl * R + r0 + q] = (double)own.v[q] * (double)acc[q];
                }
                *reinterpret_cast<Pack<VT, Q> *>(ap + (size_t)row * R + r0) = out;
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) dw[rl * R + r0 + q] = 0.0;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = sl < s.nslice ? dw[lane * R + r] : 0.0;
}

template <int R, int L>
__device__ __forceinline__ void spmv_split(const BatchSys &s, int slice0, double (&d)[R])
{
    if (s.prec)
        item_spmv_split<R, L, double, double>(s, slice0, s.val64, s.p64, s.ap64, d);
    else
        item_spmv_split<R, L, float, float>(s, slice0, s.val32, s.p, s.ap, d);
}

// lanes per row that the row-split form offers for R right-hand sides
__host__ __device__ constexpr bool split_ok(int R, int L)
{
    return (L == 2 && (R == 2 || R == 4 || R == 8)) || (L == 4 && (R == 4 || R == 8)) ||
           (L == 8 && R == 8);
}

template <int R>
__device__ __forceinline__ void spmv_item(const BatchSys &s, int slice0, double *sh, double *out)
{
    double d[R];
    if (split_ok(R, 2) && s.split) {
        if constexpr (R == 2) {
            spmv_split<2, 2>(s, slice0, d);
        } else if constexpr (R == 4) {
            if (s.split == 4) spmv_split<4, 4>(s, slice0, d);
            else spmv_split<4, 2>(s, slice0, d);
        } else if constexpr (R == 8) {
            if (s.split == 8) spmv_split<8, 8>(s, slice0, d);
            else if (s.split == 4) spmv_split<8, 4>(s, slice0, d);
            else spmv_split<8, 2>(s, slice0, d);
        }
    } else if (s.tiled) {
        if (s.prec)
            item_spmv_tiled<R, double, double>(s, slice0, s.val64, s.p64, s.ap64, d);
        else
            item_spmv_tiled<R, float, float>(s, slice0, s.val32, s.p, s.ap, d);
    } else {
        if (s.prec)
            item_spmv<R, double, double>(s, slice0, s.val64, s.p64, s.ap64, d);
        else
            item_spmv<R, float, float>(s, slice0, s.val32, s.p, s.ap, d);
    }
    block_sums<R>(d, sh, out);
}

// Ap = A p of every item with an active right-hand side; partials of p.Ap
__global__ __launch_bounds__(256) void k_b_spmv(const int2 *__restrict__ work,
                                                const BatchSys *__restrict__ sys,
                                                const BatchScal *__restrict__ sc,
                                                double *__restrict__ part)
{
    __shared__ double sh[4 * kMaxRhs];
    __shared__ ScalView sv;
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    load_scal(sc, s.slot, s.R, sv);
    if (!sv.any_active) return;
    double *out = part + (size_t)blockIdx.x * kMaxRhs;
    const int slice0 = wk.y / 64;
    switch (s.R) {
    case 1: spmv_item<1>(s, slice0, sh, out); break;
    case 2: spmv_item<2>(s, slice0, sh, out); break;
    case 4: spmv_item<4>(s, slice0, sh, out); break;
    case 6: spmv_item<6>(s, slice0, sh, out); break;
    default: spmv_item<8>(s, slice0, sh, out); break;
    }
}

// Totals of a system's per-item partials (items w0 .. w1) for each
// right-hand side, in a block of kTotThreads: thread t adds right-hand side
// t & 7 of items w0 + (t >> 3) + 128 j, K items per round with their loads
// issued together (one load of a wave covers 8 items x 8 right-hand sides,
// contiguous), then the 8 lanes of a right-hand side in each wave and the 16
// waves are added in a fixed order.  Thread r < 8 holds total r after the
// call; sh holds NQ x 16 x 8 doubles.
constexpr int kTotThreads = 1024;
template <int NQ>
__device__ __forceinline__ void system_totals(const BatchSys &s, const double *const (&part)[NQ],
                                              double (&tot)[NQ], double *sh)
{
    constexpr int K = NQ == 1 ? 8 : NQ == 2 ? 4 : 2;
    const int r = threadIdx.x & 7, il = threadIdx.x >> 3;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    if (r < s.R) {
        for (int w = s.w0 + il; w < s.w1; w += 128 * K) {
            double v[K][NQ];
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int wk = w + 128 * k;
                    v[k][q] = wk < s.w1 ? part[q][(size_t)wk * kMaxRhs + r] : 0.0;
                }
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int q = 0; q < NQ; ++q) acc[q] += v[k][q];
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double v = acc[q];
        for (int o = 32; o >= 8; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane < 8) sh[(q * 16 + wv) * 8 + lane] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double t = 0.0;
        if (threadIdx.x < 8)
            for (int j = 0; j < 16; ++j) t += sh[(q * 16 + j) * 8 + threadIdx.x];
        tot[q] = t;
    }
    __syncthreads();
}

// per system: p.Ap -> alpha (one block per system)
__global__ __launch_bounds__(kTotThreads) void k_b_alpha(const BatchSys *__restrict__ sys,
                                                         BatchScal *__restrict__ sc,
                                                         const double *__restrict__ part)
{
    __shared__ double sh[1 * 16 * 8];
    const BatchSys s = sys[blockIdx.x];
    const double *const parts[1] = {part};
    double tot[1];
    system_totals<1>(s, parts, tot, sh);
    if (threadIdx.x < s.R) {
        BatchScal &q = sc[s.slot * kMaxRhs + threadIdx.x];
        const double t = tot[0];
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

// x += alpha p (FP64 iteration: x itself; FP32 iteration: the chunk's FP32
// correction dx), r -= alpha Ap (active right-hand sides), partials of r.r
// and r.(D r); thread = row, R values per thread
template <int R, typename VT>
__device__ __forceinline__ void update_item(const BatchSys &s, int row, const ScalView &sv,
                                            VT *__restrict__ xv, VT *__restrict__ rv,
                                            const VT *__restrict__ pv,
                                            const VT *__restrict__ apv, const VT *__restrict__ dinv,
                                            double (&crr)[R], double (&crz)[R])
{
#pragma unroll
    for (int r = 0; r < R; ++r) crr[r] = crz[r] = 0.0;
    if (row >= s.n) return;
    Pack<VT, R> x = ld<VT, R>(xv, row);
    Pack<VT, R> ri = ld<VT, R>(rv, row);
    const Pack<VT, R> pi = ld<VT, R>(pv, row), ai = ld<VT, R>(apv, row);
    const double di = (double)dinv[row];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (sv.state[r] != kActive) continue;
        const double a = sv.alpha[r];
        x.v[r] += (VT)a * pi.v[r];
        ri.v[r] = ri.v[r] - (VT)a * ai.v[r];
        const double rr = (double)ri.v[r];
        crr[r] = rr * rr;
        crz[r] = rr * (di * rr);
    }
    st<VT, R>(xv, row, x);
    st<VT, R>(rv, row, ri);
}

template <int R>
__device__ __forceinline__ void update_rows(const BatchSys &s, int row0, const ScalView &sv,
                                            double *sh, double *o_rr, double *o_rz)
{
    double crr[R], crz[R];
    const int row = row0 + threadIdx.x;
    if (s.prec)
        update_item<R, double>(s, row, sv, s.x, s.r64, s.p64, s.ap64, s.dinv64, crr, crz);
    else
        update_item<R, float>(s, row, sv, s.dx, s.r, s.p, s.ap, s.dinv32, crr, crz);
    block_sums<R>(crr, sh, o_rr);
    block_sums<R>(crz, sh, o_rz);
}

__global__ __launch_bounds__(256) void k_b_update(const int2 *__restrict__ work,
                                                  const BatchSys *__restrict__ sys,
                                                  const BatchScal *__restrict__ sc,
                                                  double *__restrict__ part_rr,
                                                  double *__restrict__ part_rz)
{
    __shared__ double sh[4 * kMaxRhs];
    __shared__ ScalView sv;
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    load_scal(sc, s.slot, s.R, sv);
    if (!sv.any_active) return;
    double *o_rr = part_rr + (size_t)blockIdx.x * kMaxRhs;
    double *o_rz = part_rz + (size_t)blockIdx.x * kMaxRhs;
    switch (s.R) {
    case 1: update_rows<1>(s, wk.y, sv, sh, o_rr, o_rz); break;
    case 2: update_rows<2>(s, wk.y, sv, sh, o_rr, o_rz); break;
    case 4: update_rows<4>(s, wk.y, sv, sh, o_rr, o_rz); break;
    case 6: update_rows<6>(s, wk.y, sv, sh, o_rr, o_rz); break;
    default: update_rows<8>(s, wk.y, sv, sh, o_rr, o_rz); break;
    }
}

// per system: beta, stop tests
__global__ __launch_bounds__(kTotThreads) void k_b_beta(const BatchSys *__restrict__ sys,
                                                        BatchScal *__restrict__ sc,
                                                        const double *__restrict__ part_rr,
                                                        const double *__restrict__ part_rz)
{
    __shared__ double sh[2 * 16 * 8];
    const BatchSys s = sys[blockIdx.x];
    const double *const parts[2] = {part_rr, part_rz};
    double tot[2];
    system_totals<2>(s, parts, tot, sh);
    if (threadIdx.x < s.R) {
        BatchScal &q = sc[s.slot * kMaxRhs + threadIdx.x];
        const double trr = tot[0], trz = tot[1];
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
            } else if (trr <= q.tol2) {
                // FP32 iteration: the right-hand side stops iterating until
                // the chunk's FP64 refresh, which converges it or, if the
                // FP64 residual disagrees, lets it go on from there
                q.state = kCandidate;
                q.restart = kRecurrence32;
            } else if (q.iters >= q.max_iter) {
                // iteration cap: x stays as it is and the chunk's FP64
                // refresh measures its true residual, which decides
                // converged or stopped (finalize_solver_status, :1075-1086)
                q.state = kCandidate;
                q.restart = kCapped;
            }
        }
    }
}

// p = D r + beta p (a restart, p = D r, never reads the old direction)
template <int R, typename VT>
__device__ __forceinline__ void dir_item(const BatchSys &s, int row, const ScalView &sv,
                                         const VT *__restrict__ rv, VT *__restrict__ pv,
                                         const VT *__restrict__ dinv, bool restart_only)
{
    if (row >= s.n) return;
    const Pack<VT, R> ri = ld<VT, R>(rv, row);
    Pack<VT, R> pi = ld<VT, R>(pv, row);
    const VT di = dinv[row];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (sv.state[r] != kActive) continue;
        const bool rs = restart_only ? sv.restart[r] != 0 : false;
        if (restart_only && !rs) continue;
        const VT z = di * ri.v[r];
        pi.v[r] = rs ? z : z + (VT)sv.beta[r] * pi.v[r];
    }
    st<VT, R>(pv, row, pi);
}

template <int R>
__device__ __forceinline__ void dir_rows(const BatchSys &s, int row0, const ScalView &sv,
                                         bool restart_only)
{
    const int row = row0 + threadIdx.x;
    if (s.prec)
        dir_item<R, double>(s, row, sv, s.r64, s.p64, s.dinv64, restart_only);
    else
        dir_item<R, float>(s, row, sv, s.r, s.p, s.dinv32, restart_only);
}

// (restart_only: p = D r where the refresh asked for a restart)
__global__ __launch_bounds__(256) void k_b_dir(const int2 *__restrict__ work,
                                               const BatchSys *__restrict__ sys,
                                               const BatchScal *__restrict__ sc, int restart_only)
{
    __shared__ ScalView sv;
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    load_scal(sc, s.slot, s.R, sv);
    if (!sv.any_active) return;
    switch (s.R) {
    case 1: dir_rows<1>(s, wk.y, sv, restart_only != 0); break;
    case 2: dir_rows<2>(s, wk.y, sv, restart_only != 0); break;
    case 4: dir_rows<4>(s, wk.y, sv, restart_only != 0); break;
    case 6: dir_rows<6>(s, wk.y, sv, restart_only != 0); break;
    default: dir_rows<8>(s, wk.y, sv, restart_only != 0); break;
    }
}

// FP64 residual b - A x of the item's rows (reliable update): r <- residual
// (FP32 copy, or FP64 for a system iterating in FP64), partials of r.r,
// r.(D r), b.b and of the recurrence drift.  A fresh system whose x is zero
// takes r = b without a product.
template <int R>
__device__ __forceinline__ void resid_item(const BatchSys &s, int slice0, const ScalView &sv,
                                           bool fresh, double *sh, double *o_rr, double *o_rz,
                                           double *o_bb, double *o_dd)
{
    const int lane = threadIdx.x & 63;
    const int sl = slice0 + (threadIdx.x >> 6);
    double crr[R], crz[R], cbb[R], cdd[R];
#pragma unroll
    for (int r = 0; r < R; ++r) crr[r] = crz[r] = cbb[r] = cdd[r] = 0.0;
    if (sl < s.nslice) {
        double acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.0;
        if (!fresh || !s.x_zero) {
            const int o = s.off[sl], w = (s.off[sl + 1] - o) >> 6;
            const int *ip = s.sidx + o + lane;
            const double *vp = s.val64 + o + lane;
            constexpr int U = 4;
            for (int j = 0; j < w; j += U) {
                int c[U];
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool in = j + u < w;
                    c[u] = in ? ldm(ip + 64 * u) : 0;
                    v[u] = in ? ldm(vp + 64 * u) : 0.0;
                }
                Pack<double, R> xc[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (j + u < w) xc[u] = ld<double, R>(s.x, c[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (j + u < w)
#pragma unroll
                        for (int r = 0; r < R; ++r) acc[r] += v[u] * xc[u].v[r];
                ip += 64 * U;
                vp += 64 * U;
            }
        }
        const int row = s.srow[sl * 64 + lane];
        if (row >= 0) {
            const Pack<double, R> bi = ld<double, R>(s.b, row);
            const double di = s.dinv64[row];
            Pack<double, R> r64;
            Pack<float, R> r32;
            if (s.prec) r64 = ld<double, R>(s.r64, row);
            else r32 = ld<float, R>(s.r, row);
            bool w64 = false;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int stt = sv.state[r];
                const double ri = bi.v[r] - acc[r];
                if (stt == kActive || stt == kCandidate || stt == kFresh) {
                    // recurrence vs FP64 residual; an FP64 recurrence is kept
                    // as it is (the reference's iteration), only started from b
                    const double dr = (s.prec ? r64.v[r] : (double)r32.v[r]) - ri;
                    cdd[r] = dr * dr;
                    if (!s.prec) {
                        r32.v[r] = (float)ri;
                    } else if (stt == kFresh || sv.restart[r] == 2) {
                        r64.v[r] = ri;
                        w64 = true;
                    }
                }
                crr[r] = ri * ri;
                crz[r] = ri * (di * ri);
                cbb[r] = bi.v[r] * bi.v[r];
            }
            if (!s.prec) st<float, R>(s.r, row, r32);
            else if (w64) st<double, R>(s.r64, row, r64);
        }
    }
    block_sums<R>(crr, sh, o_rr);
    block_sums<R>(crz, sh, o_rz);
    block_sums<R>(cbb, sh, o_bb);
    block_sums<R>(cdd, sh, o_dd);
}

// Before each refresh, FP32-iterating systems: x += dx, dx = 0 (the chunk's
// steps, accumulated in FP32 by the update launch, folded into the FP64 x
// that the residual and every result read).  One block per work item, a
// thread per row, the R values of a row.
template <int R>
__device__ __forceinline__ void fold_rows(const BatchSys &s, int row)
{
    if (row >= s.n) return;
    Pack<double, R> x = ld<double, R>(s.x, row);
    const Pack<float, R> d = ld<float, R>(s.dx, row);
    Pack<float, R> z;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        x.v[r] += (double)d.v[r];
        z.v[r] = 0.0f;
    }
    st<double, R>(s.x, row, x);
    st<float, R>(s.dx, row, z);
}

__global__ __launch_bounds__(256) void k_b_fold(const int2 *__restrict__ work,
                                                const BatchSys *__restrict__ sys)
{
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    if (s.prec || !s.dx) return;
    const int row = wk.y + threadIdx.x;
    switch (s.R) {
    case 1: fold_rows<1>(s, row); break;
    case 2: fold_rows<2>(s, row); break;
    case 4: fold_rows<4>(s, row); break;
    case 6: fold_rows<6>(s, row); break;
    default: fold_rows<8>(s, row); break;
    }
}

__global__ __launch_bounds__(256) void k_b_resid(const int2 *__restrict__ work,
                                                 const BatchSys *__restrict__ sys,
                                                 const BatchScal *__restrict__ sc,
                                                 double *__restrict__ part_rr,
                                                 double *__restrict__ part_rz,
                                                 double *__restrict__ part_bb,
                                                 double *__restrict__ part_dd)
{
    __shared__ double sh[4 * kMaxRhs];
    __shared__ ScalView sv;
    const int2 wk = work[blockIdx.x];
    const BatchSys s = sys[wk.x];
    load_scal(sc, s.slot, s.R, sv);
    bool live = false, fresh = true;
    for (int r = 0; r < s.R; ++r) {
        const int stt = sv.state[r];
        live = live || stt == kActive || stt == kCandidate || stt == kFresh;
        fresh = fresh && (stt == kFresh || stt == kUnused);
    }
    if (!live) return;
    const size_t o = (size_t)blockIdx.x * kMaxRhs;
    const int slice0 = wk.y / 64;
    switch (s.R) {
    case 1: resid_item<1>(s, slice0, sv, fresh, sh, part_rr + o, part_rz + o, part_bb + o, part_dd + o); break;
    case 2: resid_item<2>(s, slice0, sv, fresh, sh, part_rr + o, part_rz + o, part_bb + o, part_dd + o); break;
    case 4: resid_item<4>(s, slice0, sv, fresh, sh, part_rr + o, part_rz + o, part_bb + o, part_dd + o); break;
    case 6: resid_item<6>(s, slice0, sv, fresh, sh, part_rr + o, part_rz + o, part_bb + o, part_dd + o); break;
    default: resid_item<8>(s, slice0, sv, fresh, sh, part_rr + o, part_rz + o, part_bb + o, part_dd + o); break;
    }
}

// per system: true residual -> converged, or continue (reliable update keeps
// p; a breakdown, or a fresh system, restarts p = D r)
__global__ __launch_bounds__(kTotThreads) void k_b_refresh(const BatchSys *__restrict__ sys,
                                                           BatchScal *__restrict__ sc,
                                                           const double *__restrict__ part_rr,
                                                           const double *__restrict__ part_rz,
                                                           const double *__restrict__ part_bb,
                                                           const double *__restrict__ part_dd)
{
    __shared__ double sh[4 * 16 * 8];
    const BatchSys s = sys[blockIdx.x];
    const double *const parts[4] = {part_rr, part_rz, part_bb, part_dd};
    double tot[4];
    system_totals<4>(s, parts, tot, sh);
    if (threadIdx.x < s.R) {
        BatchScal &q = sc[s.slot * kMaxRhs + threadIdx.x];
        const double trr = tot[0], trz = tot[1], tbb = tot[2], tdd = tot[3];
        const bool broke = q.restart == 2;
        const bool done_by_recurrence = q.state == kCandidate && q.restart == 3;
        // stopped by the FP32 recurrence: its direction missed the last
        // update, so if it goes on it restarts from D r
        const bool stale_dir = q.state == kCandidate && q.restart == kRecurrence32;
        q.restart = 0;
        if ((q.state == kActive || q.state == kCandidate || q.state == kFresh) &&
            !(trr <= 1.0e300 && tbb <= 1.0e300)) {
            // a non-finite residual (overflow, or non-finite data): this
            // right-hand side stops, unconverged, and the others go on
            q.true_rr = trr;
            q.rr = trr;
            if (q.state == kFresh) q.bb = tbb;
            q.state = kStopped;
            return;
        }
        if (q.state == kFresh) {
            // compute_solver_tolerance (:1063-1073): tol2 holds tol^2 until
            // here; relative means tol ||b - A x0||
            q.bb = tbb;
            q.init_rr = trr;
            if (!(q.flags & kAbsTol)) q.tol2 = q.tol2 * trr;
            q.iters = 0;
            q.state = trr <= q.tol2 ? kConverged : kActive;
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
                q.restart = broke || stale_dir || tdd > 1.0e-4 * trr;
                q.state = kActive;
                q.rz = trz;
            }
        }
    }
}

// Jacobi diagonal in FP64 and FP32 (|a_ii| > 1e-14, else 1: advanced_solvers.f90:1124-1155);
// use = 0 gives the identity (unpreconditioned CG)
__global__ void k_b_jacobi(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                           const double *__restrict__ val, int use, double *__restrict__ dinv,
                           float *__restrict__ dinv32)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
        if (idx[k] == i) {
            d = val[k];
            break;
        }
    const double di = use && fabs(d) > 1.0e-14 ? 1.0 / d : 1.0;
    dinv[i] = di;
    dinv32[i] = (float)di;
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

// ----------------------------------------------------------------- engine --
// Dynamic LDS of a system's product: for the tiled form its largest column
// tile, R values of its iteration precision per column; for the row-split
// form the four waves' 64 x R products p.Ap in FP64.
constexpr size_t kTileLdsCap = 65536;
size_t tile_lds_bytes(const BatchSys &b)
{
    if (b.split) return (size_t)4 * 64 * b.R * sizeof(double);
    return (size_t)b.tiled * (size_t)b.R * (b.prec ? sizeof(double) : sizeof(float));
}

// One group of the batch: its systems, work table and per-item partial sums,
// advanced on a stream of its own.
struct BatchGroup {
    DBuf<BatchSys> d_sys;
    DBuf<int2> d_work;
    DBuf<double> d_p0, d_p1, d_p2, d_p3;
    std::vector<BatchSys> h_sys; // what d_sys / d_work hold
    std::vector<int2> h_work;
    int nwork = 0, nsys = 0;
    size_t lds = 0;         // dynamic LDS of k_b_spmv: the largest of its systems' needs
    int64_t generation = 0; // bumped whenever the tables change
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;

    void set(hipStream_t s, const std::vector<BatchSys> &sys, const std::vector<int2> &work)
    {
        if (d_sys.p && sys.size() == h_sys.size() && work.size() == h_work.size() &&
            std::memcmp(sys.data(), h_sys.data(), sizeof(BatchSys) * sys.size()) == 0 &&
            std::memcmp(work.data(), h_work.data(), sizeof(int2) * work.size()) == 0)
            return;
        h_sys = sys;
        h_work = work;
        ++generation;
        d_sys.upload(sys, s);
        d_work.upload(work, s);
        const size_t np = std::max<size_t>(1, work.size() * kMaxRhs);
        if (d_p0.n < np) {
            d_p0.alloc(np);
            d_p1.alloc(np);
            d_p2.alloc(np);
            d_p3.alloc(np);
        }
        nwork = (int)work.size();
        nsys = (int)sys.size();
        lds = 0;
        for (const BatchSys &b : sys) lds = std::max(lds, tile_lds_bytes(b));
    }

    // enqueue one PCG iteration of the group's systems
    void iteration(BatchScal *scal) const
    {
        const int W = nwork, NS = nsys;
        hipStream_t s = stream;
        hipLaunchKernelGGL(k_b_spmv, dim3(W), dim3(256), lds, s, d_work.p, d_sys.p, scal, d_p0.p);
        hipLaunchKernelGGL(k_b_alpha, dim3(NS), dim3(kTotThreads), 0, s, d_sys.p, scal, d_p0.p);
        hipLaunchKernelGGL(k_b_update, dim3(W), dim3(256), 0, s, d_work.p, d_sys.p, scal, d_p0.p,
                           d_p1.p);
        hipLaunchKernelGGL(k_b_beta, dim3(NS), dim3(kTotThreads), 0, s, d_sys.p, scal, d_p0.p,
                           d_p1.p);
        hipLaunchKernelGGL(k_b_dir, dim3(W), dim3(256), 0, s, d_work.p, d_sys.p, scal, 0);
    }

    // enqueue the FP64 residual refresh (which also starts fresh systems)
    void refresh(BatchScal *scal) const
    {
        const int W = nwork, NS = nsys;
        hipStream_t s = stream;
        hipLaunchKernelGGL(k_b_fold, dim3(W), dim3(256), 0, s, d_work.p, d_sys.p);
        hipLaunchKernelGGL(k_b_resid, dim3(W), dim3(256), 0, s, d_work.p, d_sys.p, scal, d_p0.p,
                           d_p1.p, d_p2.p, d_p3.p);
        hipLaunchKernelGGL(k_b_refresh, dim3(NS), dim3(kTotThreads), 0, s, d_sys.p, scal, d_p0.p, d_p1.p,
                           d_p2.p, d_p3.p);
        hipLaunchKernelGGL(k_b_dir, dim3(W), dim3(256), 0, s, d_work.p, d_sys.p, scal, 1);
    }
};

// Device state of the server's batch: the per-(slot, right-hand side)
// scalars with their host mirror, and kGroups groups of systems.  Each
// iteration is a chain of five launches, most of them short and latency
// bound; the groups run their chains on separate streams (default priority)
// so that one group's small launches overlap another's products.  A
// system's arithmetic does not depend on its group.
struct BatchEngine {
    static constexpr int kGroups = 2;
    BatchGroup grp[kGroups];
    DBuf<BatchScal> d_scal;
    std::vector<BatchScal> h_scal;
    hipEvent_t start = nullptr;
    int nwork = 0;
    hipGraphExec_t graph = nullptr;
    int64_t graph_key = -1;

    ~BatchEngine()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (BatchGroup &g : grp) {
            if (g.done) (void)hipEventDestroy(g.done);
            if (g.stream) (void)hipStreamDestroy(g.stream);
        }
        if (start) (void)hipEventDestroy(start);
    }

    void init(int slots)
    {
        h_scal.assign((size_t)slots * kMaxRhs, BatchScal{});
        d_scal.alloc(h_scal.size());
        for (BatchGroup &g : grp) {
            FFH_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
            FFH_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
        }
        FFH_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
    }

    // append a system's work items (kRowsPerItem rows each) and the system
    static void append(BatchSys b, std::vector<BatchSys> &sys, std::vector<int2> &work)
    {
        b.w0 = (int)work.size();
        for (int row = 0; row < b.n; row += kRowsPerItem)
            work.push_back(make_int2((int)sys.size(), row));
        b.w1 = (int)work.size();
        sys.push_back(b);
    }

    void set(Context *ctx, const std::vector<BatchSys> (&sys)[kGroups],
             const std::vector<int2> (&work)[kGroups])
    {
        nwork = 0;
        for (int g = 0; g < kGroups; ++g) {
            grp[g].set(ctx->stream, sys[g], work[g]);
            nwork += grp[g].nwork;
        }
        FFH_HIP(hipStreamSynchronize(ctx->stream));
    }

    // `chunk` PCG iterations of every system with an active right-hand side,
    // then the FP64 residual refresh; ordered after and before the work on
    // the context's stream
    // one chunk (iterations + refresh) of every group, forked from and
    // joined back to the context stream
    void enqueue(Context *ctx, int chunk)
    {
        FFH_HIP(hipEventRecord(start, ctx->stream));
        for (BatchGroup &g : grp)
            if (g.nwork) FFH_HIP(hipStreamWaitEvent(g.stream, start, 0));
        for (int it = 0; it < chunk; ++it)
            for (const BatchGroup &g : grp)
                if (g.nwork) g.iteration(d_scal.p);
        for (const BatchGroup &g : grp)
            if (g.nwork) g.refresh(d_scal.p);
        FFH_HIP(hipGetLastError());
        for (BatchGroup &g : grp)
            if (g.nwork) {
                FFH_HIP(hipEventRecord(g.done, g.stream));
                FFH_HIP(hipStreamWaitEvent(ctx->stream, g.done, 0));
            }
    }

    // The chunk, launched directly or replayed from a HIP graph captured on
    // the context stream (re-captured when the batch's tables or the chunk
    // length change: the graph holds their grid sizes and addresses).
    void run(Context *ctx, int chunk, bool use_graph)
    {
        if (nwork == 0) return;
        if (!use_graph) {
            enqueue(ctx, chunk);
            return;
        }
        int64_t key = chunk;
        for (const BatchGroup &g : grp) key = key * 1000003 + g.generation * 31 + g.nwork;
        if (!graph || graph_key != key) {
            if (graph) {
                (void)hipGraphExecDestroy(graph);
                graph = nullptr;
            }
            hipGraph_t gr;
            FFH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            enqueue(ctx, chunk);
            FFH_HIP(hipStreamEndCapture(ctx->stream, &gr));
            const hipError_t e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            FFH_HIP(e);
            graph_key = key;
        }
        FFH_HIP(hipGraphLaunch(graph, ctx->stream));
    }

    void pull(Context *ctx)
    {
        FFH_HIP(hipMemcpyAsync(h_scal.data(), d_scal.p, sizeof(BatchScal) * h_scal.size(),
                               hipMemcpyDeviceToHost, ctx->stream));
        FFH_HIP(hipStreamSynchronize(ctx->stream));
    }

    void push(Context *ctx, int slot)
    {
        FFH_HIP(hipMemcpyAsync(d_scal.p + (size_t)slot * kMaxRhs, h_scal.data() + (size_t)slot * kMaxRhs,
                               sizeof(BatchScal) * kMaxRhs, hipMemcpyHostToDevice, ctx->stream));
    }
};

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
    DBuf<float> val32, dinv32, r, p, ap, dx; // SELL-64 values, FP32 iteration
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
    bool use_graph = false; // replay each chunk from a HIP graph
    int64_t switches = 0;
    std::map<ffh::ServerKey, std::unique_ptr<ffh::Problem>> cache;
    std::map<int64_t, std::unique_ptr<ffh::ServedRequest>> reqs;
    std::deque<ffh::ServedRequest *> queue;
    std::vector<ffh::ServedRequest *> active;
    std::vector<int> free_slots;
    bool changed = false;
    int64_t next_ticket = 1;
    ffh::BatchEngine eng; // device batch state
    int64_t iterations_run = 0;
    // autotuned product per system type (pattern, R, iteration precision):
    // the largest tile when the tiled product won, else 0
    // product form per system type: x = tile columns of the tiled form (or
    // 0), y = lanes per row of the row-split form (or 0)
    std::map<std::tuple<const ffh::Problem *, int, int>, int2> tiling;
    // -1 autotune, 0 plain, 1 tiled where the tiles fit, 2 / 4 / 8 row-split
    // with that many lanes per row where R allows it
    int tiling_mode = -1;
};

namespace ffh {

// worst-case device bytes of a request: both SELL value copies, both
// diagonals, b and x, and the FP32 and FP64 iteration vectors (an FP32
// request may switch to FP64)
static size_t request_bytes(int n, int64_t sell_total, int R)
{
    return (size_t)sell_total * 12 + (size_t)n * 12 + (size_t)n * R * (8 + 8 + 16 + 24) + 4096;
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

static int rhs_stride(int nrhs)
{
    return nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : nrhs <= 6 ? 6 : 8;
}

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
    FFH_REQUIRE(sell_structure(S->ctx, P->reduced.m), FFH_ERR_INVALID,
                "the reduced matrix has no SELL-64 form (rows too uneven)");
    // column tiles for the batched product's autotuner (host work, once)
    ensure_sell_tiles(S->ctx, P->reduced.m);
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
    const size_t tot = (size_t)std::max<int64_t>(A.sell_total, 1);
    q->val64.alloc(tot);
    q->val32.alloc(tot);
    sell_values(ctx, A, A.val.p, q->val64.p, q->val32.p);
    q->dinv.alloc(n);
    q->dinv32.alloc(n);
    hipLaunchKernelGGL(k_b_jacobi, dim3((n + 255) / 256), dim3(256), 0, s, n, A.col_ptr.p,
                       A.row_idx.p, A.val.p, 1, q->dinv.p, q->dinv32.p);
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
        q->dx.alloc(nr);
        FFH_HIP(hipMemsetAsync(q->r.p, 0, sizeof(float) * nr, s));
        FFH_HIP(hipMemsetAsync(q->p.p, 0, sizeof(float) * nr, s));
        FFH_HIP(hipMemsetAsync(q->dx.p, 0, sizeof(float) * nr, s));
    }
    q->best_rr.assign(q->nrhs, 1e300);
    q->stall.assign(q->nrhs, 0);
    FFH_HIP(hipStreamSynchronize(s)); // tmp and the cached problem's values are reused next
    q->slot = S->free_slots.back();
    S->free_slots.pop_back();
    for (int r = 0; r < kMaxRhs; ++r) {
        BatchScal &c = S->eng.h_scal[(size_t)q->slot * kMaxRhs + r];
        c = BatchScal{};
        c.state = r < q->nrhs ? kFresh : kUnused;
        c.max_iter = q->max_iter;
        c.tol2 = q->tol * q->tol;
    }
    S->eng.push(ctx, q->slot);
    S->used += q->bytes;
    S->active.push_back(q);
    S->changed = true;
}

// A request's system with its device arrays (slot and tiling set by the
// caller).
static BatchSys batch_sys(const ServedRequest *q)
{
    const DevMatrix &A = q->sym->reduced.m;
    BatchSys b{};
    b.off = A.sell_off.p;
    b.sidx = A.sell_idx.p;
    b.srow = A.sell_row.p;
    b.tpack = A.tile_pack.p;
    b.tlidx = A.tile_lidx.p;
    b.nslice = A.nslice;
    b.val32 = q->val32.p;
    b.val64 = q->val64.p;
    b.dinv32 = q->dinv32.p;
    b.dinv64 = q->dinv.p;
    b.b = q->b.p;
    b.x = q->x.p;
    b.r = q->r.p;
    b.p = q->p.p;
    b.ap = q->ap.p;
    b.dx = q->dx.p;
    b.r64 = q->r64.p;
    b.p64 = q->p64.p;
    b.ap64 = q->ap64.p;
    b.prec = q->prec;
    b.n = q->n;
    b.R = q->R;
    b.slot = q->slot;
    b.x_zero = 1;
    return b;
}

// Autotuner of the batched product: for each system type (pattern, R,
// iteration precision), the first time it is batched, the plain SELL-64
// product, the one with LDS column tiles (when the pattern's tiles fit
// kTileLdsCap at R values per column) and the row-split forms that R allows
// are timed on that system alone (HIP events, 2 + 5 launches each, its p as
// it stands); the fastest is kept when it beats the plain form by 2 %.  All
// forms give the same bits, so the choice changes speed only.
static int2 tuned_form(ffh_server *S, const ServedRequest *q)
{
    Context *ctx = S->ctx;
    DevMatrix &A = q->sym->reduced.m;
    const auto key = std::make_tuple((const Problem *)q->sym, q->R, q->prec);
    auto it = S->tiling.find(key);
    if (it != S->tiling.end()) return it->second;
    ensure_sell_tiles(ctx, A);
    const size_t sv = q->prec ? sizeof(double) : sizeof(float);
    const bool fits = A.tiles_ok && A.tile_max > 0 && (size_t)A.tile_max * q->R * sv <= kTileLdsCap;
    std::vector<int2> cand{make_int2(0, 0)};
    if (fits) cand.push_back(make_int2(A.tile_max, 0));
    for (int L = 2; L <= 8; L *= 2)
        if (split_ok(q->R, L)) cand.push_back(make_int2(0, L));
    int2 pick = make_int2(0, 0);
    const int mode = S->tiling_mode;
    if (mode >= 0) {
        for (const int2 &c : cand)
            if ((mode == 1 && c.x > 0) || (mode >= 2 && c.y == mode)) pick = c;
    } else if (cand.size() > 1) {
        hipStream_t s = ctx->stream;
        BatchSys b = batch_sys(q);
        b.slot = 0;
        std::vector<BatchSys> sys;
        std::vector<int2> work;
        BatchEngine::append(b, sys, work);
        std::vector<BatchScal> sc(kMaxRhs, BatchScal{});
        for (int r = 0; r < kMaxRhs; ++r) sc[r].state = r < q->R ? kActive : kUnused;
        DBuf<BatchSys> d_sys;
        DBuf<int2> d_work;
        DBuf<BatchScal> d_sc;
        DBuf<double> d_part;
        d_work.upload(work, s);
        d_sc.upload(sc, s);
        d_part.alloc(work.size() * kMaxRhs);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        FFH_HIP(hipEventCreate(&e0));
        std::vector<float> ms(cand.size(), 0.0f);
        try {
            FFH_HIP(hipEventCreate(&e1));
            for (size_t v = 0; v < cand.size(); ++v) {
                sys[0].tiled = cand[v].x;
                sys[0].split = cand[v].y;
                d_sys.upload(sys, s);
                const size_t lds = tile_lds_bytes(sys[0]);
                for (int k = 0; k < 7; ++k) {
                    if (k == 2) FFH_HIP(hipEventRecord(e0, s));
                    hipLaunchKernelGGL(k_b_spmv, dim3((unsigned)work.size()), dim3(256), lds, s,
                                       d_work.p, d_sys.p, d_sc.p, d_part.p);
                }
                FFH_HIP(hipGetLastError());
                FFH_HIP(hipEventRecord(e1, s));
                FFH_HIP(hipEventSynchronize(e1));
                FFH_HIP(hipEventElapsedTime(&ms[v], e0, e1));
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        size_t best = 0;
        for (size_t v = 1; v < cand.size(); ++v)
            if (ms[v] < ms[best]) best = v;
        if (ms[best] < 0.98f * ms[0]) pick = cand[best];
    }
    S->tiling[key] = pick;
    return pick;
}

// Systems go to the group with the least work so far (largest first), the
// work of a system being its SELL entries times (1 + R / 4).
static void rebuild_batch(ffh_server *S)
{
    constexpr int G = BatchEngine::kGroups;
    std::vector<BatchSys> sys[G];
    std::vector<int2> work[G];
    double load[G] = {};
    std::vector<ServedRequest *> order(S->active);
    auto cost = [](const ServedRequest *q) {
        return (double)q->sym->reduced.m.sell_total * (1.0 + q->R / 4.0);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](const ServedRequest *a, const ServedRequest *b) { return cost(a) > cost(b); });
    for (ServedRequest *q : order) {
        BatchSys b = batch_sys(q);
        const int2 form = tuned_form(S, q);
        b.tiled = form.x;
        b.split = form.y;
        const int g = (int)(std::min_element(load, load + G) - load);
        load[g] += cost(q);
        BatchEngine::append(b, sys[g], work[g]);
    }
    S->eng.set(S->ctx, sys, work);
    S->changed = false;
}

// One step of the batch: `chunk` PCG iterations of every active system, then
// the FP64 residual refresh (which also starts systems admitted since the
// last step: they sit out the iterations of their first step).
static void run_chunk(ffh_server *S)
{
    if (S->eng.nwork == 0 || S->active.empty()) return;
    S->eng.run(S->ctx, S->chunk, S->use_graph);
    S->iterations_run += S->chunk;
}

// An FP32-iterating request whose FP64 residual stalls (no refresh in a
// window of kStallIterations iterations brings it below kStallImprove of its
// best: the FP32 operator cannot resolve further progress) continues in FP64
// from its current x: its directions restart after the next refresh.  The
// window is counted in iterations, so that a longer chunk does not keep a
// stalled request in FP32 longer (at least 4 refreshes; 32 at a chunk of 16,
// the single solve's window of kStallRefreshes refreshes).
constexpr int kStallIterations = 512;
static void watch_stagnation(ffh_server *S, ServedRequest *q)
{
    const int window = std::max(4, (kStallIterations + S->chunk - 1) / S->chunk);
    bool stuck = false;
    for (int r = 0; r < q->nrhs; ++r) {
        const BatchScal &c = S->eng.h_scal[(size_t)q->slot * kMaxRhs + r];
        if (c.state != kActive) continue;
        if (c.true_rr < kStallImprove * q->best_rr[r]) {
            q->best_rr[r] = c.true_rr;
            q->stall[r] = 0;
        } else if (++q->stall[r] >= window) {
            stuck = true;
        }
    }
    if (!stuck) return;
    hipStream_t s = S->ctx->stream;
    alloc_fp64_vectors(q, s);
    q->prec = 1;
    for (int r = 0; r < q->nrhs; ++r) {
        BatchScal &c = S->eng.h_scal[(size_t)q->slot * kMaxRhs + r];
        if (c.state == kActive || c.state == kCandidate) {
            c.state = kCandidate;
            c.restart = 2;
        }
    }
    S->eng.push(S->ctx, q->slot);
    FFH_HIP(hipStreamSynchronize(s));
    q->r.reset(), q->p.reset(), q->ap.reset(), q->dx.reset();
    S->changed = true;
    ++S->switches;
}

static void retire_finished(ffh_server *S, int *finished)
{
    hipStream_t s = S->ctx->stream;
    S->eng.pull(S->ctx);
    std::vector<ServedRequest *> keep;
    for (ServedRequest *q : S->active) {
        bool all = true;
        for (int r = 0; r < q->nrhs; ++r) {
            const int st = S->eng.h_scal[(size_t)q->slot * kMaxRhs + r].state;
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
            const BatchScal &c = S->eng.h_scal[(size_t)q->slot * kMaxRhs + r];
            q->iters[r] = c.iters;
            q->converged[r] = c.state == kConverged;
            q->resid[r] = c.bb > 0 ? std::sqrt(c.true_rr / c.bb) : 0.0;
        }
        q->val32.reset(), q->dinv32.reset(), q->r.reset(), q->p.reset(), q->ap.reset();
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
    FFH_REQUIRE(std::isfinite(mass) && std::isfinite(tol), FFH_ERR_INVALID,
                "non-finite mass coefficient or tolerance");
    for (int c = 0; c < P->ncell; ++c)
        FFH_REQUIRE(std::isfinite(coeff[c]), FFH_ERR_INVALID,
                    "non-finite cell coefficient (cell " + std::to_string(c) + ")");
    const int nfree = P->reduced.m.n;
    for (int r = 0; r < nrhs; ++r)
        for (int i = 0; i < nfree; ++i)
            FFH_REQUIRE(std::isfinite(rhs[(size_t)r * nfree + i]), FFH_ERR_INVALID,
                        "non-finite right-hand side " + std::to_string(r) + " (row " +
                            std::to_string(i) + ")");
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
    q->bytes = request_bytes(q->n, P->reduced.m.sell_total, q->R);
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
        S->eng.init(max_systems);
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

int ffh_server_wait(ffh_server *S, int64_t ticket, int max_steps, double *solution,
                    int *iterations, int *converged, double *rel_residual)
{
    if (!S) return FFH_ERR_HANDLE;
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return FFH_ERR_HANDLE;
    if (max_steps < 0) return FFH_ERR_INVALID;
    const ffh::ServedRequest *q = it->second.get();
    for (int k = 0; k < max_steps && !q->done; ++k) {
        const int code = ffh_server_step(S, nullptr, nullptr, nullptr);
        if (code != FFH_OK) return code;
    }
    return ffh_server_poll(S, ticket, solution, iterations, converged, rel_residual);
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
            const ffh::BatchScal &c = S->eng.h_scal[(size_t)q->slot * ffh::kMaxRhs + r];
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

int ffh_server_use_graph(ffh_server *S, int enable)
{
    if (!S) return FFH_ERR_HANDLE;
    S->use_graph = enable != 0;
    return FFH_OK;
}

int ffh_server_set_tiling(ffh_server *S, int mode)
{
    if (!S) return FFH_ERR_HANDLE;
    if (!(mode >= -1 && mode <= 2) && mode != 4 && mode != 8) return FFH_ERR_INVALID;
    S->tiling_mode = mode;
    S->tiling.clear();
    S->changed = true;
    return FFH_OK;
}

int ffh_server_tiling_stats(const ffh_server *S, int *types, int *tiled)
{
    if (!S) return FFH_ERR_HANDLE;
    return ffh_server_product_forms(S, types, tiled, nullptr);
}

int ffh_server_product_forms(const ffh_server *S, int *types, int *tiled, int *split)
{
    if (!S) return FFH_ERR_HANDLE;
    int n = 0, t = 0, l = 0;
    for (const auto &kv : S->tiling) {
        ++n;
        t += kv.second.x > 0;
        l += kv.second.y > 0;
    }
    if (types) *types = n;
    if (tiled) *tiled = t;
    if (split) *split = l;
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
