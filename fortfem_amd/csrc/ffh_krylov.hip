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
#include <mutex>

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
// Thread t adds partials t, t + BS, t + 2 BS, ... in that order; the loads of
// kSumLoads of them are issued together, so a sum over a few thousand
// partials costs one or two load latencies, not one per partial.
constexpr int kSumLoads = 4;
template <int BS>
__device__ __forceinline__ double sum_partials(const double *__restrict__ partials, int nb)
{
    __shared__ double tot;
    double t = 0.0;
    for (int i0 = threadIdx.x; i0 < nb; i0 += kSumLoads * BS) {
        double v[kSumLoads];
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k) {
            const int i = i0 + k * BS;
            v[k] = i < nb ? partials[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k)
            if (i0 + k * BS < nb) t += v[k];
    }
    t = block_sum<BS>(t);
    if (threadIdx.x == 0) tot = t;
    __syncthreads();
    const double r = tot;
    __syncthreads();
    return r;
}

// Totals of two partial arrays (same order as sum_partials for each), with
// the loads of both issued together.
template <int BS>
__device__ __forceinline__ void sum_partials2(const double *__restrict__ pa,
                                              const double *__restrict__ pb, int nb, double &ta,
                                              double &tb)
{
    __shared__ double tot[2];
    double a = 0.0, b = 0.0;
    for (int i0 = threadIdx.x; i0 < nb; i0 += kSumLoads * BS) {
        double va[kSumLoads], vb[kSumLoads];
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k) {
            const int i = i0 + k * BS;
            va[k] = i < nb ? pa[i] : 0.0;
            vb[k] = i < nb ? pb[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k)
            if (i0 + k * BS < nb) {
                a += va[k];
                b += vb[k];
            }
    }
    a = block_sum<BS>(a);
    b = block_sum<BS>(b);
    if (threadIdx.x == 0) {
        tot[0] = a;
        tot[1] = b;
    }
    __syncthreads();
    ta = tot[0];
    tb = tot[1];
    __syncthreads();
}

__device__ __forceinline__ bool is_breakdown(double pap, double rz, int mode)
{
    if (mode == 1) return !(fabs(pap) > 1.0e-14 * fabs(rz));
    return fabs(pap) < 1.0e-14;
}

// ------------------------------------------------------------- gathers --
// What a row engine multiplies the matrix entries by, in two steps so that
// the loads can be issued before the scalars they are combined with are
// known: raw(c) loads column c's data, val(raw) forms the value.  XPlain is a
// plain vector.  XDir forms the new CG direction on the fly,
// p_new = z + beta p_old with z = D r (p_new = z at a restart), so that the
// direction update needs no launch of its own; z and p_old are interleaved
// (one load per column).  The owner of row i stores p_new[i] from the same
// expression, so a gathered value and the stored one are the same number.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }

template <typename VT>
struct XPlain {
    using Raw = VT;
    const VT *__restrict__ x;
    __device__ __forceinline__ Raw raw(int c) const { return x[c]; }
    __device__ __forceinline__ VT val(Raw v) const { return v; }
    __device__ __forceinline__ VT operator()(int c) const { return x[c]; }
};

template <typename VT>
struct alignas(2 * sizeof(VT)) ZP {
    VT z, p;
};

template <typename VT>
struct XDir {
    using Raw = ZP<VT>;
    const ZP<VT> *__restrict__ zp;
    VT beta;
    int restart;
    __device__ __forceinline__ Raw raw(int c) const { return zp[c]; }
    __device__ __forceinline__ VT val(Raw v) const
    {
        return restart ? v.z : fma_t(beta, v.p, v.z);
    }
    __device__ __forceinline__ VT operator()(int c) const { return val(zp[c]); }
};

// The start hook of an engine: called once by every thread of the block,
// after the block's first loads are issued and before any value is formed;
// false stops the block.  NoHook for plain products.
struct NoHook {
    __device__ __forceinline__ bool operator()() const { return true; }
};

// start hook that ends the launch once the solve has stopped: the control
// block is read after the block's first matrix loads are issued, so its
// latency overlaps them (*stop records that the launch ended here)
struct DoneHook {
    const KrylovCtl *ctl;
    bool *stop;
    __device__ __forceinline__ bool operator()() const
    {
        if (ctl->done) {
            *stop = true;
            return false;
        }
        return true;
    }
};

// ------------------------------------------------------------ CSR stream --
// A row block is a run of consecutive rows whose nonzeros fit in LDS.  The
// block streams the run's values and column indices with fully coalesced
// loads (each lane takes every BS-th nonzero), multiplies by the gathered
// values into LDS, then one thread per row adds its row's products in column
// order.  A row longer than the LDS stage is summed by the whole block.
// Epilogue E receives (row, row value) and accumulates up to three partial
// sums in d.
template <typename MT, typename VT, typename XF, typename E, typename H>
__device__ __forceinline__ void csr_stream(const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val, const XF &xf,
                                           const int *__restrict__ rowblocks, int nrb,
                                           const E &epi, double *d, const H &hook)
{
    constexpr int BS = kStreamBlock, NZ = kStreamNnz, U = NZ / BS;
    __shared__ VT prod[NZ];
    bool hooked = false;
    for (int blk = xcd_block(); blk < nrb; blk += gridDim.x) {
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
            // the first index/value loads are in flight during the hook
            // (holding the gathered values across it as well costs more
            // registers than it saves)
            if (!hooked) {
                if (!hook()) return;
                hooked = true;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = threadIdx.x + u * BS;
                if (k < cnt) prod[k] = (VT)v[u] * xf(c[u]);
            }
            __syncthreads();
            const int nrows = r1 - r0;
            if (2 * nrows > BS) {
                for (int r = r0 + threadIdx.x; r < r1; r += BS) {
                    const int b = ptr[r] - k0, e = ptr[r + 1] - k0;
                    VT s = VT(0);
                    for (int k = b; k < e; ++k) s += prod[k];
                    epi(r, s, d);
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
                if (lane == 0 && r < r1) epi(r, s, d);
            }
            __syncthreads();
        } else {
            // one long row (r1 == r0 + 1)
            if (!hooked) {
                if (!hook()) return;
                hooked = true;
            }
            VT s = VT(0);
            for (int k = k0 + threadIdx.x; k < k1; k += BS) s += (VT)val[k] * xf(idx[k]);
            double t = block_sum<BS>((double)s);
            __shared__ double tot;
            if (threadIdx.x == 0) tot = t;
            __syncthreads();
            if (threadIdx.x == 0) epi(r0, (VT)tot, d);
            __syncthreads();
        }
    }
    if (!hooked) (void)hook();
}

// Vector CSR: a group of G lanes per row (grid-stride over rows); each lane
// issues two independent (index, value) loads before its gathers, then the
// group reduces by shuffles.  Used when it beats the stream kernel for a
// matrix (chosen by the autotuner below).
template <int G, typename MT, typename VT, typename XF, typename E, typename H>
__device__ __forceinline__ void csr_vector(int n, const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val, const XF &xf,
                                           const E &epi, double *d, const H &hook)
{
    if (!hook()) return;
    constexpr int RPB = kStreamBlock / G;
    const int lane = threadIdx.x % G;
    const int ngrp = gridDim.x * RPB;
    for (int row = xcd_block() * RPB + threadIdx.x / G; row < n; row += ngrp) {
        const int b = ptr[row], e = ptr[row + 1];
        VT s = VT(0);
        int k = b + lane;
        for (; k + G < e; k += 2 * G) {
            const int c0 = __builtin_nontemporal_load(idx + k);
            const int c1 = __builtin_nontemporal_load(idx + k + G);
            const MT v0 = __builtin_nontemporal_load(val + k);
            const MT v1 = __builtin_nontemporal_load(val + k + G);
            s += (VT)v0 * xf(c0);
            s += (VT)v1 * xf(c1);
        }
        if (k < e) s += (VT)val[k] * xf(idx[k]);
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
        if (lane == 0) epi(row, s, d);
    }
}

// SELL-64: one wave per slice of 64 rows, one lane per row.  Every index
// and value load of a wave reads 64 consecutive words; each lane adds its
// row's products in column order (the CSR order).  Entries go in steps of
// eight: the sixteen index/value loads of a step and its eight gathers are
// issued together, so a wave has many loads in flight.  Here ptr is the
// slice offset array, idx/val the SELL entries, srow the row of each slice
// position and nrb the slice count.
template <typename MT, typename VT, typename XF, typename E, typename H>
__device__ __forceinline__ void sell_rows(int n, const int *__restrict__ off,
                                          const int *__restrict__ sidx,
                                          const MT *__restrict__ sval,
                                          const int *__restrict__ srow, const XF &xf,
                                          int nslice, const E &epi, double *d, const H &hook)
{
    constexpr int WPB = kStreamBlock / 64, U = 8;
    const int lane = threadIdx.x & 63;
    bool hooked = false;
    // all waves of the block call the hook at the same point (it may
    // synchronise the block): the loop runs the same trip count in every
    // wave of the block, idle waves included
    const int first = xcd_block() * WPB;
    for (int base = first; base < nslice || !hooked; base += gridDim.x * WPB) {
        const int sl = base + (int)(threadIdx.x >> 6);
        const bool live = sl < nslice;
        const int o = live ? off[sl] : 0;
        const int w = live ? (off[sl + 1] - o) >> 6 : 0;
        const int *ip = sidx + o + lane;
        const MT *vp = sval + o + lane;
        VT s = VT(0);
        int j = 0;
        for (; j < w || !hooked; j += U) {
            int c[U];
            MT v[U];
            typename XF::Raw q[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) {
                    c[u] = __builtin_nontemporal_load(ip + 64 * u);
                    v[u] = __builtin_nontemporal_load(vp + 64 * u);
                }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) q[u] = xf.raw(c[u]);
            if (!hooked) {
                if (!hook()) return;
                hooked = true;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < w) s += (VT)v[u] * xf.val(q[u]);
            ip += 64 * U;
            vp += 64 * U;
        }
        if (live) {
            const int row = srow[sl * 64 + lane];
            if (row >= 0) epi(row, s, d);
        }
    }
}

// SELL-64 with column tiles: a block takes a group of four slices (256
// rows); the group's distinct columns (its tile, sorted) are loaded once,
// their values formed (for the direction, with the beta the start hook
// finds) and staged in LDS, then each wave adds its slice's rows from LDS
// through 16-bit tile positions.  The matrix stream is the value plus a
// 2-byte position per entry, and the vector leaves L2 once per group.  Here
// ptr is the slice offset array, idx the 16-bit positions (as int words),
// val the SELL values, and pack holds the row of each slice position
// (nslice * 64), then the tile offsets (ngroup + 1), then the tile columns.
template <typename MT, typename VT, typename XF, typename E, typename H>
__device__ __forceinline__ void sell_tiled(const int *__restrict__ off,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ sval,
                                           const int *__restrict__ pack, const XF &xf,
                                           int nslice, const E &epi, double *d, const H &hook)
{
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    VT *lds = reinterpret_cast<VT *>(dyn_lds);
    const uint16_t *lidx = reinterpret_cast<const uint16_t *>(idx);
    const int ngroup = (nslice + 3) >> 2;
    const int *srow = pack;
    const int *toff = pack + (size_t)nslice * 64;
    const int *tcols = toff + ngroup + 1;
    const int lane = threadIdx.x & 63;
    bool hooked = false;
    // the tile is staged SB columns per thread at a time: the batch's
    // column-list loads, then its vector gathers, then its LDS stores, so a
    // thread has SB gathers in flight instead of one
    constexpr int SB = 2;
    for (int g = xcd_block(); g < ngroup || !hooked; g += gridDim.x) {
        const bool live = g < ngroup;
        const int c0 = live ? toff[g] : 0;
        const int nc = live ? toff[g + 1] - c0 : 0;
        int cc[SB];
        typename XF::Raw qq[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int k = (int)threadIdx.x + u * kStreamBlock;
            cc[u] = k < nc ? tcols[c0 + k] : 0;
        }
#pragma unroll
        for (int u = 0; u < SB; ++u)
            if ((int)threadIdx.x + u * kStreamBlock < nc) qq[u] = xf.raw(cc[u]);
        // the first batch's loads are in flight during the hook
        if (!hooked) {
            if (!hook()) return;
            hooked = true;
        }
        if (!live) break;
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int k = (int)threadIdx.x + u * kStreamBlock;
            if (k < nc) lds[k] = xf.val(qq[u]);
        }
        for (int k0 = (int)threadIdx.x + SB * kStreamBlock; k0 < nc; k0 += SB * kStreamBlock) {
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int k = k0 + u * kStreamBlock;
                cc[u] = k < nc ? tcols[c0 + k] : 0;
            }
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (k0 + u * kStreamBlock < nc) qq[u] = xf.raw(cc[u]);
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int k = k0 + u * kStreamBlock;
                if (k < nc) lds[k] = xf.val(qq[u]);
            }
        }
        __syncthreads();
        const int sl = 4 * g + (int)(threadIdx.x >> 6);
        if (sl < nslice) {
            const int o = off[sl], w = (off[sl + 1] - o) >> 6;
            const uint16_t *ip = lidx + o + lane;
            const MT *vp = sval + o + lane;
            VT s = VT(0);
            constexpr int U = 8;
            int j = 0;
            for (; j + U <= w; j += U) {
                int c[U];
                MT v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    c[u] = __builtin_nontemporal_load(ip + 64 * u);
                    v[u] = __builtin_nontemporal_load(vp + 64 * u);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) s += (VT)v[u] * lds[c[u]];
                ip += 64 * U;
                vp += 64 * U;
            }
            for (; j < w; ++j) {
                s += (VT)vp[0] * lds[ip[0]];
                ip += 64;
                vp += 64;
            }
            const int row = srow[sl * 64 + lane];
            if (row >= 0) epi(row, s, d);
        }
        __syncthreads();
    }
}

// Row engine of a launch: V = 0 stream, V = 1 SELL-64 (rowblocks = the row
// of each slice position), V = 2 SELL-64 with column tiles (rowblocks = the
// tile pack), otherwise vector CSR with G = V lanes.
template <int V, typename MT, typename VT, typename XF, typename E, typename H = NoHook>
__device__ __forceinline__ void row_engine(int n, const int *__restrict__ ptr,
                                           const int *__restrict__ idx,
                                           const MT *__restrict__ val, const XF &xf,
                                           const int *__restrict__ rowblocks, int nrb,
                                           const E &epi, double *d, const H &hook = H())
{
    if constexpr (V == 0)
        csr_stream<MT, VT>(ptr, idx, val, xf, rowblocks, nrb, epi, d, hook);
    else if constexpr (V == 1)
        sell_rows<MT, VT>(n, ptr, idx, val, rowblocks, xf, nrb, epi, d, hook);
    else if constexpr (V == 2)
        sell_tiled<MT, VT>(ptr, idx, val, rowblocks, xf, nrb, epi, d, hook);
    else
        csr_vector<V, MT, VT>(n, ptr, idx, val, xf, epi, d, hook);
}

// ------------------------------------------------------------ epilogues --
template <typename VT>
struct EpiStore { // y = A x
    VT *y;
    __device__ void operator()(int r, VT s, double *) const { y[r] = s; }
};

template <typename VT>
struct EpiInit { // r = b - A x ; z = D r ; r.r, r.z
    const VT *b, *dinv;
    VT *r;
    VT *z; // z of row i at z[i * zs]: zs = 2 in (z, p) pairs, 1 dense
    int zs;
    __device__ void operator()(int i, VT s, double *d) const
    {
        const VT ri = b[i] - s;
        const VT zi = dinv[i] * ri;
        r[i] = ri;
        z[(size_t)i * zs] = zi;
        d[0] += (double)ri * (double)ri;
        d[1] += (double)ri * (double)zi;
    }
};

template <typename VT>
struct EpiDirAp { // p_new(i) stored ; Ap = A p_new ; p_new.Ap
    const XDir<VT> *g;
    ZP<VT> *zpnew;
    VT *ap;
    __device__ void operator()(int i, VT s, double *d) const
    {
        const VT pi = (*g)(i);
        zpnew[i] = ZP<VT>{VT(0), pi}; // z is written by the update launch
        ap[i] = s;
        d[0] += (double)pi * (double)s;
    }
};

template <typename VT>
struct EpiApDot { // Ap = A p ; p.Ap
    const VT *p;
    VT *ap;
    __device__ void operator()(int i, VT s, double *d) const
    {
        ap[i] = s;
        d[0] += (double)p[i] * (double)s;
    }
};

struct EpiResidual { // r = b - A x ; r.r
    const double *b;
    double *r;
    __device__ void operator()(int i, double s, double *d) const
    {
        const double ri = b[i] - s;
        r[i] = ri;
        d[0] += ri * ri;
    }
};

struct EpiRefresh { // mixed: FP64 residual replaces the FP32 one ; z ; rr, rz, drift
    const double *b;
    const float *dinv32;
    float *r32;
    float *zo; // z of row i at zo[i * zs] (as EpiInit)
    int zs;
    __device__ void operator()(int i, double s, double *d) const
    {
        const double ri = b[i] - s;
        const double dr = (double)r32[i] - ri;
        const float r = (float)ri;
        const float z = dinv32[i] * r;
        r32[i] = r;
        zo[(size_t)i * zs] = z;
        d[0] += ri * ri;
        d[1] += (double)r * (double)z;
        d[2] += dr * dr;
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
    double d[3] = {0.0, 0.0, 0.0};
    row_engine<V, MT, VT>(n, ptr, idx, val, XPlain<VT>{x}, rowblocks, nrb, epi, d);
}

// Partial-sum slots (stride apart): 0 p.Ap, 1 r.r, 2 r.(D r), 3 drift.
//
// r = b - A x ; z = D r into z (stride zs: the first (z, p) array, or the
// dense z of the three-launch form) ; partials of r.r and r.z
// (initialize_pcg_state :859-878)
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_pcg_init(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx, const MT *__restrict__ val,
    const VT *__restrict__ x, const VT *__restrict__ b, VT *__restrict__ r,
    const VT *__restrict__ dinv, VT *__restrict__ z, int zs, const int *__restrict__ rowblocks,
    int nrb, double *__restrict__ partials, int stride)
{
    EpiInit<VT> epi{b, dinv, r, z, zs};
    double d[3] = {0.0, 0.0, 0.0};
    row_engine<V, MT, VT>(n, ptr, idx, val, XPlain<VT>{x}, rowblocks, nrb, epi, d);
    const double s0 = block_sum<kStreamBlock>(d[0]);
    const double s1 = block_sum<kStreamBlock>(d[1]);
    if (threadIdx.x == 0) {
        partials[stride + blockIdx.x] = s0;
        partials[2 * stride + blockIdx.x] = s1;
    }
}

// initial norm and tolerance (compute_solver_tolerance, :1063-1073); the
// first direction is D r
__global__ __launch_bounds__(kBlock) void k_pcg_init_finish(const double *__restrict__ partials,
                                                            int stride, int nb, KrylovCtl *ctl)
{
    const double trr = sum_partials<kBlock>(partials + stride, nb);
    if (threadIdx.x == 0) {
        const double init = sqrt(trr);
        ctl->rr = trr;
        ctl->initial_norm = init;
        ctl->residual_norm = init;
        ctl->tol = ctl->tol_type == 1 ? ctl->tolerance : ctl->tolerance * init;
        ctl->iter = 0;
        ctl->k2_done = 0;
        ctl->restart = 1;
        ctl->nb_rz = nb;
        ctl->done = ctl->max_iter <= 0 ? 1 : 0;
    }
}

// The tail of an iteration, run by every block of the next launch from the
// r.r / r.(D r) partials of the last update (or initialisation or refresh):
// the stop test on the recursive residual, then beta = rz_new / rz_old
// (pcg_iteration_loop :906-922).  Block 0 records it; the records are the
// same however often it runs for one update.  True when the solve stops.
template <int BS>
__device__ __forceinline__ bool pcg_finish(const double *__restrict__ partials, int stride,
                                           KrylovCtl *ctl, double &beta)
{
    const int nb = ctl->nb_rz;
    double trr, trz;
    sum_partials2<BS>(partials + stride, partials + 2 * stride, nb, trr, trz);
    const int done_it = ctl->k2_done;
    const int par = done_it & 1;
    const bool restart = ctl->restart != 0;
    const double rn = sqrt(trr);
    bool stop = false, conv = false;
    if (!restart) {
        conv = !ctl->mixed && rn <= ctl->tol;
        // FP32 iterations of a mixed solve: the recurrence residual reaching
        // the tolerance ends the chunk at once (the rest of its launches
        // return); the FP64 refresh that follows decides convergence, and
        // clears the stop if it does not agree.  The refresh then comes at
        // this iteration instead of the chunk's end, nothing else changes.
        const bool candidate = ctl->mixed && rn <= ctl->tol;
        stop = conv || candidate || done_it >= ctl->max_iter;
    }
    beta = restart ? 0.0 : trz / ctl->rzs[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->iter = done_it;
        if (!restart) {
            ctl->rr = trr;
            ctl->residual_norm = rn;
        }
        ctl->rzs[par ^ 1] = trz;
        ctl->beta = beta;
        ctl->cur = par ^ 1;
        if (stop) {
            ctl->done = 1;
            if (conv) ctl->converged = 1;
        }
    }
    return stop;
}

// the tail of the previous iteration as a start hook: beta into the gather
template <typename VT>
struct DirHook {
    const double *partials;
    int stride;
    KrylovCtl *ctl;
    XDir<VT> *g;
    __device__ __forceinline__ bool operator()() const
    {
        double beta;
        if (pcg_finish<kStreamBlock>(partials, stride, ctl, beta)) return false;
        g->beta = (VT)beta;
        return true;
    }
};

// Launch 1 of an iteration: the tail of the previous one, then
// p_new = z + beta p_old (own rows stored, neighbours formed on the fly),
// Ap = A p_new and the partials of p_new.Ap.  (z, p) pairs alternate between
// zp0 and zp1: the update's z and the previous direction are read from slot
// k2_done & 1, the new direction is written to the other slot.
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_pcg_dirspmv(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx, const MT *__restrict__ val,
    ZP<VT> *__restrict__ zp0, ZP<VT> *__restrict__ zp1, VT *__restrict__ ap,
    const int *__restrict__ rowblocks, int nrb, double *__restrict__ partials, int stride,
    KrylovCtl *ctl)
{
    if (ctl->done) return;
    const int par = ctl->k2_done & 1;
    XDir<VT> g{par ? zp1 : zp0, VT(0), ctl->restart};
    DirHook<VT> hook{partials, stride, ctl, &g};
    EpiDirAp<VT> epi{&g, par ? zp0 : zp1, ap};
    double d[3] = {0.0, 0.0, 0.0};
    // the block's first matrix and (z, p) loads are in flight while the
    // hook sums the previous partials
    row_engine<V, MT, VT>(n, ptr, idx, val, g, rowblocks, nrb, epi, d, hook);
    const double s0 = block_sum<kStreamBlock>(d[0]);
    if (threadIdx.x == 0) partials[blockIdx.x] = s0;
}

// Three-launch form of an iteration (where forming the direction inside the
// SpMV's gathers costs more than a launch of its own), on dense z and p.
// Launch 1: the tail of the previous iteration (every block), then
// p = z + beta p in place (the expression of XDir).
// A thread's first kDirRows rows are loaded before the tail's partial sums,
// so that their fetch overlaps the sums' latency.
constexpr int kDirRows = 8;

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, const VT *__restrict__ z,
                                                    VT *__restrict__ p,
                                                    const double *__restrict__ partials,
                                                    int stride, KrylovCtl *ctl)
{
    const int i0 = xcd_block() * kBlock + threadIdx.x;
    const int step = gridDim.x * kBlock;
    VT zk[kDirRows], pk[kDirRows];
#pragma unroll
    for (int k = 0; k < kDirRows; ++k) {
        const int i = i0 + k * step;
        zk[k] = i < n ? z[i] : VT(0);
        pk[k] = i < n ? p[i] : VT(0);
    }
    // the control block is read with the rows in flight
    if (ctl->done) return;
    const int restart = ctl->restart;
    double beta_d;
    if (pcg_finish<kBlock>(partials, stride, ctl, beta_d)) return;
    const VT beta = (VT)beta_d;
#pragma unroll
    for (int k = 0; k < kDirRows; ++k) {
        const int i = i0 + k * step;
        if (i < n) p[i] = restart ? zk[k] : fma_t(beta, pk[k], zk[k]);
    }
    for (int i = i0 + kDirRows * step; i < n; i += step)
        p[i] = restart ? z[i] : fma_t(beta, p[i], z[i]);
}

// Launch 2 of the three-launch form: Ap = A p and the partials of p.Ap.
template <int V, typename MT, typename VT>
__global__ __launch_bounds__(kStreamBlock) void k_pcg_apdot(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx, const MT *__restrict__ val,
    const VT *__restrict__ p, VT *__restrict__ ap, const int *__restrict__ rowblocks,
    int nrb, double *__restrict__ partials, const KrylovCtl *ctl)
{
    EpiApDot<VT> epi{p, ap};
    double d[3] = {0.0, 0.0, 0.0};
    bool stop = false;
    row_engine<V, MT, VT>(n, ptr, idx, val, XPlain<VT>{p}, rowblocks, nrb, epi, d,
                          DoneHook{ctl, &stop});
    if (stop) return;
    const double s0 = block_sum<kStreamBlock>(d[0]);
    if (threadIdx.x == 0) partials[blockIdx.x] = s0;
}

// p.Ap from the SpMV launch's partials -> alpha (or the breakdown stop,
// :897-899), recorded by block 0; false when the solve stops here.
__device__ __forceinline__ bool update_alpha(const double *__restrict__ partials, int nb_spmv,
                                             KrylovCtl *ctl, int cur, double &alpha_d)
{
    const double pap = sum_partials<kBlock>(partials, nb_spmv);
    const double rz_old = ctl->rzs[cur];
    if (is_breakdown(pap, rz_old, ctl->breakdown_mode)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->pap = pap;
            ctl->done = 1;
            ctl->breakdown_hit = 1;
        }
        return false;
    }
    alpha_d = rz_old / pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->pap = pap;
        ctl->alpha = alpha_d;
        ctl->k2_done = ctl->k2_done + 1;
        ctl->restart = 0;
        ctl->nb_rz = gridDim.x;
    }
    return true;
}

// Launch 3 of the three-launch form, on dense z and p: alpha ; x += alpha p ;
// r -= alpha Ap ; z = D r ; partials of r.r and r.z.  x may be FP64 while
// the iteration vectors are FP32 (mixed precision).
template <typename VT, typename XT>
__global__ __launch_bounds__(kBlock) void k_pcg_update3(
    int n, XT *__restrict__ x, VT *__restrict__ r, VT *__restrict__ z, const VT *__restrict__ p,
    const VT *__restrict__ ap, const VT *__restrict__ dinv, double *__restrict__ partials,
    int stride, int nb_spmv, KrylovCtl *ctl)
{
    // this thread's first kDirRows rows, loaded while the control block is
    // read and the partials are summed
    const int i0 = xcd_block() * kBlock + threadIdx.x;
    const int step = gridDim.x * kBlock;
    XT xk[kDirRows];
    VT rk[kDirRows], ak[kDirRows], dk[kDirRows], pk[kDirRows];
#pragma unroll
    for (int k = 0; k < kDirRows; ++k) {
        const int i = i0 + k * step;
        const bool in = i < n;
        xk[k] = in ? x[i] : XT(0);
        rk[k] = in ? r[i] : VT(0);
        ak[k] = in ? ap[i] : VT(0);
        dk[k] = in ? dinv[i] : VT(0);
        pk[k] = in ? p[i] : VT(0);
    }
    if (ctl->done) return;
    const int cur = ctl->cur;
    double alpha_d;
    if (!update_alpha(partials, nb_spmv, ctl, cur, alpha_d)) return;
    const VT alpha = (VT)alpha_d;
    const XT alpha_x = (XT)alpha_d;
    double rr = 0.0, rz = 0.0;
#pragma unroll
    for (int k = 0; k < kDirRows; ++k) {
        const int i = i0 + k * step;
        if (i >= n) continue;
        x[i] = xk[k] + alpha_x * (XT)pk[k];
        const VT ri = rk[k] - alpha * ak[k];
        r[i] = ri;
        const VT zi = dk[k] * ri;
        z[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    for (int i = i0 + kDirRows * step; i < n; i += step) {
        x[i] = x[i] + alpha_x * (XT)p[i];
        const VT ri = r[i] - alpha * ap[i];
        r[i] = ri;
        const VT zi = dinv[i] * ri;
        z[i] = zi;
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

// Launch 2 of the two-launch form: p.Ap -> alpha (or breakdown, :897-899);
// x += alpha p ; r -= alpha Ap ; z = D r next to p ; partials of r.r and
// r.z.  x may be FP64 while the iteration vectors are FP32 (mixed precision).
template <typename VT, typename XT>
__global__ __launch_bounds__(kBlock) void k_pcg_update(
    int n, XT *__restrict__ x, VT *__restrict__ r, ZP<VT> *__restrict__ zp0,
    ZP<VT> *__restrict__ zp1, const VT *__restrict__ ap, const VT *__restrict__ dinv,
    double *__restrict__ partials, int stride, int nb_spmv, KrylovCtl *ctl)
{
    if (ctl->done) return;
    const int cur = ctl->cur;
    ZP<VT> *zp = cur ? zp1 : zp0;
    // this thread's first element, loaded while the partials are summed
    const int i0 = xcd_block() * kBlock + threadIdx.x;
    XT x0 = XT(0);
    VT r0 = VT(0), a0 = VT(0), d0 = VT(0);
    ZP<VT> q0{};
    if (i0 < n) {
        x0 = x[i0];
        r0 = r[i0];
        a0 = ap[i0];
        d0 = dinv[i0];
        q0 = zp[i0];
    }
    double alpha_d;
    if (!update_alpha(partials, nb_spmv, ctl, cur, alpha_d)) return;
    const VT alpha = (VT)alpha_d;
    const XT alpha_x = (XT)alpha_d;
    double rr = 0.0, rz = 0.0;
    if (i0 < n) {
        x[i0] = x0 + alpha_x * (XT)q0.p;
        const VT ri = r0 - alpha * a0;
        r[i0] = ri;
        const VT zi = d0 * ri;
        zp[i0] = ZP<VT>{zi, q0.p};
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    for (int i = i0 + gridDim.x * kBlock; i < n; i += gridDim.x * kBlock) {
        const ZP<VT> q = zp[i];
        x[i] = x[i] + alpha_x * (XT)q.p;
        const VT ri = r[i] - alpha * ap[i];
        r[i] = ri;
        const VT zi = dinv[i] * ri;
        zp[i] = ZP<VT>{zi, q.p};
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

// end of a chunk: the tail of the last iteration (stop test) for the host
__global__ __launch_bounds__(kBlock) void k_pcg_fin(const double *__restrict__ partials,
                                                    int stride, KrylovCtl *ctl)
{
    if (ctl->done) return;
    double beta;
    (void)pcg_finish<kBlock>(partials, stride, ctl, beta);
}

// Jacobi diagonal (build_sparse_preconditioner 'jacobi', :1130-1144):
// first entry of row i with column i; 1/a_ii if |a_ii| > 1e-14 else 1.
// dinv32 (optional) receives the FP32 copy.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const double *__restrict__ val,
                                                   int use, VT *__restrict__ dinv,
                                                   float *__restrict__ dinv32)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double a = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (idx[k] == i) {
                a = val[k];
                break;
            }
        const double di = (use && fabs(a) > 1.0e-14) ? 1.0 / a : 1.0;
        dinv[i] = (VT)di;
        if (dinv32) dinv32[i] = (float)di;
    }
}

// r = b - A x (FP64) and partials of ||r||^2 -- the true residual of the
// final FP64 report.
template <int V>
__global__ __launch_bounds__(kStreamBlock) void k_residual(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    const int *__restrict__ rowblocks, int nrb, double *__restrict__ partials)
{
    EpiResidual epi{b, r};
    double d[3] = {0.0, 0.0, 0.0};
    row_engine<V, double, double>(n, ptr, idx, val, XPlain<double>{x}, rowblocks, nrb, epi, d);
    const double s0 = block_sum<kStreamBlock>(d[0]);
    if (threadIdx.x == 0) partials[blockIdx.x] = s0;
}

__global__ __launch_bounds__(kBlock) void k_residual_finish(const double *__restrict__ partials,
                                                            int nb, KrylovCtl *ctl)
{
    const double t = sum_partials<kBlock>(partials, nb);
    if (threadIdx.x == 0) ctl->true_rr = t;
}

// Mixed precision: the FP32 iterations of a chunk add their steps alpha p to
// an FP32 correction dx (4 bytes read and written per row and iteration
// instead of FP64 x's 8); before each refresh x += dx in FP64 and dx = 0, so
// the refresh, and every result, sees x with the whole chunk's correction.
__global__ __launch_bounds__(kBlock) void k_fold_dx(int n, double *__restrict__ x,
                                                    float *__restrict__ dx)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        x[i] += (double)dx[i];
        dx[i] = 0.0f;
    }
}

// Mixed precision, reliable update: the FP64 residual b - A x (FP64 matrix
// and x) replaces the FP32 recurrence residual; partials of its r.r and
// r.(D r) go where the next direction launch reads them, with the drift
// ||r_recurrence - r||^2.
template <int V>
__global__ __launch_bounds__(kStreamBlock) void k_mixed_refresh(
    int n, const int *__restrict__ ptr, const int *__restrict__ idx,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, float *__restrict__ r32, const float *__restrict__ dinv32,
    float *__restrict__ z0, float *__restrict__ z1, int zs, const int *__restrict__ rowblocks,
    int nrb, double *__restrict__ partials, int stride, const KrylovCtl *ctl)
{
    if (ctl->converged) return;
    // z goes where the next direction launch reads it: the dense z of the
    // three-launch form (z0 = z1, zs = 1), or the (z, p) array of slot
    // k2_done & 1 (zs = 2)
    EpiRefresh epi{b, dinv32, r32, (ctl->k2_done & 1) ? z1 : z0, zs};
    double d[3] = {0.0, 0.0, 0.0};
    row_engine<V, double, double>(n, ptr, idx, val, XPlain<double>{x}, rowblocks, nrb, epi, d);
    const double s0 = block_sum<kStreamBlock>(d[0]);
    const double s1 = block_sum<kStreamBlock>(d[1]);
    const double s2 = block_sum<kStreamBlock>(d[2]);
    if (threadIdx.x == 0) {
        partials[stride + blockIdx.x] = s0;
        partials[2 * stride + blockIdx.x] = s1;
        partials[3 * stride + blockIdx.x] = s2;
    }
}

// Mixed precision: convergence on the FP64 residual (||r|| <= tol, with tol
// from the first one, :1063-1073), the iteration cap, and whether the
// direction is kept (the recurrence tracked the FP64 residual to 1 %) or CG
// restarts from D r (larger drift, a breakdown, the first refresh).
__global__ __launch_bounds__(kBlock) void k_mixed_refresh_finish(const double *__restrict__ partials,
                                                                 int stride, int nb, int first,
                                                                 KrylovCtl *ctl)
{
    if (ctl->converged) return;
    double trr, tdd;
    sum_partials2<kBlock>(partials + stride, partials + 3 * stride, nb, trr, tdd);
    if (threadIdx.x == 0) {
        const double rn = sqrt(trr);
        if (first) {
            ctl->initial_norm = rn;
            ctl->tol = ctl->tol_type == 1 ? ctl->tolerance : ctl->tolerance * rn;
            ctl->iter = 0;
            ctl->k2_done = 0;
        }
        ctl->true_rr = trr;
        ctl->rr = trr;
        ctl->residual_norm = rn;
        ctl->nb_rz = nb;
        ctl->refreshes = ctl->refreshes + 1;
        if (rn <= ctl->tol) {
            ctl->converged = 1;
            ctl->done = 1;
        } else if (ctl->k2_done >= ctl->max_iter) {
            ctl->done = 1;
        } else {
            if (first || ctl->breakdown_hit || tdd > 1.0e-4 * trr) ctl->restart = 1;
            ctl->breakdown_hit = 0;
            ctl->done = 0;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ a,
                                                   float *__restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

// SELL-64 entries (padding: zero value on the row's own column, source
// position -1), one thread per slice position.
__global__ __launch_bounds__(kBlock) void k_sell_fill(const int *__restrict__ ptr,
                                                      const int *__restrict__ idx, int nslice,
                                                      const int *__restrict__ off,
                                                      const int *__restrict__ srow,
                                                      int *__restrict__ sidx,
                                                      int *__restrict__ src)
{
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (t >= (int64_t)nslice * 64) return;
    const int sl = (int)(t >> 6), lane = (int)(t & 63);
    const int row = srow[t];
    const int o = off[sl], w = (off[sl + 1] - o) >> 6;
    const int b = row >= 0 ? ptr[row] : 0;
    const int len = row >= 0 ? ptr[row + 1] - b : 0;
    const int pad = row >= 0 ? row : 0;
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
                                        {1, 2}, {2, 1}, {2, 2}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// Bytes one SpMV launch of variant vi must move in the format that variant
// reads (values s_v bytes each; x read and y written once): CSR engines
// nnz (s_v + 4) + 4 (n + 1) + 2 n s_v; SELL-64 its padded entries
// (s_v + 4 each), slice offsets and row map; tiled SELL-64 its padded
// entries (s_v + 2 each: 16-bit tile positions), slice offsets, row map,
// tile offsets and tile column lists.
double variant_bytes(const DevMatrix &A, int vi, int sv)
{
    const int engine = kVariants[vi].engine;
    const double vec = 2.0 * sv * (double)A.n;
    if (engine == 1 && A.sell_ok)
        return (double)A.sell_total * (sv + 4) + 4.0 * (A.nslice + 1) + 256.0 * A.nslice + vec;
    if (engine == 2 && A.tiles_ok) {
        const int ngroup = (A.nslice + 3) / 4;
        return (double)A.sell_total * (sv + 2) + 4.0 * (A.nslice + 1) + 256.0 * A.nslice +
               4.0 * (ngroup + 1) + 4.0 * (double)A.tile_total + vec;
    }
    return (double)A.nnz * (sv + 4) + 4.0 * (A.n + 1) + vec;
}

static int variant_grid(const Context *ctx, const DevMatrix &A, int vi)
{
    const SpmvVariant v = kVariants[vi];
    const int64_t cap = (int64_t)grid_cap(ctx) * v.gmul;
    int64_t need;
    if (v.engine == 0)
        need = A.nrb;
    else if (v.engine == 1 || v.engine == 2)
        need = ((int64_t)A.nslice + kStreamBlock / 64 - 1) / (kStreamBlock / 64);
    else
        need = ((int64_t)A.n * v.engine + kStreamBlock - 1) / kStreamBlock;
    return (int)std::max<int64_t>(1, std::min(need, cap));
}

// dynamic LDS of a launch: the tiled engine stages one value (bytes vbytes)
// per distinct column of a group
static size_t variant_lds(const DevMatrix &A, int vi, size_t vbytes)
{
    return kVariants[vi].engine == 2 ? (size_t)A.tile_max * vbytes : 0;
}

// The grid of a grid-stride row-engine launch, capped at the blocks of that
// kernel (at its register and LDS use) that fit on the device at once: a
// grid past that runs its last blocks after the others, so every block's
// share of the rows would wait on a second round (occupancy per kernel and
// dynamic LDS size, cached).
// (a fixed table, so that no call allocates: the allocation-free solve
// launches through here)
static int resident_grid(const Context *ctx, const void *kernel, int grid, size_t lds)
{
    struct Entry {
        const void *kernel;
        size_t lds;
        int per_cu;
    };
    constexpr int kEntries = 256;
    static std::mutex mu;
    static Entry table[kEntries];
    static int used = 0;
    int per_cu = -1;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (int i = 0; i < used; ++i)
            if (table[i].kernel == kernel && table[i].lds == lds) per_cu = table[i].per_cu;
        if (per_cu < 0) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kStreamBlock, lds) !=
                hipSuccess)
                per_cu = 0;
            if (used < kEntries) table[used++] = Entry{kernel, lds, per_cu};
        }
    }
    if (per_cu <= 0) return grid;
    return std::max(1, std::min(grid, per_cu * ctx->cu_count));
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
    // a SELL variant launches only on the arrays ready_variant made current
    FFH_REQUIRE(kVariants[vi].engine != 1 || A.sell_ok, FFH_ERR_DEVICE,
                "SELL-64 variant launched without its SELL copy");
    FFH_REQUIRE(kVariants[vi].engine != 2 || (A.sell_ok && A.tiles_ok && A.tile_pack.p &&
                                              A.tile_lidx.p),
                FFH_ERR_DEVICE, "tiled SELL-64 variant launched without its column tiles");
    if (kVariants[vi].engine == 1) {
        if constexpr (sizeof(MT) == 8)
            return {A.sell_off.p, A.sell_idx.p, A.sell_val.p, A.sell_row.p, A.nslice};
        else
            return {A.sell_off.p, A.sell_idx.p, A.sell_val32.p, A.sell_row.p, A.nslice};
    }
    if (kVariants[vi].engine == 1 || kVariants[vi].engine == 2) {
        const void *sv = sizeof(MT) == 8 ? (const void *)A.sell_val.p : (const void *)A.sell_val32.p;
        FFH_REQUIRE(sv != nullptr, FFH_ERR_DEVICE, "SELL-64 variant launched without its values");
    }
    if (kVariants[vi].engine == 2) {
        const int *lidx = reinterpret_cast<const int *>(A.tile_lidx.p);
        if constexpr (sizeof(MT) == 8)
            return {A.sell_off.p, lidx, A.sell_val.p, A.tile_pack.p, A.nslice};
        else
            return {A.sell_off.p, lidx, A.sell_val32.p, A.tile_pack.p, A.nslice};
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
    case 2: { constexpr int VV = 2; __VA_ARGS__; } break;                     \
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
    A.variant_dir[0] = A.variant_dir[1] = -1;
    A.sell_built = A.sell_ok = false;
    A.tiles_built = A.tiles_ok = false;
    A.sell64_valid = A.sell32_valid = false;
    A.tr_built = false;
    FFH_HIP(hipStreamSynchronize(ctx->stream));
}

Workspace::~Workspace()
{
    for (auto &g : graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto &e : ctl_ev)
        if (e) (void)hipEventDestroy(e);
    if (host_ctl) (void)hipHostFree(host_ctl);
}

static int partial_stride(const Context *ctx) { return grid_cap(ctx) * 4; }

void ensure_workspace(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int n = M.m.n;
    if (w.n != n) invalidate_graphs(w);
    w.n = n;
    // p and p2 hold the interleaved (z, p) pairs of the PCG iteration
    w.r.alloc(n), w.p.alloc(2 * (size_t)n), w.p2.alloc(2 * (size_t)n), w.ap.alloc(n),
        w.x.alloc(n), w.b.alloc(n), w.dinv.alloc(n);
    if (!w.partials.p) {
        w.partials.alloc(4 * (size_t)partial_stride(ctx));
        FFH_HIP(hipMemsetAsync(w.partials.p, 0, sizeof(double) * w.partials.n, ctx->stream));
    }
    w.ctl.alloc(1);
    if (!w.host_ctl)
        FFH_HIP(hipHostMalloc((void **)&w.host_ctl, 2 * sizeof(KrylovCtl),
                              hipHostMallocDefault));
    for (auto &e : w.ctl_ev)
        if (!e) FFH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

static void ensure_workspace32(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int n = M.m.n;
    w.r32.alloc(n), w.p32.alloc(2 * (size_t)n), w.p32b.alloc(2 * (size_t)n), w.ap32.alloc(n),
        w.x32.alloc(n), w.dinv32.alloc(n);
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

// SELL-64 structure of the gather arrays (rows sorted by length within
// windows of 1024, which keeps the padding near 5-15 % for tetrahedral
// matrices whose vertex, edge and face rows differ in length).  Not built
// (sell_ok false, the SELL variants fall back to the stream engine) when the
// padding would more than triple the stored entries or the entry count
// exceeds int32.
static void ensure_sell_structure(Context *ctx, DevMatrix &A)
{
    if (A.sell_built) return;
    A.sell_built = true;
    A.sell_ok = false;
    const int n = A.n;
    if (n == 0 || A.nnz == 0) return;
    constexpr int kSigma = 1024; // rows sorted by length within windows of 16 slices
    const int nslice = (n + 63) / 64;
    hipStream_t s = ctx->stream;
    std::vector<int> ptr(n + 1);
    FFH_HIP(hipMemcpyAsync(ptr.data(), A.gptr(), sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<int> order((size_t)nslice * 64, -1);
    for (int w0 = 0; w0 < n; w0 += kSigma) {
        const int w1 = std::min(n, w0 + kSigma);
        std::vector<int> rows(w1 - w0);
        for (int i = w0; i < w1; ++i) rows[i - w0] = i;
        std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) {
            return ptr[a + 1] - ptr[a] > ptr[b + 1] - ptr[b];
        });
        std::copy(rows.begin(), rows.end(), order.begin() + w0);
    }
    std::vector<int> off(nslice + 1, 0);
    int64_t total = 0;
    for (int sl = 0; sl < nslice; ++sl) {
        int w = 0;
        for (int l = 0; l < 64; ++l) {
            const int r = order[(size_t)sl * 64 + l];
            if (r >= 0) w = std::max(w, ptr[r + 1] - ptr[r]);
        }
        total += 64ll * w;
        if (total > INT32_MAX) return;
        off[sl + 1] = (int)total;
    }
    if (total > 3 * A.nnz + (1 << 20)) return;
    A.sell_off.upload(off, s);
    A.sell_row.upload(order, s);
    A.sell_idx.alloc(std::max<int64_t>(total, 1));
    A.sell_src.alloc(std::max<int64_t>(total, 1));
    hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)(((int64_t)nslice * 64 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, A.gptr(), A.gidx(), nslice, A.sell_off.p, A.sell_row.p,
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

// Column tiles of the SELL-64 copy (see DevMatrix::tile_*): for each group of
// four slices the sorted distinct columns of its entries and the 16-bit
// position of every entry's column in that list.  Host work, once per
// structure.
void ensure_sell_tiles(Context *ctx, DevMatrix &A)
{
    if (A.tiles_built) return;
    ensure_sell_structure(ctx, A);
    A.tiles_built = true;
    A.tiles_ok = false;
    if (!A.sell_ok) return;
    const int nslice = A.nslice;
    const int ngroup = (nslice + 3) / 4;
    hipStream_t s = ctx->stream;
    std::vector<int> off(nslice + 1), idx((size_t)A.sell_total), srow((size_t)nslice * 64);
    FFH_HIP(hipMemcpyAsync(off.data(), A.sell_off.p, sizeof(int) * (nslice + 1), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipMemcpyAsync(idx.data(), A.sell_idx.p, sizeof(int) * idx.size(), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipMemcpyAsync(srow.data(), A.sell_row.p, sizeof(int) * srow.size(), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<std::vector<int>> cols(ngroup);
    std::vector<uint16_t> lidx(idx.size());
    parallel_for(ngroup, [&](int64_t lo, int64_t hi) {
        for (int64_t g = lo; g < hi; ++g) {
            const int e0 = off[4 * g], e1 = off[std::min<int64_t>(4 * g + 4, nslice)];
            std::vector<int> &c = cols[g];
            c.assign(idx.begin() + e0, idx.begin() + e1);
            std::sort(c.begin(), c.end());
            c.erase(std::unique(c.begin(), c.end()), c.end());
            if (c.size() > 65536) continue;
            for (int e = e0; e < e1; ++e)
                lidx[e] = (uint16_t)(std::lower_bound(c.begin(), c.end(), idx[e]) - c.begin());
        }
    }, 64);
    std::vector<int> toff(ngroup + 1, 0);
    int mx = 0;
    for (int g = 0; g < ngroup; ++g) {
        if (cols[g].size() > 65536) return;
        toff[g + 1] = toff[g] + (int)cols[g].size();
        mx = std::max(mx, (int)cols[g].size());
    }
    std::vector<int> pack(srow.size() + (size_t)ngroup + 1 + (size_t)toff[ngroup]);
    std::copy(srow.begin(), srow.end(), pack.begin());
    std::copy(toff.begin(), toff.end(), pack.begin() + srow.size());
    int *tc = pack.data() + srow.size() + ngroup + 1;
    for (int g = 0; g < ngroup; ++g) std::copy(cols[g].begin(), cols[g].end(), tc + toff[g]);
    A.tile_off.upload(toff, s);
    A.tile_cols.upload(tc, (size_t)toff[ngroup], s);
    A.tile_lidx.upload(lidx, s);
    A.tile_pack.upload(pack, s);
    FFH_HIP(hipStreamSynchronize(s));
    A.tile_max = mx;
    A.tile_total = toff[ngroup];
    A.tiles_ok = true;
}

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
    const int eng = kVariants[vi].engine;
    if (eng != 1 && eng != 2) return vi;
    ensure_sell_values(ctx, A, f32);
    if (!A.sell_ok) return 0;
    if (eng == 2) {
        ensure_sell_tiles(ctx, A);
        // the FP64 launches stage 8 bytes per column in at most 64 KiB of
        // dynamic LDS (their static LDS comes on top, within the 160 KiB a
        // workgroup may hold)
        if (!A.tiles_ok || (size_t)A.tile_max * 8 > 65536) return 0;
    }
    return vi;
}

// y = A x with variant vi (its arrays current: see ready_variant)
template <typename MT, typename VT>
static void launch_spmv(Context *ctx, const DevMatrix &A, const VT *x, VT *y, int vi)
{
    const GArgs<MT> g = gargs<MT>(A, vi);
    const size_t lds = variant_lds(A, vi, sizeof(VT));
    FFH_ENGINE(kVariants[vi].engine, {
        const int grid = resident_grid(ctx, (const void *)k_spmv<VV, MT, VT>,
                                       variant_grid(ctx, A, vi), lds);
        hipLaunchKernelGGL((k_spmv<VV, MT, VT>), dim3(grid), dim3(kStreamBlock), lds,
                           ctx->stream, A.n, g.ptr, g.idx, g.val, x, y, g.rb, g.nrb);
    });
    FFH_HIP(hipGetLastError());
}

// What the other launches of a PCG iteration stream between two of its SpMV
// launches: the direction p rewritten in place (the direction launch) and m
// more values read and written (the update's vectors), so that an SpMV
// launch timed after it meets the caches as it does inside the solve.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_tune_stream(int n, VT *__restrict__ p, int64_t m,
                                                        float *__restrict__ buf)
{
    const int64_t step = (int64_t)gridDim.x * kBlock;
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    for (int64_t i = i0; i < n; i += step) p[i] = p[i] + VT(0);
    for (int64_t j = i0; j < m; j += step) buf[j] = buf[j] + 1.0f;
}

// Autotuner: the fastest SpMV variant for this matrix and precision, timed
// with events of its own on the context stream, into scratch vectors,
// partials and control block of its own, so that no solver vector or timing
// event is touched.  Each variant's launch of the solver's SpMV kernel is
// timed 10 times after a warm-up, each time after k_tune_stream has moved
// the bytes of the iteration's other two launches (about 56 bytes per row in
// FP32 iterations, 88 in FP64): timed back to back instead, a variant whose
// matrix partly stays in the memory-side cache between launches can look as
// fast as one that is 20 % faster inside the solve.
// FFH_SPMV_VARIANT=<index> forces a variant.
// best_dir receives the fastest variant that gathers from L2 (not the tiled
// engine): the CG direction launch forms each gathered column's direction on
// the fly, which the tiled engine's staging serialises behind its start
// hook, so that launch takes the best gathering variant.
template <typename MT, typename VT>
static int tune_spmv(Context *ctx, DevMatrix &A, int &best_dir)
{
    constexpr bool f32 = sizeof(MT) == 4;
    best_dir = 0;
    if (const char *env = std::getenv("FFH_SPMV_VARIANT")) {
        int v = std::atoi(env);
        if (v >= 0 && v < kNumVariants) return best_dir = ready_variant(ctx, A, v, f32);
    }
    if (A.nnz == 0 || A.n == 0) return 0;
    DBuf<VT> tx, ty;
    tx.alloc(A.n);
    ty.alloc(A.n);
    FFH_HIP(hipMemsetAsync(tx.p, 0, sizeof(VT) * A.n, ctx->stream));
    // the launch timed is the solver's own SpMV launch (k_pcg_apdot: the
    // product, its p.Ap partials and its start hook), whose register use,
    // and so occupancy, can differ from the plain product's
    DBuf<double> part;
    part.alloc((size_t)grid_cap(ctx) * 8);
    DBuf<KrylovCtl> tctl;
    tctl.alloc(1);
    FFH_HIP(hipMemsetAsync(tctl.p, 0, sizeof(KrylovCtl), ctx->stream));
    // the update's and direction's other vector bytes, as float values
    const int64_t m = (int64_t)A.n * ((f32 ? 56 : 88) - 2 * (int64_t)sizeof(VT)) / 8;
    DBuf<float> scrub;
    scrub.alloc((size_t)std::max<int64_t>(1, m));
    FFH_HIP(hipMemsetAsync(scrub.p, 0, sizeof(float) * std::max<int64_t>(1, m), ctx->stream));
    const int gstream = grid_cap(ctx);
    auto stream_others = [&]() {
        hipLaunchKernelGGL((k_tune_stream<VT>), dim3(gstream), dim3(kBlock), 0, ctx->stream, A.n,
                           tx.p, m, scrub.p);
    };
    auto launch = [&](int vi) {
        const GArgs<MT> g = gargs<MT>(A, vi);
        const size_t lds = variant_lds(A, vi, sizeof(VT));
        FFH_ENGINE(kVariants[vi].engine, {
            const int grid = resident_grid(ctx, (const void *)k_pcg_apdot<VV, MT, VT>,
                                           variant_grid(ctx, A, vi), lds);
            hipLaunchKernelGGL((k_pcg_apdot<VV, MT, VT>), dim3(grid), dim3(kStreamBlock), lds,
                               ctx->stream, A.n, g.ptr, g.idx, g.val, tx.p, ty.p, g.rb, g.nrb,
                               part.p, tctl.p);
        });
        FFH_HIP(hipGetLastError());
    };
    constexpr int kReps = 10;
    hipEvent_t ev[2 * kReps] = {};
    int best = 0;
    auto destroy = [&]() {
        for (hipEvent_t &e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    try {
        for (hipEvent_t &e : ev) FFH_HIP(hipEventCreate(&e));
        float best_ms = 1e30f, best_dir_ms = 1e30f;
        for (int vi = 0; vi < kNumVariants; ++vi) {
            if (kVariants[vi].engine == 0 && kVariants[vi].gmul > 1 &&
                A.nrb <= grid_cap(ctx) * (kVariants[vi].gmul / 2))
                continue; // same grid as the variant before it
            if ((kVariants[vi].engine == 1 || kVariants[vi].engine == 2) &&
                ready_variant(ctx, A, vi, f32) != vi)
                continue; // no SELL copy (or no column tiles) of this matrix
            launch(vi);
            for (int r = 0; r < kReps; ++r) {
                stream_others();
                FFH_HIP(hipEventRecord(ev[2 * r], ctx->stream));
                launch(vi);
                FFH_HIP(hipEventRecord(ev[2 * r + 1], ctx->stream));
            }
            FFH_HIP(hipGetLastError());
            FFH_HIP(hipEventSynchronize(ev[2 * kReps - 1]));
            float ms = 0.0f;
            for (int r = 0; r < kReps; ++r) {
                float t = 0.0f;
                FFH_HIP(hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
                ms += t;
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = vi;
            }
            if (kVariants[vi].engine != 2 && ms < best_dir_ms) {
                best_dir_ms = ms;
                best_dir = vi;
            }
        }
    } catch (...) {
        destroy();
        throw;
    }
    destroy();
    return best;
}

// The variant of a matrix for a precision, tuned on first use.  Call outside
// graph capture and before a solve's timed region.
static int spmv_variant(Context *ctx, Matrix &M, bool f32)
{
    DevMatrix &A = M.m;
    int &v = A.variant[f32 ? 1 : 0];
    int &vd = A.variant_dir[f32 ? 1 : 0];
    if (v < 0) {
        if (f32) {
            ensure_val32(ctx, A);
            v = tune_spmv<float, float>(ctx, A, vd);
        } else {
            v = tune_spmv<double, double>(ctx, A, vd);
        }
    }
    v = ready_variant(ctx, A, v, f32);
    return v;
}

// the variant of the CG direction launch (tuned with spmv_variant)
static int dir_variant(Context *ctx, Matrix &M, bool f32)
{
    (void)spmv_variant(ctx, M, f32);
    DevMatrix &A = M.m;
    int &vd = A.variant_dir[f32 ? 1 : 0];
    if (vd < 0) vd = A.variant[f32 ? 1 : 0];
    vd = ready_variant(ctx, A, vd, f32);
    return vd;
}

int select_spmv_variant(Context *ctx, Matrix &M, bool f32, int request)
{
    static_assert(kNumVariants == FFH_SPMV_VARIANTS, "variant table and header disagree");
    ensure_val64(ctx, M.m);
    if (f32) ensure_val32(ctx, M.m);
    if (request >= 0) {
        request = ready_variant(ctx, M.m, request, f32);
        M.m.variant[f32 ? 1 : 0] = request;
        M.m.variant_dir[f32 ? 1 : 0] = request;
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

// The form of a Jacobi-PCG iteration: three launches on dense z and p
// (k_pcg_dir, k_pcg_apdot with the plain SpMV variant, k_pcg_update3), the
// faster form in FP32 and FP64 on the P2 32^3 x 6 and P1 128^3 x 6 Poisson
// systems; FFH_PCG_LAUNCHES=2 (read once per process) selects the two-launch
// form on (z, p) pairs (k_pcg_dirspmv, k_pcg_update) instead.
bool three_launch(int64_t n, bool f32)
{
    (void)n;
    (void)f32;
    static const bool two = [] {
        const char *e = std::getenv("FFH_PCG_LAUNCHES");
        return e && std::atoi(e) == 2;
    }();
    return !two;
}

// `iters` iterations of Jacobi-PCG (matrix MT, vectors VT, x in XT).  Two
// launches each (direction + SpMV with variant vi, then the x/r update), or
// for large systems three (direction in place, SpMV with variant vi_sp,
// update), both with the reference's recurrence.
template <typename MT, typename VT, typename XT>
static void launch_iterations(Context *ctx, const DevMatrix &A, double *partials, KrylovCtl *ctl,
                              int iters, XT *x, VT *r, VT *zpa, VT *zpb, VT *ap, const VT *dinv,
                              int vi, int vi_sp, int prof_slot)
{
    ZP<VT> *zp0 = reinterpret_cast<ZP<VT> *>(zpa), *zp1 = reinterpret_cast<ZP<VT> *>(zpb);
    const int n = A.n;
    if (three_launch(n, sizeof(VT) == 4)) {
        // dense z (the first n values of the first array) and p (those of
        // the second)
        VT *zl = zpa, *pl = zpb;
        const int gridv = grid_for(ctx, n, kBlock);
        const int gridd = (int)std::min<int64_t>(
            gridv, std::max<int64_t>(ctx->cu_count, ((int64_t)n + 4 * kBlock - 1) / (4 * kBlock)));
        int grids = variant_grid(ctx, A, vi_sp);
        FFH_ENGINE(kVariants[vi_sp].engine, {
            grids = resident_grid(ctx, (const void *)k_pcg_apdot<VV, MT, VT>, grids,
                                  variant_lds(A, vi_sp, sizeof(VT)));
        });
        const int stride = partial_stride(ctx);
        const GArgs<MT> g = gargs<MT>(A, vi_sp);
        for (int it = 0; it < iters; ++it) {
            hipLaunchKernelGGL((k_pcg_dir<VT>), dim3(gridd), dim3(kBlock), 0, ctx->stream, n, zl,
                               pl, partials, stride, ctl);
            if (prof_slot >= 0) ctx->prof_mark(prof_slot, true);
            FFH_ENGINE(kVariants[vi_sp].engine, {
                hipLaunchKernelGGL((k_pcg_apdot<VV, MT, VT>), dim3(grids), dim3(kStreamBlock),
                                   variant_lds(A, vi_sp, sizeof(VT)), ctx->stream, n, g.ptr,
                                   g.idx, g.val, pl, ap, g.rb, g.nrb, partials, ctl);
            });
            if (prof_slot >= 0) ctx->prof_mark(prof_slot, false);
            hipLaunchKernelGGL((k_pcg_update3<VT, XT>), dim3(gridd), dim3(kBlock), 0, ctx->stream,
                               n, x, r, zl, pl, ap, dinv, partials, stride, grids, ctl);
        }
        FFH_HIP(hipGetLastError());
        return;
    }
    // about four rows per thread, at least one block per CU: every block of
    // the next direction launch re-sums this launch's partials, so fewer
    // blocks shorten that launch's start (c2: 977 -> 256 blocks)
    const int gridv = (int)std::min<int64_t>(
        grid_for(ctx, n, kBlock),
        std::max<int64_t>(ctx->cu_count, ((int64_t)n + 4 * kBlock - 1) / (4 * kBlock)));
    int grids = variant_grid(ctx, A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        grids = resident_grid(ctx, (const void *)k_pcg_dirspmv<VV, MT, VT>, grids,
                              variant_lds(A, vi, sizeof(VT)));
    });
    const int stride = partial_stride(ctx);
    const GArgs<MT> g = gargs<MT>(A, vi);
    for (int it = 0; it < iters; ++it) {
        if (prof_slot >= 0) ctx->prof_mark(prof_slot, true);
        FFH_ENGINE(kVariants[vi].engine, {
            hipLaunchKernelGGL((k_pcg_dirspmv<VV, MT, VT>), dim3(grids), dim3(kStreamBlock),
                               variant_lds(A, vi, sizeof(VT)),
                               ctx->stream, n, g.ptr, g.idx, g.val, zp0, zp1, ap, g.rb, g.nrb,
                               partials, stride, ctl);
        });
        if (prof_slot >= 0) ctx->prof_mark(prof_slot, false);
        hipLaunchKernelGGL((k_pcg_update<VT, XT>), dim3(gridv), dim3(kBlock), 0, ctx->stream, n, x,
                           r, zp0, zp1, ap, dinv, partials, stride, grids, ctl);
    }
    FFH_HIP(hipGetLastError());
}

// The FP64 residual refresh of the mixed solve (FP64 matrix, variant vi64).
static void launch_refresh(Context *ctx, Matrix &M, int vi64, int first)
{
    const DevMatrix &A = M.m;
    Workspace &w = M.ws;
    int gr = variant_grid(ctx, A, vi64);
    FFH_ENGINE(kVariants[vi64].engine, {
        gr = resident_grid(ctx, (const void *)k_mixed_refresh<VV>, gr,
                           variant_lds(A, vi64, sizeof(double)));
    });
    const int stride = partial_stride(ctx);
    const GArgs<double> g = gargs<double>(A, vi64);
    if (ctx->profiling) ctx->prof_mark(0, true);
    FFH_ENGINE(kVariants[vi64].engine, {
        hipLaunchKernelGGL((k_mixed_refresh<VV>), dim3(gr), dim3(kStreamBlock),
                           variant_lds(A, vi64, sizeof(double)), ctx->stream,
                           A.n, g.ptr, g.idx, g.val, w.x.p, w.b.p, w.r32.p, w.dinv32.p,
                           w.p32.p, three_launch(A.n, true) ? w.p32.p : w.p32b.p,
                           three_launch(A.n, true) ? 1 : 2, g.rb, g.nrb, w.partials.p,
                           stride, w.ctl.p);
    });
    if (ctx->profiling) ctx->prof_mark(0, false);
    hipLaunchKernelGGL(k_mixed_refresh_finish, dim3(1), dim3(kBlock), 0, ctx->stream, w.partials.p,
                       stride, gr, first, w.ctl.p);
    FFH_HIP(hipGetLastError());
}

// Runs chunks (enqueue_chunk) until the control block says done.  The next
// chunk is queued before the previous one's control block is read (it
// returns at once after the stop), so the host poll overlaps the GPU.
// `watch` sees each polled control block and may return false to stop early
// (the stream is then drained).  A chunk may be replayed from a HIP graph.
template <typename Enq, typename Watch>
static KrylovCtl run_chunks(Context *ctx, Matrix &M, int gslot, int64_t graph_key, bool use_graph,
                            Enq &&enqueue_chunk, Watch &&watch)
{
    Workspace &w = M.ws;
    hipStream_t s = ctx->stream;
    if (ctx->profiling) use_graph = false;
    if (use_graph && (w.graph[gslot] == nullptr || w.graph_chunk[gslot] != graph_key)) {
        if (w.graph[gslot]) {
            (void)hipGraphExecDestroy(w.graph[gslot]);
            w.graph[gslot] = nullptr;
        }
        hipGraph_t g;
        FFH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk();
        FFH_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&w.graph[gslot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FFH_HIP(e);
        w.graph_chunk[gslot] = graph_key;
    }
    auto one = [&](int k) {
        if (use_graph)
            FFH_HIP(hipGraphLaunch(w.graph[gslot], s));
        else
            enqueue_chunk();
        FFH_HIP(hipMemcpyAsync(w.host_ctl + (k & 1), w.ctl.p, sizeof(KrylovCtl),
                               hipMemcpyDeviceToHost, s));
        FFH_HIP(hipEventRecord(w.ctl_ev[k & 1], s));
    };
    one(0);
    for (int k = 0;; ++k) {
        one(k + 1);
        FFH_HIP(hipEventSynchronize(w.ctl_ev[k & 1]));
        const KrylovCtl c = w.host_ctl[k & 1];
        if (c.done || !watch(c)) break;
    }
    FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    return *w.host_ctl;
}

// r = b - A x in FP64 with variant vi and ||r||^2 into ctl->true_rr
static void launch_residual_on(Context *ctx, const DevMatrix &A, int vi, const double *x,
                               const double *b, double *r, double *partials, KrylovCtl *ctl)
{
    int gr = variant_grid(ctx, A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        gr = resident_grid(ctx, (const void *)k_residual<VV>, gr, variant_lds(A, vi, sizeof(double)));
    });
    const GArgs<double> g = gargs<double>(A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_residual<VV>), dim3(gr), dim3(kStreamBlock),
                           variant_lds(A, vi, sizeof(double)), ctx->stream, A.n,
                           g.ptr, g.idx, g.val, x, b, r, g.rb, g.nrb, partials);
    });
    hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kBlock), 0, ctx->stream, partials, gr,
                       ctl);
    FFH_HIP(hipGetLastError());
}

static void launch_residual(Context *ctx, Matrix &M)
{
    Workspace &w = M.ws;
    const int vi = spmv_variant(ctx, M, false);
    launch_residual_on(ctx, M.m, vi, w.x.p, w.b.p, w.r.p, w.partials.p, w.ctl.p);
}

// The device vectors of one FP64 Jacobi-PCG solve: x (initial guess,
// overwritten), b, r, the two interleaved (z, p) arrays (2 n each), A p,
// D^-1, the partial sums (4 banks of partial_stride) and the control block.
struct PcgBufs {
    double *x, *b, *r, *zpa, *zpb, *ap, *dinv, *partials;
    KrylovCtl *ctl;
};

__global__ void k_ctl_set(KrylovCtl *ctl, KrylovCtl v) { *ctl = v; }

// FP64 Jacobi-PCG (or CG) of the reference on the vectors of w with SpMV
// variant vi: initialisation, then chunks of iterations each ending with the
// stop test; run(enqueue_chunk, chunk) drives the chunks until the control
// block says done and returns its last state.
template <typename Run>
static KrylovCtl pcg_fp64_on(Context *ctx, const DevMatrix &A, int vi, int vi_it, const PcgBufs &w,
                             int method, int precond, double tolerance, int tol_type,
                             int max_iter, int breakdown_mode, int check_interval, Run &&run)
{
    const int n = A.n;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL((k_jacobi<double>), dim3(grid_for(ctx, n, kBlock)), dim3(kBlock), 0, s, n,
                       A.gptr(), A.gidx(), A.gval(), (method == 1 && precond == 1) ? 1 : 0,
                       w.dinv, (float *)nullptr);
    KrylovCtl init{};
    init.max_iter = max_iter;
    init.breakdown_mode = breakdown_mode;
    init.tol_type = tol_type;
    init.tolerance = tolerance;
    hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(1), 0, s, w.ctl, init);
    int gi = variant_grid(ctx, A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        gi = resident_grid(ctx, (const void *)k_pcg_init<VV, double, double>, gi,
                           variant_lds(A, vi, sizeof(double)));
    });
    const int stride = partial_stride(ctx);
    const GArgs<double> g = gargs<double>(A, vi);
    FFH_ENGINE(kVariants[vi].engine, {
        hipLaunchKernelGGL((k_pcg_init<VV, double, double>), dim3(gi), dim3(kStreamBlock),
                           variant_lds(A, vi, sizeof(double)), s, n,
                           g.ptr, g.idx, g.val, w.x, w.b, w.r, w.dinv, w.zpa,
                           three_launch(n, false) ? 1 : 2, g.rb, g.nrb, w.partials, stride);
    });
    hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), dim3(kBlock), 0, s, w.partials, stride, gi,
                       w.ctl);
    FFH_HIP(hipGetLastError());
    const int chunk = check_interval > 0 ? check_interval : 16;
    const int prof = ctx->profiling ? 0 : -1;
    auto enqueue = [&]() {
        launch_iterations<double, double, double>(ctx, A, w.partials, w.ctl, chunk, w.x, w.r,
                                                  w.zpa, w.zpb, w.ap, w.dinv, vi_it, vi, prof);
        hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(kBlock), 0, s, w.partials, stride, w.ctl);
        FFH_HIP(hipGetLastError());
    };
    return run(enqueue, chunk);
}

// FP64 Jacobi-PCG on a matrix's workspace vectors b, x (x = initial guess,
// overwritten), chunks polled one behind the GPU (run_chunks).
static KrylovCtl pcg_fp64(Context *ctx, Matrix &M, int method, int precond, double tolerance,
                          int tol_type, int max_iter, int breakdown_mode, int check_interval,
                          int use_graph)
{
    Workspace &w = M.ws;
    const int vi = spmv_variant(ctx, M, false);
    const int vd = dir_variant(ctx, M, false);
    const PcgBufs bufs{w.x.p, w.b.p, w.r.p, w.p.p, w.p2.p, w.ap.p, w.dinv.p, w.partials.p, w.ctl.p};
    return pcg_fp64_on(ctx, M.m, vi, vd, bufs, method, precond, tolerance, tol_type, max_iter,
                       breakdown_mode, check_interval, [&](auto &enqueue, int chunk) {
                           return run_chunks(ctx, M, 0, ((int64_t)chunk * 64 + vi) * 64 + vd,
                                             use_graph != 0, enqueue,
                                             [](const KrylovCtl &) { return true; });
                       });
}

// Public FP64 solve on device vectors.
void solve_fp64(Context *ctx, Matrix &M, const double *b, double *x,
                const ffh_solver_options &o, ffh_solver_stats &st)
{
    ensure_workspace(ctx, M);
    ensure_val64(ctx, M.m);
    spmv_variant(ctx, M, false);
    Workspace &w = M.ws;
    const int n = M.m.n;
    hipStream_t s = ctx->stream;
    if (ctx->profiling) ctx->prof_begin();
    FFH_HIP(hipEventRecord(ctx->ev_start, s));
    // graphs capture the workspace vectors; copy b/x in and out
    FFH_HIP(hipMemcpyAsync(w.b.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FFH_HIP(hipMemcpyAsync(w.x.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const KrylovCtl out = pcg_fp64(ctx, M, o.method, o.preconditioner, o.tolerance,
                                   o.tolerance_type, o.max_iterations, o.breakdown_mode,
                                   o.check_interval, o.use_graph);
    FFH_HIP(hipMemcpyAsync(x, w.x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_residual(ctx, M); // true residual in FP64
    FFH_HIP(hipEventRecord(ctx->ev_stop, s));
    FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
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
    st.true_residual = std::sqrt(w.host_ctl->true_rr);
    st.initial_residual = out.initial_norm;
}

// Mixed precision: FP32 matrix and FP32 r, p, Ap, D^-1 in the two-launch
// Jacobi-PCG; x and every reduction in FP64.  After every chunk the residual
// b - A x is recomputed in FP64 with the FP64 matrix and replaces the FP32
// one (reliable update; the direction is kept while the recurrence tracked
// it), and it alone decides convergence: ||b - A x|| <= tol ||b - A x0||
// (relative) or <= tol (absolute).  A solve whose FP64 residual stalls
// (kStallRefreshes refreshes without falling below kStallImprove of its
// best squared norm) finishes with the FP64 iteration from its current x.
void solve_mixed(Context *ctx, Matrix &M, const double *b, double *x,
                 const ffh_solver_options &o, ffh_solver_stats &st)
{
    ensure_workspace(ctx, M);
    ensure_workspace32(ctx, M);
    DevMatrix &A = M.m;
    ensure_val64(ctx, A);
    ensure_val32(ctx, A);
    const int vi64 = spmv_variant(ctx, M, false);
    const int vi32 = dir_variant(ctx, M, true);
    const int vi32sp = spmv_variant(ctx, M, true);
    Workspace &w = M.ws;
    const int n = A.n;
    hipStream_t s = ctx->stream;
    if (ctx->profiling) ctx->prof_begin();
    FFH_HIP(hipEventRecord(ctx->ev_start, s));
    FFH_HIP(hipMemcpyAsync(w.b.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FFH_HIP(hipMemcpyAsync(w.x.p, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FFH_HIP(hipMemsetAsync(w.r32.p, 0, sizeof(float) * n, s));
    FFH_HIP(hipMemsetAsync(w.x32.p, 0, sizeof(float) * n, s)); // dx
    const int use = o.method == 1 && o.preconditioner == 1;
    hipLaunchKernelGGL((k_jacobi<float>), dim3(grid_for(ctx, n, kBlock)), dim3(kBlock), 0, s, n,
                       A.gptr(), A.gidx(), A.gval(), use, w.dinv32.p, (float *)nullptr);
    KrylovCtl init{};
    init.max_iter = o.max_iterations;
    init.breakdown_mode = 1;
    init.tol_type = o.tolerance_type;
    init.tolerance = o.tolerance;
    init.mixed = 1;
    *w.host_ctl = init;
    FFH_HIP(hipMemcpyAsync(w.ctl.p, w.host_ctl, sizeof(KrylovCtl), hipMemcpyHostToDevice, s));
    launch_refresh(ctx, M, vi64, 1);
    FFH_HIP(hipGetLastError());
    const int chunk = o.check_interval > 0 ? o.check_interval : 16;
    const int prof = ctx->profiling ? 1 : -1;
    const int gfold = grid_for(ctx, n, kBlock);
    auto enqueue = [&]() {
        // the FP32 iterations accumulate their x steps in dx (w.x32)
        launch_iterations<float, float, float>(ctx, A, w.partials.p, w.ctl.p, chunk, w.x32.p,
                                               w.r32.p, w.p32.p, w.p32b.p, w.ap32.p, w.dinv32.p,
                                               vi32, vi32sp, prof);
        hipLaunchKernelGGL(k_fold_dx, dim3(gfold), dim3(kBlock), 0, s, n, w.x.p, w.x32.p);
        launch_refresh(ctx, M, vi64, 0);
    };
    double best = 1e300;
    int stall = 0;
    bool stagnated = false;
    KrylovCtl out = run_chunks(ctx, M, 1, (((int64_t)chunk * 64 + vi32) * 64 + vi64) * 64 + vi32sp,
                               o.use_graph != 0, enqueue,
                               [&](const KrylovCtl &c) {
                                   if (c.true_rr < kStallImprove * best) {
                                       best = c.true_rr;
                                       stall = 0;
                                   } else if (++stall >= kStallRefreshes) {
                                       stagnated = true;
                                       return false;
                                   }
                                   return true;
                               });
    int iters = out.k2_done;
    const int refreshes = out.refreshes;
    if (stagnated && !out.converged && !out.done) {
        // FP32 iterations stopped making progress: FP64 from the current x
        // to the same absolute tolerance
        const KrylovCtl f = pcg_fp64(ctx, M, o.method, o.preconditioner, out.tol, 1,
                                     std::max(0, o.max_iterations - iters), 1, o.check_interval,
                                     o.use_graph);
        iters += f.iter;
        const double init_norm = out.initial_norm;
        out = f;
        out.initial_norm = init_norm;
        launch_residual(ctx, M);
        FFH_HIP(hipMemcpyAsync(w.host_ctl, w.ctl.p, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
        FFH_HIP(hipStreamSynchronize(s));
        out.true_rr = w.host_ctl->true_rr;
        if (!out.converged && std::sqrt(out.true_rr) <= out.tol) out.converged = 1;
    }
    FFH_HIP(hipMemcpyAsync(x, w.x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FFH_HIP(hipEventRecord(ctx->ev_stop, s));
    FFH_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FFH_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (ctx->profiling) ctx->prof_end();
    st.converged = out.converged;
    st.iterations = iters;
    st.final_residual = std::sqrt(out.true_rr);
    st.true_residual = std::sqrt(out.true_rr);
    st.initial_residual = out.initial_norm;
    st.solve_time = ms * 1e-3;
    st.refinements = refreshes;
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
    const int vi = ready_variant(ctx, A, spmv_variant(ctx, M, f32), f32);
    *bytes = variant_bytes(A, vi, sv);
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

// Bytes one plain SpMV launch moves in the format of the variant in use for
// the precision (tuned first if needed), as variant_bytes.
double spmv_launch_bytes(Context *ctx, Matrix &M, bool f32)
{
    ensure_workspace(ctx, M);
    ensure_val64(ctx, M.m);
    if (f32) ensure_val32(ctx, M.m);
    const int vi = ready_variant(ctx, M.m, spmv_variant(ctx, M, f32), f32);
    return variant_bytes(M.m, vi, f32 ? 4 : 8);
}

// ------------------------------------------------ allocation-free solve --
// FP64 Jacobi-PCG on a caller's device CSC and device b / x with every work
// vector in `scratch`, a fixed array in the library's code object: each GPU
// the library runs on holds its own copy from the moment the code object is
// loaded there until the process ends.  Layout (in doubles): the control
// block and a status word (256 B), the partial sums (4 banks of
// kNoAllocStride), then r, (z, p) a, (z, p) b, A p and D^-1 for up to
// FFH_NOALLOC_MAX_ROWS rows.  Calls on one GPU share it, so calls that
// overlap on one GPU race.
constexpr int64_t kNoAllocHead = 32;     // doubles: KrylovCtl + status word
constexpr int64_t kNoAllocStride = 8192; // partial-sum bank (>= partial_stride)
constexpr int64_t kNoAllocRows = FFH_NOALLOC_MAX_ROWS;
static_assert(kNoAllocRows % 32 == 0, "rows are padded to multiples of 32");
constexpr int64_t kNoAllocDoubles = kNoAllocHead + 4 * kNoAllocStride + 7 * kNoAllocRows;
static_assert(sizeof(KrylovCtl) + sizeof(int) <= kNoAllocHead * sizeof(double),
              "control block does not fit the scratch head");

static __device__ __align__(256) double scratch[kNoAllocDoubles];

// device address of `scratch` per device, looked up once per device with
// that device current (the array's name is only a host-side stand-in)
static std::mutex g_scratch_lock;
static std::vector<double *> g_scratch_ptr;

void scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_lock);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return;
    void *ptr = nullptr;
    FFH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    FFH_REQUIRE(ptr != nullptr, FFH_ERR_DEVICE, "scratch array has no device address");
    g_scratch_ptr[device] = static_cast<double *>(ptr);
}

static double *scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_lock);
    FFH_REQUIRE(device >= 0 && device < (int)g_scratch_ptr.size() && g_scratch_ptr[device],
                FFH_ERR_DEVICE, "scratch array not set up on this device (ffh_init)");
    return g_scratch_ptr[device];
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(int64_t n, double *__restrict__ a)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        a[i] = 0.0;
}

// Checks a 0-based CSC for the products of the allocation-free solve:
// col_ptr(0) = 0, col_ptr(n) = nnz, columns monotone, rows in range and
// strictly ascending in each column, and every entry (i, j) mirrored by an
// entry (j, i) of the same value (A = A^T, so the CSC reads as CSR).  Any
// failure writes 1 to *bad; every read is bounds-checked first.
__global__ __launch_bounds__(kBlock) void k_check_symmetric(int n, int64_t nnz,
                                                             const int *__restrict__ ptr,
                                                             const int *__restrict__ idx,
                                                             const double *__restrict__ val,
                                                             int *__restrict__ bad)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    if (j == 0 && (ptr[0] != 0 || (int64_t)ptr[n] != nnz)) *bad = 1;
    const int b = ptr[j], e = ptr[j + 1];
    if (b < 0 || e < b || (int64_t)e > nnz) {
        *bad = 1;
        return;
    }
    int prev = -1;
    for (int k = b; k < e; ++k) {
        const int i = idx[k];
        if (i <= prev || i >= n) {
            *bad = 1;
            return;
        }
        prev = i;
        const int cb = ptr[i], ce = ptr[i + 1];
        if (cb < 0 || ce < cb || (int64_t)ce > nnz) {
            *bad = 1;
            return;
        }
        int lo = cb, hi = ce;
        while (lo < hi) { // first entry of column i with row >= j
            const int mid = lo + ((hi - lo) >> 1);
            if (idx[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= ce || idx[lo] != j || val[lo] != val[k]) {
            *bad = 1;
            return;
        }
    }
}

// SpMV engine of the allocation-free solve: vector CSR (no row blocks or
// SELL copy to build), lanes per row from the mean row length.
int noalloc_variant(int n, int64_t nnz)
{
    const double avg = n > 0 ? (double)nnz / n : 0.0;
    if (avg <= 6.0) return 3;  // 4 lanes, grid x2
    if (avg <= 16.0) return 5; // 8 lanes, grid x2
    if (avg <= 40.0) return 6; // 16 lanes, grid x2
    return 7;                  // 32 lanes, grid x2
}

void solve_fp64_noalloc(Context *ctx, int n, int64_t nnz, const int *col_ptr,
                        const int *row_idx, const double *val, const double *b, double *x,
                        const ffh_solver_options &o, ffh_solver_stats &st)
{
    FFH_REQUIRE(n >= 1 && n <= kNoAllocRows, FFH_ERR_INVALID,
                "allocation-free solve: n must be 1 .. FFH_NOALLOC_MAX_ROWS");
    FFH_REQUIRE(nnz >= 1 && nnz <= INT32_MAX, FFH_ERR_INVALID,
                "allocation-free solve: nnz must be 1 .. 2^31-1");
    FFH_REQUIRE(partial_stride(ctx) <= kNoAllocStride, FFH_ERR_DEVICE,
                "allocation-free solve: more compute units than the scratch partials cover");
    FFH_HIP(hipSetDevice(ctx->device));
    double *base = scratch_of(ctx->device);
    hipStream_t s = ctx->stream;
    KrylovCtl *ctl = reinterpret_cast<KrylovCtl *>(base);
    int *bad = reinterpret_cast<int *>(reinterpret_cast<char *>(base) + sizeof(KrylovCtl));
    double *partials = base + kNoAllocHead;
    double *v = partials + 4 * kNoAllocStride;
    // each vector starts on a 256-byte boundary (the (z, p) pairs are read
    // as 16-byte words)
    const int64_t nr = ((int64_t)n + 31) & ~(int64_t)31;
    const PcgBufs w{x, const_cast<double *>(b), v, v + nr, v + 3 * nr, v + 5 * nr, v + 6 * nr,
                    partials, ctl};
    // the caller's arrays, borrowed (never freed here)
    struct Borrowed {
        DevMatrix m;
        ~Borrowed()
        {
            m.col_ptr.p = m.row_idx.p = nullptr;
            m.val.p = nullptr;
            m.col_ptr.n = m.row_idx.n = m.val.n = 0;
        }
    } A;
    A.m.n = n;
    A.m.nnz = nnz;
    A.m.symmetric = true;
    A.m.col_ptr.p = const_cast<int *>(col_ptr);
    A.m.row_idx.p = const_cast<int *>(row_idx);
    A.m.val.p = const_cast<double *>(val);
    const int vi = noalloc_variant(n, nnz);

    // what must start at zero: the head (control block, status word) and the
    // partial banks, cleared by a kernel (the array keeps the last call's data)
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ctx, kNoAllocHead + 4 * kNoAllocStride, kBlock)),
                       dim3(kBlock), 0, s, kNoAllocHead + 4 * kNoAllocStride, base);
    hipLaunchKernelGGL(k_check_symmetric, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                       nnz, col_ptr, row_idx, val, bad);
    FFH_HIP(hipGetLastError());
    int h_bad = 0;
    FFH_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    FFH_REQUIRE(h_bad == 0, FFH_ERR_INVALID,
                "allocation-free solve: the CSC must be 0-based, rows ascending in each column, "
                "and exactly symmetric");

    FFH_HIP(hipEventRecord(ctx->ev_start, s));
    KrylovCtl out = pcg_fp64_on(
        ctx, A.m, vi, vi, w, o.method, o.preconditioner, o.tolerance, o.tolerance_type,
        o.max_iterations, o.breakdown_mode, o.check_interval, [&](auto &enqueue, int) {
            // one chunk at a time, the control block read back into host
            // stack memory after each
            KrylovCtl c{};
            do {
                enqueue();
                FFH_HIP(hipMemcpyAsync(&c, ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
                FFH_HIP(hipStreamSynchronize(s));
            } while (!c.done);
            return c;
        });
    launch_residual_on(ctx, A.m, vi, x, b, w.r, partials, ctl);
    FFH_HIP(hipEventRecord(ctx->ev_stop, s));
    KrylovCtl fin{};
    FFH_HIP(hipMemcpyAsync(&fin, ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FFH_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    st.converged = out.converged;
    if (!st.converged && out.residual_norm <= out.tol) st.converged = 1;
    st.iterations = out.iter;
    st.final_residual = out.residual_norm;
    st.solve_time = ms * 1e-3;
    st.refinements = 0;
    st.true_residual = std::sqrt(fin.true_rr);
    st.initial_residual = out.initial_norm;
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

namespace ffh {

// ------------------------------------------------- differentiable products --
// Column form of the reference's loops (fortfem_sparse_matrix.f90:60-126):
// one thread per CSC column j.
//   jvp: y(row) += val(k) x_dot(j) + val_dot(k) x(j)   (val or x_dot may be
//        absent: the A_dot x term of the implicit solve tangent)
//   vjp: val_bar(k) = scale y_bar(row) x(j) ; x_bar(j) = sum_k val(k) y_bar(row)
// The jvp scatters into rows with FP64 atomics (summation order free, as
// for any sum); the vjp gathers per column (deterministic).
// y = A xdot + Adot x row by row through the transposed index (row i's
// entries in ascending column order: column tcol[k], CSC position tpos[k]),
// so every y_i is summed in one fixed order.
__global__ __launch_bounds__(kBlock) void k_csc_jvp(int nrow, const int *__restrict__ tptr,
                                                    const int *__restrict__ tcol,
                                                    const int *__restrict__ tpos,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ vdot,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ xdot,
                                                    double *__restrict__ y)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrow) return;
    double s = 0.0;
    for (int k = tptr[i]; k < tptr[i + 1]; ++k) {
        const int j = tcol[k], p = tpos[k];
        if (val && xdot) s += val[p] * xdot[j];
        if (vdot && x) s += vdot[p] * x[j];
    }
    y[i] = s;
}

__global__ __launch_bounds__(kBlock) void k_csc_vjp(int ncol, const int *__restrict__ ptr,
                                                    const int *__restrict__ idx,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ ybar,
                                                    double scale, double *__restrict__ vbar,
                                                    double *__restrict__ xbar)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= ncol) return;
    const double xj = x[j];
    double s = 0.0;
    for (int k = ptr[j]; k < ptr[j + 1]; ++k) {
        const double yb = ybar[idx[k]];
        if (vbar) vbar[k] = scale * yb * xj;
        s += val[k] * yb;
    }
    if (xbar) xbar[j] = s;
}

// The row-major (transposed) index of a CSC pattern, built on the host once
// per pattern: row pointers, and per entry its column and CSC position.
static void ensure_transpose(Context *ctx, DevMatrix &A)
{
    if (A.tr_built) return;
    hipStream_t s = ctx->stream;
    const int n = A.n;
    const int64_t nnz = A.nnz;
    std::vector<int> cp((size_t)n + 1), ri((size_t)nnz);
    FFH_HIP(hipMemcpyAsync(cp.data(), A.col_ptr.p, sizeof(int) * cp.size(), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipMemcpyAsync(ri.data(), A.row_idx.p, sizeof(int) * ri.size(), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    std::vector<int> tp((size_t)n + 1, 0), tc((size_t)nnz), tq((size_t)nnz);
    for (int64_t k = 0; k < nnz; ++k) {
        FFH_REQUIRE(ri[k] >= 0 && ri[k] < n, FFH_ERR_INVALID, "row index out of range");
        ++tp[(size_t)ri[k] + 1];
    }
    for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
    std::vector<int> next(tp.begin(), tp.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int k = cp[j]; k < cp[j + 1]; ++k) {
            const int at = next[ri[k]]++;
            tc[at] = j;
            tq[at] = k;
        }
    A.tr_ptr.upload(tp, s);
    A.tr_col.upload(tc, s);
    A.tr_pos.upload(tq, s);
    FFH_HIP(hipStreamSynchronize(s));
    A.tr_built = true;
}

void csc_jvp(Context *ctx, DevMatrix &A, const double *vdot, const double *x,
             const double *xdot, double *y)
{
    const int n = A.n;
    ensure_transpose(ctx, A);
    hipLaunchKernelGGL(k_csc_jvp, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, n,
                       A.tr_ptr.p, A.tr_col.p, A.tr_pos.p, xdot ? A.val.p : nullptr, vdot, x, xdot,
                       y);
    FFH_HIP(hipGetLastError());
}

void csc_vjp(Context *ctx, const DevMatrix &A, const double *x, const double *ybar, double scale,
             double *vbar, double *xbar)
{
    const int n = A.n;
    hipLaunchKernelGGL(k_csc_vjp, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, n,
                       A.col_ptr.p, A.row_idx.p, A.val.p, x, ybar, scale, vbar, xbar);
    FFH_HIP(hipGetLastError());
}

} // namespace ffh
