// ffh_internal.hpp -- internal types of libfortfem_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ffh_nedelec.hpp"

#include "../../include/fortfem_hip.h"

namespace ffh {

// Matrix load of the batched SpMV kernels (index and value arrays): an
// ordinary load, so that a batch's matrices stay in the 256 MB memory-side
// cache between iterations (the single-system kernels stream theirs with
// non-temporal loads, which measured faster there).
template <typename T>
__device__ __forceinline__ T ldm(const T *p)
{
    return *p;
}

// Workgroups are dispatched to the 8 XCDs round-robin (XCD = blockIdx % 8)
// and each XCD has its own L2.  The logical block index gives every XCD one
// contiguous range of blocks (so of rows), so the vector entries an XCD's
// rows gather stay in its own L2 instead of every L2 holding all of them.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block()
{
    const int g = gridDim.x, b = blockIdx.x;
    const int per = g / kXcds, rem = g % kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x * per + min(x, rem) + i;
}

// Mixed-precision stagnation rule (single solves and the served batch): an
// FP32-iterating solve whose FP64 true residual has not fallen below 0.9 of
// its best (0.81 in squared norm) for kStallRefreshes consecutive refreshes
// continues in FP64.  Slowly converging systems (a 1e-10 reduction over tens
// of thousands of iterations gains ~1 % per refresh) keep their FP32
// iterations; a residual held at the FP32 operator's accuracy still switches.
constexpr double kStallImprove = 0.81;
constexpr int kStallRefreshes = 32;

// ---------------------------------------------------------------- errors --
struct Error {
    int code;
    std::string msg;
};

#define FFH_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            throw ::ffh::Error{FFH_ERR_DEVICE,                               \
                               std::string(#call) + ": " +                   \
                                   hipGetErrorString(e_)};                   \
    } while (0)

#define FFH_REQUIRE(cond, code, message)                                     \
    do {                                                                     \
        if (!(cond)) throw ::ffh::Error{(code), (message)};                  \
    } while (0)

// Device buffer owned by RAII; hipFree on destruction.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        reset();
        if (count == 0) return;
        FFH_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t s)
    {
        alloc(count);
        if (count)
            FFH_HIP(hipMemcpyAsync(p, h, count * sizeof(T),
                                   hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s)
    {
        upload(h.data(), h.size(), s);
    }
    void swap(DBuf &o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void download(T *h, size_t count, hipStream_t s) const
    {
        if (count)
            FFH_HIP(hipMemcpyAsync(h, p, count * sizeof(T),
                                   hipMemcpyDeviceToHost, s));
    }
};

// Control block of one Krylov solve, resident in device memory.  Each field
// has one writer per launch (block 0 of the launch that owns it), and no
// other block of that launch reads it.
struct KrylovCtl {
    double rzs[2], pap, alpha, beta, rr, rz;
    double residual_norm, initial_norm, tol;
    int iter, max_iter, converged, done, breakdown_hit;
    int breakdown_mode, tol_type, mixed;
    double tolerance;
    double true_rr;  // FP64 ||b - A x||^2 (final report, mixed refreshes)
    int k2_done;     // x/r updates completed (written by the update launch)
    int cur;         // r.z slot of the current iteration (direction launch)
    int restart;     // the next direction is D r (init, refresh, breakdown)
    int nb_rz;       // partial count of the last r.r / r.z producer
    int refreshes;   // mixed: FP64 residual refreshes so far
    int pad0;
};

// ---------------------------------------------------------------- matrix --
// A CSC matrix in HBM (0-based).  Products are row gathers: for an exactly
// symmetric matrix the CSC arrays are read as CSR; otherwise g_* hold the
// transposed copy.
struct DevMatrix {
    int n = 0;
    int64_t nnz = 0;
    DBuf<int> col_ptr, row_idx;
    DBuf<double> val;
    DBuf<float> val32;
    bool symmetric = true;
    DBuf<int> t_ptr, t_idx;
    DBuf<double> t_val;
    DBuf<float> t_val32;
    const int *gptr() const { return symmetric ? col_ptr.p : t_ptr.p; }
    const int *gidx() const { return symmetric ? row_idx.p : t_idx.p; }
    const double *gval() const { return symmetric ? val.p : t_val.p; }
    const float *gval32() const { return symmetric ? val32.p : t_val32.p; }
    bool val32_valid = false;
    bool val64_stale = false; // FP32 assembly: val holds no FP64 view yet
    DBuf<int> rowblocks; // CSR-stream row blocks of the gather arrays
    int nrb = 0;
    int variant[2] = {-1, -1};     // autotuned SpMV variant (FP64, FP32)
    int variant_dir[2] = {-1, -1}; // that of the CG direction launch
    // SELL-64 copy of the gather arrays, built on first use: within each
    // window of 1024 rows the rows are ordered by decreasing length and cut
    // into slices of 64 (sell_row: the row at each slice position, -1 for
    // padding); entry j of a slice's row l sits at sell_off[s] + 64 j + l,
    // each row padded to its slice's longest row with zero values on its own
    // column.  sell_src gives the gather-array position of each entry (-1 for
    // padding); the value copies are refreshed when the values change.
    bool sell_built = false, sell_ok = false;
    bool sell64_valid = false, sell32_valid = false;
    int nslice = 0;
    int64_t sell_total = 0;
    DBuf<int> sell_off, sell_idx, sell_src, sell_row;
    DBuf<double> sell_val;
    DBuf<float> sell_val32;
    // Column tiles of the SELL copy for the batched products (built on first
    // use by the server): per group of four slices (256 rows, one work item)
    // the sorted distinct columns its entries read (tile_cols from
    // tile_off[g]), and per SELL entry its column's position in that list
    // (16 bits), so a block stages each column once in LDS.
    bool tiles_built = false, tiles_ok = false;
    // row-major index of the pattern for the deterministic CSC tangent
    // product (built on first use)
    bool tr_built = false;
    DBuf<int> tr_ptr, tr_col, tr_pos;
    int tile_max = 0; // most distinct columns of one group
    int64_t tile_total = 0; // entries of tile_cols
    DBuf<int> tile_off, tile_cols;
    DBuf<uint16_t> tile_lidx;
    // the row of each slice position, the tile offsets and the tile columns
    // in one array (the tiled row engine's argument)
    DBuf<int> tile_pack;
    // the values (FP64, FP32 or both) were rewritten
    void values_changed()
    {
        val32_valid = false;
        sell64_valid = sell32_valid = false;
    }
};

// Solver workspace attached to a matrix (kept resident between solves).
struct Workspace {
    int n = 0;
    DBuf<double> r, p, p2, ap, x, b, dinv;
    DBuf<float> r32, p32, p32b, ap32, x32, dinv32;
    DBuf<double> partials;
    DBuf<KrylovCtl> ctl;
    KrylovCtl *host_ctl = nullptr; // pinned mailboxes (two, polled alternately)
    hipEvent_t ctl_ev[2] = {nullptr, nullptr};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int64_t graph_chunk[2] = {0, 0};
    ~Workspace();
};

struct Context;

struct Matrix {
    Context *ctx = nullptr;
    DevMatrix m;
    Workspace ws;
    bool borrowed = false; // owned by a problem
};

struct Context {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool profiling = false;
    // launch timing inside solves: every prof_stride-th bracketed launch of
    // each kind is timed (events cost host time; sampling keeps the host
    // ahead of the GPU)
    std::vector<hipEvent_t> prof_ev[2];
    int prof_used[2] = {0, 0};
    int prof_seen[2] = {0, 0};
    bool prof_on[2] = {false, false};
    int prof_stride = 8;
    double prof_avg[2] = {0.0, 0.0};
    int prof_count[2] = {0, 0};
    void prof_begin();
    void prof_mark(int which, bool start);
    void prof_end();
};

// --------------------------------------------------------------- problem --
struct Problem {
    Context *ctx = nullptr;
    int dim = 3;
    int nvert = 0, ncell = 0;
    std::vector<double> vertices; // 3*nv (or 2*nv)
    std::vector<int> cells;       // 4*T (or 3*T), 1-based
    DBuf<double> d_vertices;
    int family = 0, order = 0;
    int n_local = 0;
    int n_global = 0;
    std::vector<int> global_dofs; // n_local*T, 0-based
    // colouring
    int n_colors = 0;
    std::vector<int> color_start; // n_colors+1 positions
    std::vector<int> perm;        // pos -> cell
    DBuf<int> d_perm;
    DBuf<int> d_cells_c;   // colour-sorted cell vertex ids (0-based), nv_cell*T
    DBuf<int> d_gdofs_c;   // [l][pos], 0-based dof ids
    DBuf<int> d_slots;     // [r*n+c][pos] value slots
    // Nedelec: orientation class per cell, classes present (slot order),
    // class slot per colour position, entry slots [pos][c * n + r],
    // class tables (slot x 12 terms x packed pairs), 12 terms per cell
    std::vector<unsigned char> cell_class;
    std::vector<int> class_list;
    DBuf<unsigned char> d_cls_c;
    DBuf<int> d_slots_cm;
    DBuf<double> d_ned64;
    DBuf<float> d_ned32;
    int ned_rule = -1;
    DBuf<double> d_gterm;
    int n_edges = 0, n_faces = 0;
    // Nedelec vector load: oriented load tables per class slot, the load
    // points (x, y, z, w), and the per-cell g_q = det w_q J^-1 f(x_q)
    DBuf<double> d_ned_load_tab, d_ned_load_pts, d_ned_g;
    int ned_load_nq = 0;
    // Nedelec with a tensor at every point of the weighted rule: oriented
    // reference curls per class slot (slot x nq x 3 x ldn, zero padded to
    // ldn = 32 ceil(n / 32)), FP64 and FP32, the rule (x, y, z, w per
    // point) and the uploaded point tensors (9 per point, cell-major)
    DBuf<double> d_ned_ct64, d_ned_pt_pts, d_ptens;
    DBuf<float> d_ned_ct32;
    int ned_pt_nq = 0, ned_pt_ldn = 0;
    bool ned_pt_ready = false;
    // full matrix + reduced matrix
    Matrix full, reduced;
    DBuf<double> rhs_full, rhs_red;
    bool have_dirichlet = false;
    std::vector<int> constrained;  // n_global flags
    std::vector<int> free_dofs;    // 0-based
    DBuf<int> d_free_dofs;
    DBuf<int> d_red_orig;           // reduced entry -> full entry
    DBuf<int> d_lift_ptr, d_lift_k, d_lift_c; // per free row lift entries
    DBuf<double> d_bvals;           // boundary values (n_global)
    int dirichlet_mode = 0;
    // element tables
    DBuf<double> d_tab64;
    DBuf<float> d_tab32;
    int tab_qdeg = -1, tab_nq = 0;
    DBuf<double> d_load_tab; // per-point xi (3) + w*phi (n_local)
    DBuf<double> d_elem;     // element matrices / vectors, [entry][pos]
    // H1 gather assembly.  Column blocks (runs of at most 256 CSC columns
    // whose values fit one block's LDS, gat_blk) are cut into slices of 64
    // columns (gat_bslice: first slice of a block); slice s holds its
    // columns' occurrences (cells e holding the column's dof, local index l,
    // as e * n_local + l, ascending e; -1 pads) step-major in gat_occ
    // [(gat_soff[s] + k) * 64 + lane] and their n_local value positions
    // relative to the block's first value as groups of four in gat_slot
    // [((gat_soff[s] + k) * Q + q) * 64 + lane], Q = ceil(n_local / 4).
    bool gat_ok = false;
    int n_gblk = 0, gat_nz_max = 0;
    DBuf<int> d_gat_occ, d_gat_blk, d_gat_bslice, d_gat_soff;
    DBuf<uint16_t> d_gat_slot;
    DBuf<double> d_h1_terms; // 8 per cell (7 used), cell order
    DBuf<int> d_cells_e;     // 4 vertex ids per cell (0-based), cell order
    int load_qdeg = -1, load_nq = 0;
    DBuf<int> d_err;
    DBuf<double> d_coeff;
    int coeff_kind = -1;
    DBuf<double> x_red, x_full, d_samples;
    bool red_dirty = true;
    bool assembled = false, loaded = false;
    // P1_2D symmetric elimination (mode 1)
    std::vector<int> boundary2d;
};

// host-side symbolic routines (ffh_host.cpp)
void box_mesh(const double bounds[6], const int counts[3], double *vertices,
              int *tets);
void unit_square_mesh(int n, double *vertices, int *tris, int *boundary);
void gauss_legendre01(int n, double *x, double *w);
int tetra_duffy(int p, std::vector<double> &x, std::vector<double> &y,
                std::vector<double> &z, std::vector<double> &w);
int lagrange_dof_count(int d);
void lagrange_indices(int d, std::vector<int> &bary, std::vector<double> &nodes);
void lagrange_eval(int d, const std::vector<int> &bary, const double *pt,
                   double *values, double *grads);
int lagrange_dof_map(int d, int ntet, const int *tets, std::vector<int> &gd,
                     int &count);
void build_pattern(int n_global, int ncell, int nloc,
                   const std::vector<int> &gd, std::vector<int> &col_ptr,
                   std::vector<int> &row_idx);
int color_cells(int n_global, int ncell, int nloc, const std::vector<int> &gd,
                std::vector<int> &color_start, std::vector<int> &perm);
// FP64 view of FP32-assembled values, made on first use
void ensure_val64(Context *ctx, DevMatrix &A);
void build_entry_slots(int ncell, int nloc, const std::vector<int> &gd,
                       const std::vector<int> &perm, const std::vector<int> &color_start,
                       const std::vector<int> &col_ptr, const std::vector<int> &row_idx,
                       std::vector<int> &slots);
void build_slots(int ncell, int nloc, const std::vector<int> &gd,
                 const std::vector<int> &perm, const std::vector<int> &col_ptr,
                 const std::vector<int> &row_idx, std::vector<int> &slots);
// device symbolic phase (ffh_symbolic.hip); tets 1-based on the device
// (eid_out, when given, receives the first-appearance edge numbers, T x 6)
void device_nedelec_dofs(Context *ctx, int k, int64_t T, const int *d_tets, std::vector<int> &gd,
                         int &ndofs, int &nedges, int &nfaces, DBuf<int> *eid_out = nullptr);
void device_lagrange_dofs(Context *ctx, int d, int64_t T, const int *d_tets,
                          const std::vector<int> &bary, std::vector<int> &gd, int &count);
int64_t device_pattern_bytes(int64_t T, int nl);
void device_pattern(Context *ctx, int G, int64_t T, int nl, const int *d_gdc,
                    DBuf<int> &col_ptr, DBuf<int> &row_idx, std::vector<int> &h_col_ptr);
void device_slots_h1(Context *ctx, int64_t T, int nl, const int *d_gdc, const DBuf<int> &col_ptr,
                     const DBuf<int> &row_idx, DBuf<int> &slots);
void device_slots_cm(Context *ctx, int64_t T, int nl, const int *d_gdc, const DBuf<int> &col_ptr,
                     const DBuf<int> &row_idx, int64_t nnz, const std::vector<int> &cstart,
                     DBuf<int> &slots);
// colouring of T cells on the device from their incidence lists (T x m
// entity ids, starting at base; entities 0 .. n_ent - 1): cells colour by
// colour in perm (ascending index inside a colour), offsets in cstart; false
// when a cell would need more than 256 colours
bool device_colour(Context *ctx, int64_t T, int m, const int *d_inc, int base, int n_ent,
                   std::vector<int> &cstart, std::vector<int> &perm);
void lagrange_boundary(int d, int ntet, const std::vector<unsigned char> &bf,
                       const std::vector<int> &gd, int n_global,
                       std::vector<int> &constrained);

} // namespace ffh

// opaque handle of the C-ABI (include/fortfem_hip.h)
struct ffh_context : ffh::Context {};
