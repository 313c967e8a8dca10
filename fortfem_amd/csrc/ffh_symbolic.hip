// Symbolic phase on the device: the CSC pattern of an assembled operator and
// the value slot of every element entry, from the cell dof map.
//
// The pattern is the reference's (csc_from_triplet over the element
// triplets of every cell, duplicates summed: src/assembly/
// assembly_tetra_lagrange_arbitrary_order_3d.f90:276-334,
// assembly_tetra_nedelec_3d.f90:2042-2112): the distinct (row, column)
// pairs, columns ascending, rows ascending within a column.  Here every
// element entry becomes the 64-bit key (column << 32) | row, one radix sort
// orders the keys, a unique pass keeps one key per pair, and a run-start pass
// turns the columns into col_ptr.  Slots are a binary search of the row in
// its column, one thread per element entry.
#include <hipcub/hipcub.hpp>

#include <climits>

#include "ffh_internal.hpp"

namespace ffh {

namespace {

constexpr int kSymBlock = 256;

int sym_grid(int64_t items)
{
    const int64_t g = (items + kSymBlock - 1) / kSymBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

// keys of all element entries; gdc is the coloured dof layout
// gdc[l * T + pos]
__global__ __launch_bounds__(kSymBlock) void k_entry_keys(int64_t T, int nl,
                                                          const int *__restrict__ gdc,
                                                          uint64_t *__restrict__ keys)
{
    const int64_t nn = (int64_t)nl * nl, total = T * nn;
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t pos = t / nn;
        const int rem = (int)(t - pos * nn);
        const int c = rem / nl, r = rem - c * nl;
        const uint32_t col = (uint32_t)gdc[(int64_t)c * T + pos];
        const uint32_t row = (uint32_t)gdc[(int64_t)r * T + pos];
        keys[t] = ((uint64_t)col << 32) | row;
    }
}

// row indices and column starts of the sorted distinct keys
__global__ __launch_bounds__(kSymBlock) void k_pattern_from_keys(int64_t nu,
                                                                 const uint64_t *__restrict__ u,
                                                                 int G, int *__restrict__ row_idx,
                                                                 int *__restrict__ col_ptr)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < nu;
         t += (int64_t)gridDim.x * kSymBlock) {
        const uint64_t k = u[t];
        const int col = (int)(k >> 32);
        row_idx[t] = (int)(uint32_t)k;
        const int prev = t ? (int)(u[t - 1] >> 32) : -1;
        for (int c = prev + 1; c <= col; ++c) col_ptr[c] = (int)t;
        if (t == nu - 1)
            for (int c = col + 1; c <= G; ++c) col_ptr[c] = (int)nu;
    }
}

__device__ __forceinline__ int find_slot(const int *__restrict__ col_ptr,
                                         const int *__restrict__ row_idx, int col, int row)
{
    int lo = col_ptr[col], hi = col_ptr[col + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row_idx[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    return lo; // the pattern holds every element entry
}

// H1 layout: slots[(r * nl + c) * T + pos]
__global__ __launch_bounds__(kSymBlock) void k_slots_h1(int64_t T, int nl,
                                                        const int *__restrict__ gdc,
                                                        const int *__restrict__ col_ptr,
                                                        const int *__restrict__ row_idx,
                                                        int *__restrict__ slots)
{
    const int64_t total = T * nl * nl;
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t e = t / T, pos = t - e * T;
        const int r = (int)(e / nl), c = (int)(e - (int64_t)r * nl);
        slots[t] = find_slot(col_ptr, row_idx, gdc[(int64_t)c * T + pos],
                             gdc[(int64_t)r * T + pos]);
    }
}

__device__ __forceinline__ int colour_of(const int *__restrict__ cstart, int ncol, int pos)
{
    int lo = 0, hi = ncol - 1; // cstart[lo] <= pos < cstart[hi + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Cell-major layout slots[pos * nl * nl + c * nl + r], and the first colour
// that writes each slot
__global__ __launch_bounds__(kSymBlock) void k_slots_cm(int64_t T, int nl,
                                                        const int *__restrict__ gdc,
                                                        const int *__restrict__ col_ptr,
                                                        const int *__restrict__ row_idx,
                                                        const int *__restrict__ cstart, int ncol,
                                                        int *__restrict__ slots,
                                                        int *__restrict__ first_colour)
{
    const int64_t nn = (int64_t)nl * nl, total = T * nn;
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t pos = t / nn;
        const int rem = (int)(t - pos * nn);
        const int c = rem / nl, r = rem - c * nl;
        const int sl = find_slot(col_ptr, row_idx, gdc[(int64_t)c * T + pos],
                                 gdc[(int64_t)r * T + pos]);
        slots[t] = sl;
        atomicMin(first_colour + sl, colour_of(cstart, ncol, (int)pos));
    }
}

// First touch: the entry of the slot's first colour (one cell of a colour
// writes a slot at most once) stores ~slot.
__global__ __launch_bounds__(kSymBlock) void k_mark_first(int64_t T, int nl,
                                                          const int *__restrict__ cstart,
                                                          int ncol,
                                                          const int *__restrict__ first_colour,
                                                          int *__restrict__ slots)
{
    const int64_t nn = (int64_t)nl * nl, total = T * nn;
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int sl = slots[t];
        if (first_colour[sl] == colour_of(cstart, ncol, (int)(t / nn))) slots[t] = ~sl;
    }
}

__global__ __launch_bounds__(kSymBlock) void k_fill_int(int64_t n, int v, int *__restrict__ a)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        a[t] = v;
}

int bits_for(int64_t v)
{
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

// ------------------------------------------------------------- dof maps --
// First-appearance numbering of keyed occurrences: occurrence i (in the
// order cells, then local entities) carries a key; entities are the distinct
// keys, numbered in the order of their first occurrence.  A stable radix
// sort of (key, i) puts each key's first occurrence at the head of its run;
// the heads are flagged in occurrence order and an exclusive scan of the
// flags numbers them.  This is the reference's numbering (its linear search
// and the host's hash give each new key the next number, in the same order).

__global__ __launch_bounds__(kSymBlock) void k_iota(int64_t n, int *__restrict__ a)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        a[t] = (int)t;
}

__global__ __launch_bounds__(kSymBlock) void k_heads(int64_t n, const uint64_t *__restrict__ k,
                                                     int *__restrict__ h)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        h[t] = (t == 0 || k[t] != k[t - 1]) ? 1 : 0;
}

// run g = gid[j] - 1: its first occurrence, flagged in occurrence order
__global__ __launch_bounds__(kSymBlock) void k_first_flags(int64_t n, const int *__restrict__ h,
                                                           const int *__restrict__ gid,
                                                           const int *__restrict__ occ,
                                                           int *__restrict__ first,
                                                           int *__restrict__ flag)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        if (h[t]) {
            first[gid[t] - 1] = occ[t];
            flag[occ[t]] = 1;
        }
}

__global__ __launch_bounds__(kSymBlock) void k_number(int64_t n, const int *__restrict__ gid,
                                                      const int *__restrict__ occ,
                                                      const int *__restrict__ first,
                                                      const int *__restrict__ num,
                                                      int *__restrict__ out)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        out[occ[t]] = num[first[gid[t] - 1]];
}

// number[i] = first-appearance number of keys[i]; returns the count
int first_appearance(Context *ctx, int64_t n, const uint64_t *keys, int end_bit, int *number)
{
    hipStream_t s = ctx->stream;
    if (n == 0) return 0;
    DBuf<uint64_t> ks;
    DBuf<int> occ_in, occ, h, gid, first, flag, num;
    ks.alloc(n), occ_in.alloc(n), occ.alloc(n), h.alloc(n), gid.alloc(n), flag.alloc(n),
        num.alloc(n), first.alloc(n);
    hipLaunchKernelGGL(k_iota, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, occ_in.p);
    size_t b1 = 0, b2 = 0, b3 = 0;
    FFH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, keys, ks.p, occ_in.p, occ.p, n, 0,
                                               end_bit, s));
    FFH_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, b2, h.p, gid.p, n, s));
    FFH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, flag.p, num.p, n, s));
    DBuf<unsigned char> tmp;
    tmp.alloc(std::max<size_t>({b1, b2, b3, (size_t)1}));
    FFH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, b1, keys, ks.p, occ_in.p, occ.p, n, 0,
                                               end_bit, s));
    hipLaunchKernelGGL(k_heads, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, ks.p, h.p);
    FFH_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, b2, h.p, gid.p, n, s));
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, 0, flag.p);
    hipLaunchKernelGGL(k_first_flags, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, h.p, gid.p,
                       occ.p, first.p, flag.p);
    FFH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, b3, flag.p, num.p, n, s));
    hipLaunchKernelGGL(k_number, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, gid.p, occ.p,
                       first.p, num.p, number);
    FFH_HIP(hipGetLastError());
    int count = 0;
    FFH_HIP(hipMemcpyAsync(&count, gid.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    FFH_HIP(hipStreamSynchronize(s));
    return count;
}

// local edges (pairs of local vertices) and faces, in the reference order
__constant__ int c_edge_v[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int c_face_v[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

__device__ __forceinline__ int local_edge(int a, int b) // a != b local positions
{
    const int p = min(a, b), q = max(a, b);
    return p == 0 ? q - 1 : (p == 1 ? q + 1 : 5);
}

// edge keys (min id, max id) over (cell, local edge); tets 1-based
__global__ __launch_bounds__(kSymBlock) void k_edge_keys(int64_t T, const int *__restrict__ tets,
                                                         uint64_t *__restrict__ keys)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < 6 * T;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t e = t / 6;
        const int l = (int)(t - 6 * e);
        const uint32_t a = (uint32_t)tets[4 * e + c_edge_v[l][0]];
        const uint32_t b = (uint32_t)tets[4 * e + c_edge_v[l][1]];
        keys[t] = ((uint64_t)min(a, b) << 32) | max(a, b);
    }
}

// face keys (smallest id, number of the edge of the other two) over
// (cell, local face)
__global__ __launch_bounds__(kSymBlock) void k_face_keys(int64_t T, const int *__restrict__ tets,
                                                         const int *__restrict__ eid,
                                                         uint64_t *__restrict__ keys)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < 4 * T;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t e = t / 4;
        const int f = (int)(t - 4 * e);
        int p[3] = {c_face_v[f][0], c_face_v[f][1], c_face_v[f][2]};
        // the local position of the smallest id first
        int m = 0;
        for (int i = 1; i < 3; ++i)
            if (tets[4 * e + p[i]] < tets[4 * e + p[m]]) m = i;
        const int a = p[m], b = p[(m + 1) % 3], c = p[(m + 2) % 3];
        keys[t] = ((uint64_t)(uint32_t)tets[4 * e + a] << 32) |
                  (uint32_t)eid[6 * e + local_edge(b, c)];
    }
}

__global__ __launch_bounds__(kSymBlock) void k_nedelec_gd(int64_t T, int k,
                                                          const int *__restrict__ eid,
                                                          const int *__restrict__ fid, int ne,
                                                          int nf, int *__restrict__ gd)
{
    const int fdc = k * (k - 1), cdc = k * (k - 1) * (k - 2) / 2;
    const int n = 6 * k + 4 * fdc + cdc;
    const int64_t edc = (int64_t)k * ne;
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n * T;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t e = t / n;
        const int l = (int)(t - e * n);
        int64_t g;
        if (l < 6 * k) {
            g = (int64_t)eid[6 * e + l / k] * k + l % k;
        } else if (l < 6 * k + 4 * fdc) {
            const int q = l - 6 * k;
            g = edc + (int64_t)fid[4 * e + q / fdc] * fdc + q % fdc;
        } else {
            g = edc + (int64_t)nf * fdc + e * cdc + (l - 6 * k - 4 * fdc);
        }
        gd[t] = (int)g;
    }
}

// Lagrange dof keys: support (vertices of nonzero barycentric index) and the
// indices in ascending vertex-id order.  type 0 vertex (its id), 1 edge (its
// number, index at the smaller id), 2 face (its number, indices at the two
// smallest ids), 3 interior (the occurrence itself: never shared).
__global__ __launch_bounds__(kSymBlock) void k_lagrange_keys(int64_t T, int n,
                                                             const int *__restrict__ tets,
                                                             const int *__restrict__ bary,
                                                             const int *__restrict__ eid,
                                                             const int *__restrict__ fid,
                                                             uint64_t *__restrict__ keys)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n * T;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int64_t e = t / n;
        const int l = (int)(t - e * n);
        int pos[4], w[4], m = 0;
        for (int c = 0; c < 4; ++c)
            if (bary[4 * l + c]) {
                pos[m] = c;
                w[m] = bary[4 * l + c];
                ++m;
            }
        // ascending vertex id
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b)
                if (tets[4 * e + pos[b]] < tets[4 * e + pos[a]]) {
                    int x = pos[a];
                    pos[a] = pos[b];
                    pos[b] = x;
                    x = w[a];
                    w[a] = w[b];
                    w[b] = x;
                }
        uint64_t type, id, wa = 0, wb = 0;
        if (m == 1) {
            type = 0;
            id = (uint64_t)(tets[4 * e + pos[0]] - 1);
        } else if (m == 2) {
            type = 1;
            id = (uint64_t)eid[6 * e + local_edge(pos[0], pos[1])];
            wa = (uint64_t)w[0];
        } else if (m == 3) {
            type = 2;
            const int missing = 6 - pos[0] - pos[1] - pos[2];
            // local face opposite `missing` in the reference order
            const int f = 3 - missing;
            id = (uint64_t)fid[4 * e + f];
            wa = (uint64_t)w[0];
            wb = (uint64_t)w[1];
        } else {
            type = 3;
            id = (uint64_t)t;
        }
        keys[t] = (type << 62) | (id << 8) | (wa << 4) | wb;
    }
}

} // namespace

// First-appearance edge and face numbers of a tetrahedral mesh (tets
// 1-based on the device): eid (T x 6), fid (T x 4).
static void edge_face_numbers(Context *ctx, int64_t T, const int *d_tets, DBuf<int> &eid,
                              DBuf<int> &fid, int &ne, int &nf)
{
    hipStream_t s = ctx->stream;
    DBuf<uint64_t> keys;
    keys.alloc(6 * T);
    eid.alloc(6 * T);
    fid.alloc(4 * T);
    hipLaunchKernelGGL(k_edge_keys, dim3(sym_grid(6 * T)), dim3(kSymBlock), 0, s, T, d_tets,
                       keys.p);
    ne = first_appearance(ctx, 6 * T, keys.p, 64, eid.p);
    hipLaunchKernelGGL(k_face_keys, dim3(sym_grid(4 * T)), dim3(kSymBlock), 0, s, T, d_tets,
                       eid.p, keys.p);
    nf = first_appearance(ctx, 4 * T, keys.p, 64, fid.p);
}

void device_nedelec_dofs(Context *ctx, int k, int64_t T, const int *d_tets, std::vector<int> &gd,
                         int &ndofs, int &nedges, int &nfaces, DBuf<int> *eid_out)
{
    hipStream_t s = ctx->stream;
    DBuf<int> eid, fid, dgd;
    edge_face_numbers(ctx, T, d_tets, eid, fid, nedges, nfaces);
    const int fdc = k * (k - 1), cdc = k * (k - 1) * (k - 2) / 2;
    const int n = 6 * k + 4 * fdc + cdc;
    const int64_t total = (int64_t)k * nedges + (int64_t)nfaces * fdc + T * cdc;
    FFH_REQUIRE(total < INT_MAX, FFH_ERR_INVALID, "Nedelec dof count exceeds int32");
    ndofs = (int)total;
    dgd.alloc(n * T);
    hipLaunchKernelGGL(k_nedelec_gd, dim3(sym_grid(n * T)), dim3(kSymBlock), 0, s, T, k, eid.p,
                       fid.p, nedges, nfaces, dgd.p);
    FFH_HIP(hipGetLastError());
    gd.resize((size_t)n * T);
    dgd.download(gd.data(), gd.size(), s);
    FFH_HIP(hipStreamSynchronize(s));
    if (eid_out) eid_out->swap(eid);
}

void device_lagrange_dofs(Context *ctx, int d, int64_t T, const int *d_tets,
                          const std::vector<int> &bary, std::vector<int> &gd, int &count)
{
    hipStream_t s = ctx->stream;
    const int n = (int)bary.size() / 4;
    DBuf<int> eid, fid, dbary, dgd;
    int ne = 0, nf = 0;
    if (d >= 2) edge_face_numbers(ctx, T, d_tets, eid, fid, ne, nf);
    dbary.upload(bary, s);
    DBuf<uint64_t> keys;
    keys.alloc(n * T);
    dgd.alloc(n * T);
    hipLaunchKernelGGL(k_lagrange_keys, dim3(sym_grid(n * T)), dim3(kSymBlock), 0, s, T, n,
                       d_tets, dbary.p, eid.p, fid.p, keys.p);
    FFH_HIP(hipGetLastError());
    count = first_appearance(ctx, n * T, keys.p, 64, dgd.p);
    gd.resize((size_t)n * T);
    dgd.download(gd.data(), gd.size(), s);
    FFH_HIP(hipStreamSynchronize(s));
}

int64_t device_pattern_bytes(int64_t T, int nl)
{
    return T * nl * nl * (int64_t)(2 * sizeof(uint64_t));
}

// Pattern of G dofs from the coloured dof layout d_gdc (T x nl); col_ptr is
// also returned on the host.
void device_pattern(Context *ctx, int G, int64_t T, int nl, const int *d_gdc,
                    DBuf<int> &col_ptr, DBuf<int> &row_idx, std::vector<int> &h_col_ptr)
{
    hipStream_t s = ctx->stream;
    const int64_t N = T * nl * nl;
    DBuf<uint64_t> keys, sorted;
    keys.alloc(N);
    sorted.alloc(N);
    hipLaunchKernelGGL(k_entry_keys, dim3(sym_grid(N)), dim3(kSymBlock), 0, s, T, nl, d_gdc,
                       keys.p);
    FFH_HIP(hipGetLastError());
    const int end_bit = 32 + bits_for(G);
    size_t tmp_sort = 0, tmp_uniq = 0;
    DBuf<int64_t> nsel;
    nsel.alloc(1);
    FFH_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_sort, keys.p, sorted.p, N, 0, end_bit,
                                              s));
    FFH_HIP(hipcub::DeviceSelect::Unique(nullptr, tmp_uniq, sorted.p, keys.p, nsel.p, N, s));
    DBuf<unsigned char> tmp;
    tmp.alloc(std::max<size_t>(std::max(tmp_sort, tmp_uniq), 1));
    FFH_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_sort, keys.p, sorted.p, N, 0, end_bit, s));
    FFH_HIP(hipcub::DeviceSelect::Unique(tmp.p, tmp_uniq, sorted.p, keys.p, nsel.p, N, s));
    int64_t nu = 0;
    nsel.download(&nu, 1, s);
    FFH_HIP(hipStreamSynchronize(s));
    FFH_REQUIRE(nu < INT_MAX, FFH_ERR_INVALID, "pattern exceeds int32 nonzeros");
    row_idx.alloc(nu);
    col_ptr.alloc((size_t)G + 1);
    hipLaunchKernelGGL(k_pattern_from_keys, dim3(sym_grid(nu)), dim3(kSymBlock), 0, s, nu, keys.p,
                       G, row_idx.p, col_ptr.p);
    FFH_HIP(hipGetLastError());
    h_col_ptr.resize((size_t)G + 1);
    col_ptr.download(h_col_ptr.data(), (size_t)G + 1, s);
    FFH_HIP(hipStreamSynchronize(s));
}

void device_slots_h1(Context *ctx, int64_t T, int nl, const int *d_gdc, const DBuf<int> &col_ptr,
                     const DBuf<int> &row_idx, DBuf<int> &slots)
{
    const int64_t N = T * nl * nl;
    slots.alloc(N);
    hipLaunchKernelGGL(k_slots_h1, dim3(sym_grid(N)), dim3(kSymBlock), 0, ctx->stream, T, nl,
                       d_gdc, col_ptr.p, row_idx.p, slots.p);
    FFH_HIP(hipGetLastError());
}

void device_slots_cm(Context *ctx, int64_t T, int nl, const int *d_gdc, const DBuf<int> &col_ptr,
                     const DBuf<int> &row_idx, int64_t nnz, const std::vector<int> &cstart,
                     DBuf<int> &slots)
{
    hipStream_t s = ctx->stream;
    const int64_t N = T * nl * nl;
    const int ncol = (int)cstart.size() - 1;
    DBuf<int> dcs, first;
    dcs.upload(cstart, s);
    first.alloc(std::max<int64_t>(nnz, 1));
    slots.alloc(N);
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(nnz)), dim3(kSymBlock), 0, s, nnz, INT_MAX,
                       first.p);
    hipLaunchKernelGGL(k_slots_cm, dim3(sym_grid(N)), dim3(kSymBlock), 0, s, T, nl, d_gdc,
                       col_ptr.p, row_idx.p, dcs.p, ncol, slots.p, first.p);
    hipLaunchKernelGGL(k_mark_first, dim3(sym_grid(N)), dim3(kSymBlock), 0, s, T, nl, dcs.p, ncol,
                       first.p, slots.p);
    FFH_HIP(hipGetLastError());
    FFH_HIP(hipStreamSynchronize(s)); // dcs and first are freed on return
}

// ------------------------------------------------------------ colouring --
// Cell colouring on the device.  Two cells conflict when they share an
// entity of the incidence lists: vertices for the Lagrange and 2D families
// (every dof they can share lies on a shared vertex, edge or face, whose
// vertices are then shared too) and edges for Nedelec (its shared dofs lie on
// shared edges or faces).  No two cells of one colour conflict, so each
// colour scatters into the CSC values without atomics.
//
// Jones-Plassmann rounds with first-fit colours (M. T. Jones and P. E.
// Plassmann, "A parallel graph coloring heuristic", SIAM J. Sci. Comput. 14,
// 1993): an uncoloured cell takes the smallest colour its coloured
// neighbours do not hold once every uncoloured neighbour has a lower
// priority.  Priorities are a 32-bit integer hash of the cell index with the
// index as tie-break, so the colouring is the same on every run.  A cell
// coloured in round k stores (k << 16) | colour, and the cells of round k
// read such an entry as unset: each round sees the colours of the rounds
// before it only, whatever the order the threads run in.
namespace {

constexpr int kMaxColours = 256;
constexpr int kColourRoundsCap = 16384;

// lowbias32 integer hash (C. Wellons, "Prospecting for hash functions",
// 2018) of the cell index, the index in the low word
__device__ __forceinline__ uint64_t jp_priority(uint32_t e)
{
    uint32_t x = e;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return ((uint64_t)x << 32) | e;
}

__global__ __launch_bounds__(kSymBlock) void k_inc_count(int64_t n, const int *__restrict__ inc,
                                                         int base, int *__restrict__ cnt)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock)
        atomicAdd(cnt + (inc[t] - base), 1);
}

// entity -> cells lists (order inside a list is immaterial to the rounds)
__global__ __launch_bounds__(kSymBlock) void k_inc_fill(int64_t n, int m,
                                                        const int *__restrict__ inc, int base,
                                                        const int *__restrict__ ptr,
                                                        int *__restrict__ cur,
                                                        int *__restrict__ list)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kSymBlock) {
        const int ent = inc[t] - base;
        const int q = atomicAdd(cur + ent, 1);
        list[ptr[ent] + q] = (int)(t / m);
    }
}

// One round, one thread per cell.  stat[0] += cells coloured, stat[1] =
// max(stat[1], colour), stat[2] = 1 if a cell found all kMaxColours colours
// taken.  A cell reads its neighbours eight at a time (list entries, then
// their colours), so each thread keeps several loads in flight.
constexpr int kMaxInc = 6; // entities per cell (tetrahedron edges)

__global__ __launch_bounds__(kSymBlock) void k_jp_round(int64_t T, int m,
                                                        const int *__restrict__ inc, int base,
                                                        const int *__restrict__ ptr,
                                                        const int *__restrict__ list,
                                                        int *__restrict__ col, int round,
                                                        int *__restrict__ stat)
{
    const int64_t e = blockIdx.x * (int64_t)kSymBlock + threadIdx.x;
    int done = 0, cmax = -1, over = 0;
    if (e < T && col[e] < 0) {
        const uint64_t ke = jp_priority((uint32_t)e);
        int q0[kMaxInc], q1[kMaxInc];
#pragma unroll
        for (int k = 0; k < kMaxInc; ++k)
            if (k < m) {
                const int ent = inc[e * m + k] - base;
                q0[k] = ptr[ent];
                q1[k] = ptr[ent + 1];
            }
        uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        bool ready = true;
#pragma unroll
        for (int k = 0; k < kMaxInc; ++k) {
            if (k >= m || !ready) continue;
            for (int q = q0[k]; q < q1[k] && ready; q += 8) {
                int f[8], cf[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) f[u] = q + u < q1[k] ? list[q + u] : -1;
#pragma unroll
                for (int u = 0; u < 8; ++u) cf[u] = f[u] >= 0 ? col[f[u]] : 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (f[u] < 0 || f[u] == (int)e) continue;
                    if (cf[u] < 0 || (cf[u] >> 16) == round) {
                        if (jp_priority((uint32_t)f[u]) > ke) ready = false;
                    } else {
                        const int c = cf[u] & 0xffff; // < kMaxColours
                        const uint64_t bit = 1ull << (c & 63);
                        const int w = c >> 6;
                        m0 |= w == 0 ? bit : 0ull;
                        m1 |= w == 1 ? bit : 0ull;
                        m2 |= w == 2 ? bit : 0ull;
                        m3 |= w == 3 ? bit : 0ull;
                    }
                }
            }
        }
        if (ready) {
            int c = -1;
            if (~m0) c = __builtin_ctzll(~m0);
            else if (~m1) c = 64 + __builtin_ctzll(~m1);
            else if (~m2) c = 128 + __builtin_ctzll(~m2);
            else if (~m3) c = 192 + __builtin_ctzll(~m3);
            if (c < 0) {
                over = 1;
            } else {
                col[e] = (round << 16) | c;
                done = 1;
                cmax = c;
            }
        }
    }
    __shared__ int s_done[kSymBlock / 64], s_max[kSymBlock / 64], s_over[kSymBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        done += __shfl_down(done, o, 64);
        cmax = max(cmax, __shfl_down(cmax, o, 64));
        over |= __shfl_down(over, o, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_done[wv] = done;
        s_max[wv] = cmax;
        s_over[wv] = over;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int d = 0, mx = -1, ov = 0;
        for (int w = 0; w < kSymBlock / 64; ++w) {
            d += s_done[w];
            mx = max(mx, s_max[w]);
            ov |= s_over[w];
        }
        if (d) atomicAdd(stat, d);
        if (mx >= 0) atomicMax(stat + 1, mx);
        if (ov) atomicOr(stat + 2, 1);
    }
}

__global__ __launch_bounds__(kSymBlock) void k_colour_keys(int64_t T, const int *__restrict__ col,
                                                           unsigned *__restrict__ key)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * kSymBlock)
        key[t] = (unsigned)(col[t] & 0xffff);
}

// cstart[c] = first position of colour c in the sorted keys (colours are
// 0 .. C-1 without gaps: first fit takes colour c only next to a c - 1)
__global__ __launch_bounds__(kSymBlock) void k_colour_starts(int64_t T,
                                                             const unsigned *__restrict__ sk,
                                                             int *__restrict__ cstart)
{
    for (int64_t t = blockIdx.x * (int64_t)kSymBlock + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * kSymBlock)
        if (t == 0 || sk[t] != sk[t - 1]) cstart[sk[t]] = (int)t;
}

} // namespace

// Colours of T cells from their incidence lists d_inc (T x m entity ids,
// base-based, entities 0 .. n_ent - 1): perm lists the cells colour by
// colour (ascending cell index inside a colour), cstart the colour offsets.
// False (nothing returned) when a cell would need more than kMaxColours
// colours.
bool device_colour(Context *ctx, int64_t T, int m, const int *d_inc, int base, int n_ent,
                   std::vector<int> &cstart, std::vector<int> &perm)
{
    hipStream_t s = ctx->stream;
    if (T == 0) {
        cstart.assign(1, 0);
        perm.clear();
        return true;
    }
    FFH_REQUIRE(m >= 1 && m <= kMaxInc, FFH_ERR_INVALID, "colouring: 1 to 6 entities per cell");
    FFH_REQUIRE(T * m < INT_MAX, FFH_ERR_INVALID, "colouring: incidence lists exceed int32");
    const int64_t n = T * m;
    DBuf<int> cnt, ptr, cur, list, col, stat;
    cnt.alloc(n_ent + 1);
    ptr.alloc(n_ent + 1);
    cur.alloc(n_ent);
    list.alloc(n);
    col.alloc(T);
    stat.alloc(3 * (size_t)kColourRoundsCap);
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(n_ent + 1)), dim3(kSymBlock), 0, s,
                       (int64_t)n_ent + 1, 0, cnt.p);
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(n_ent)), dim3(kSymBlock), 0, s, (int64_t)n_ent,
                       0, cur.p);
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(T)), dim3(kSymBlock), 0, s, T, -1, col.p);
    hipLaunchKernelGGL(k_fill_int, dim3(sym_grid(3 * kColourRoundsCap)), dim3(kSymBlock), 0, s,
                       (int64_t)3 * kColourRoundsCap, 0, stat.p);
    hipLaunchKernelGGL(k_inc_count, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, d_inc, base,
                       cnt.p);
    size_t bs = 0;
    FFH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bs, cnt.p, ptr.p, n_ent + 1, s));
    {
        DBuf<unsigned char> tmp;
        tmp.alloc(std::max<size_t>(bs, 1));
        FFH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bs, cnt.p, ptr.p, n_ent + 1, s));
    }
    hipLaunchKernelGGL(k_inc_fill, dim3(sym_grid(n)), dim3(kSymBlock), 0, s, n, m, d_inc, base,
                       ptr.p, cur.p, list.p);
    FFH_HIP(hipGetLastError());
    // rounds in batches of 8 between checks of the coloured count
    FFH_REQUIRE((T + kSymBlock - 1) / kSymBlock <= INT_MAX, FFH_ERR_INVALID, "colouring: too many cells");
    const int grid = (int)((T + kSymBlock - 1) / kSymBlock);
    int64_t coloured = 0;
    int ncol = 0, round = 0;
    std::vector<int> h(3 * 8);
    while (coloured < T) {
        FFH_REQUIRE(round + 8 <= kColourRoundsCap, FFH_ERR_DEVICE, "colouring did not finish");
        for (int r = 0; r < 8; ++r)
            hipLaunchKernelGGL(k_jp_round, dim3(grid), dim3(kSymBlock), 0, s, T, m, d_inc, base,
                               ptr.p, list.p, col.p, round + r, stat.p + 3 * (round + r));
        FFH_HIP(hipGetLastError());
        FFH_HIP(hipMemcpyAsync(h.data(), stat.p + 3 * round, sizeof(int) * h.size(),
                               hipMemcpyDeviceToHost, s));
        FFH_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < 8; ++r) {
            if (h[3 * r + 2]) return false;
            coloured += h[3 * r];
            ncol = std::max(ncol, h[3 * r + 1] + 1);
        }
        round += 8;
    }
    // cells sorted by colour (stable: ascending index inside a colour)
    DBuf<unsigned> key, skey;
    DBuf<int> idx, sidx, dstart;
    key.alloc(T);
    skey.alloc(T);
    idx.alloc(T);
    sidx.alloc(T);
    dstart.alloc(ncol);
    hipLaunchKernelGGL(k_colour_keys, dim3(sym_grid(T)), dim3(kSymBlock), 0, s, T, col.p, key.p);
    hipLaunchKernelGGL(k_iota, dim3(sym_grid(T)), dim3(kSymBlock), 0, s, T, idx.p);
    const int end_bit = bits_for(ncol);
    bs = 0;
    FFH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bs, key.p, skey.p, idx.p, sidx.p, T, 0,
                                               end_bit, s));
    {
        DBuf<unsigned char> tmp;
        tmp.alloc(std::max<size_t>(bs, 1));
        FFH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bs, key.p, skey.p, idx.p, sidx.p, T, 0,
                                                   end_bit, s));
        hipLaunchKernelGGL(k_colour_starts, dim3(sym_grid(T)), dim3(kSymBlock), 0, s, T, skey.p,
                           dstart.p);
        FFH_HIP(hipGetLastError());
        perm.resize(T);
        cstart.resize(ncol + 1);
        sidx.download(perm.data(), T, s);
        dstart.download(cstart.data(), ncol, s);
        FFH_HIP(hipStreamSynchronize(s));
    }
    cstart[ncol] = (int)T;
    return true;
}

} // namespace ffh
