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

} // namespace

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

} // namespace ffh
