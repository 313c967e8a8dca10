// ffh_host.cpp -- host-side symbolic stage of libfortfem_hip.so: meshes,
// quadrature tables, the reference's global numbering with O(N) hashing, the
// CSC pattern, the element colouring and the element -> value-slot map.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <exception>
#include <thread>
#include <unordered_map>

#include "ffh_internal.hpp"

namespace ffh {

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &f, int64_t grain)
{
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
    if (n < grain || nt == 1) {
        f(0, n);
        return;
    }
    nt = (int)std::min<int64_t>(nt, n);
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs(nt);
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&f, &errs, t, lo, hi] {
            try {
                f(lo, hi);
            } catch (...) {
                errs[t] = std::current_exception();
            }
        });
    }
    for (auto &x : th) x.join();
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
}

// ------------------------------------------------------------------ meshes
// generate_structured_tetra_box_mesh
// (src/mesh/structured_tetra_box_mesh.f90:12-70): vertex id
// 1 + i + (n1+1)(j + (n2+1)k); six tets per brick sharing corners 1 and 8
// (:84-89); local vertices 3 and 4 swap when det(J) < 0 (:56-68).
void box_mesh(const double bounds[6], const int counts[3], double *vertices,
              int *tets)
{
    const int n1 = counts[0] + 1, n2 = counts[1] + 1;
    auto vid = [&](int i, int j, int k) { return 1 + i + n1 * (j + n2 * k); };
    for (int k = 0; k <= counts[2]; ++k)
        for (int j = 0; j <= counts[1]; ++j)
            for (int i = 0; i <= counts[0]; ++i) {
                int v = vid(i, j, k) - 1;
                const int ijk[3] = {i, j, k};
                for (int c = 0; c < 3; ++c) {
                    double t = (double)ijk[c] / (double)counts[c];
                    vertices[3 * (size_t)v + c] =
                        bounds[c] + t * (bounds[3 + c] - bounds[c]);
                }
            }
    static const int pick[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                   {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    const int64_t nbrick = (int64_t)counts[0] * counts[1] * counts[2];
    parallel_for(nbrick, [&](int64_t lo, int64_t hi) {
        for (int64_t b = lo; b < hi; ++b) {
            int i = (int)(b % counts[0]);
            int j = (int)((b / counts[0]) % counts[1]);
            int k = (int)(b / ((int64_t)counts[0] * counts[1]));
            const int brick[8] = {vid(i, j, k),         vid(i + 1, j, k),
                                  vid(i, j + 1, k),     vid(i + 1, j + 1, k),
                                  vid(i, j, k + 1),     vid(i + 1, j, k + 1),
                                  vid(i, j + 1, k + 1), vid(i + 1, j + 1, k + 1)};
            for (int t = 0; t < 6; ++t) {
                int *c = tets + 4 * (6 * b + t);
                for (int q = 0; q < 4; ++q) c[q] = brick[pick[t][q]];
                double jac[9];
                for (int col = 0; col < 3; ++col)
                    for (int r = 0; r < 3; ++r)
                        jac[r + 3 * col] =
                            vertices[3 * (size_t)(c[col + 1] - 1) + r] -
                            vertices[3 * (size_t)(c[0] - 1) + r];
                double det =
                    jac[0] * (jac[4] * jac[8] - jac[7] * jac[5]) -
                    jac[3] * (jac[1] * jac[8] - jac[7] * jac[2]) +
                    jac[6] * (jac[1] * jac[5] - jac[4] * jac[2]);
                if (det < 0.0) std::swap(c[2], c[3]);
            }
        }
    });
}

// unit_square_mesh -> create_rectangular (src/mesh/mesh_2d.f90:90-148):
// lower triangle (v1, v2, v3), upper (v1, v3, v4); find_boundary (:553-603)
// marks both ends of every edge owned by one triangle, which on this grid are
// exactly the vertices of the outer rows and columns.
void unit_square_mesh(int n, double *v, int *tris, int *boundary)
{
    const double h = 1.0 / (double)(n - 1);
    for (int j = 1, k = 0; j <= n; ++j)
        for (int i = 1; i <= n; ++i, ++k) {
            v[2 * k] = 0.0 + (i - 1) * h;
            v[2 * k + 1] = 0.0 + (j - 1) * h;
            if (boundary) boundary[k] = (i == 1 || i == n || j == 1 || j == n);
        }
    for (int j = 1, k = 0; j < n; ++j)
        for (int i = 1; i < n; ++i) {
            int v1 = (j - 1) * n + i, v2 = v1 + 1, v3 = j * n + i + 1,
                v4 = j * n + i;
            tris[3 * k] = v1, tris[3 * k + 1] = v2, tris[3 * k + 2] = v3, ++k;
            tris[3 * k] = v1, tris[3 * k + 1] = v3, tris[3 * k + 2] = v4, ++k;
        }
}

// -------------------------------------------------------------- quadrature
// Gauss-Legendre on [0, 1] by Newton's method on the three-term recurrence;
// the rule of fortnum gauss_legendre_ab used by
// src/quadrature/tetra_duffy_quadrature.f90:27 (ascending nodes).
void gauss_legendre01(int n, double *x, double *w)
{
    const double pi = 3.14159265358979323846264338327950288;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            pp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pp = n * (z * p1 - p0) / (z * z - 1.0);
        double wt = 2.0 / ((1.0 - z * z) * pp * pp);
        int lo = i, hi = n - 1 - i;
        x[lo] = 0.5 - 0.5 * z;
        x[hi] = 0.5 + 0.5 * z;
        w[lo] = w[hi] = 0.5 * wt;
        if (lo == hi) x[lo] = 0.5;
    }
}

// tetra_duffy_quadrature (src/quadrature/tetra_duffy_quadrature.f90:12-51):
// m = (p+4)/2 points per direction; x = s1, y = (1-s1)s2,
// z = (1-s1)(1-s2)s3, w = w1 w2 w3 (1-s1)^2 (1-s2).
int tetra_duffy(int p, std::vector<double> &x, std::vector<double> &y,
                std::vector<double> &z, std::vector<double> &w)
{
    if (p < 0) return 1;
    int m = (p + 4) / 2;
    std::vector<double> s(m), sw(m);
    gauss_legendre01(m, s.data(), sw.data());
    x.resize(m * m * m), y.resize(m * m * m), z.resize(m * m * m),
        w.resize(m * m * m);
    int q = 0;
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            for (int c = 0; c < m; ++c, ++q) {
                double c1 = 1.0 - s[a], c2 = 1.0 - s[b];
                x[q] = s[a];
                y[q] = c1 * s[b];
                z[q] = c1 * c2 * s[c];
                w[q] = sw[a] * sw[b] * sw[c] * (c1 * c1) * c2;
            }
    return 0;
}

// ------------------------------------------------------------ Lagrange H1
int lagrange_dof_count(int d) { return (d + 1) * (d + 2) * (d + 3) / 6; }

// initialize_tetra_lagrange (src/elements/tetra_lagrange_arbitrary_order.f90:
// 29-66): local order first -> second -> third, fourth = d - rest.
void lagrange_indices(int d, std::vector<int> &bary, std::vector<double> &nodes)
{
    int n = lagrange_dof_count(d);
    bary.assign(4 * n, 0);
    nodes.assign(3 * n, 0.0);
    int k = 0;
    for (int a = 0; a <= d; ++a)
        for (int b = 0; b <= d - a; ++b)
            for (int c = 0; c <= d - a - b; ++c, ++k) {
                bary[4 * k] = a, bary[4 * k + 1] = b, bary[4 * k + 2] = c;
                bary[4 * k + 3] = d - a - b - c;
                nodes[3 * k] = (double)b / d, nodes[3 * k + 1] = (double)c / d;
                nodes[3 * k + 2] = (double)(d - a - b - c) / d;
            }
}

// evaluate_tetra_lagrange (:68-116): products of cardinal factors
// l_i(lambda) = prod_{f<i}(d lambda - f)/i!, lambda = [1-x-y-z, x, y, z].
void lagrange_eval(int d, const std::vector<int> &bary, const double *pt,
                   double *values, double *grads)
{
    int n = lagrange_dof_count(d);
    double lam[4] = {1.0 - (pt[0] + pt[1] + pt[2]), pt[0], pt[1], pt[2]};
    for (int i = 0; i < n; ++i) {
        double f[4], df[4];
        for (int c = 0; c < 4; ++c) {
            int idx = bary[4 * i + c];
            double fact = 1.0;
            for (int t = 2; t <= idx; ++t) fact *= t;
            double v = 1.0, dv = 0.0;
            for (int t = 0; t < idx; ++t) v *= (d * lam[c] - t);
            for (int s = 0; s < idx; ++s) {
                double term = d;
                for (int t = 0; t < idx; ++t)
                    if (t != s) term *= (d * lam[c] - t);
                dv += term;
            }
            f[c] = v / fact;
            df[c] = dv / fact;
        }
        values[i] = f[0] * f[1] * f[2] * f[3];
        double ld[4];
        for (int c = 0; c < 4; ++c) {
            ld[c] = df[c];
            for (int k = 0; k < 4; ++k)
                if (k != c) ld[c] *= f[k];
        }
        for (int k = 0; k < 3; ++k) grads[3 * i + k] = ld[k + 1] - ld[0];
    }
}

// build_tetra_lagrange_dof_map (src/elements/tetra_lagrange_global_dof_map.f90:
// 13-79): key = vertex ids with nonzero barycentric index sorted ascending
// (sort_key :81-99) plus those indices; ids by first appearance over
// (cell, local dof).  The reference's linear search (:62) becomes a hash
// lookup with the same first-appearance order.  Rejects degree outside 1..4
// and cells with repeated vertices (:20-34).  Output is 0-based.
namespace {
struct Key {
    int v[4];
    int w[4];
    bool operator==(const Key &o) const
    {
        return std::memcmp(this, &o, sizeof(Key)) == 0;
    }
};
struct KeyHash {
    size_t operator()(const Key &k) const
    {
        uint64_t h = 0xcbf29ce484222325ull;
        for (int i = 0; i < 4; ++i) {
            h ^= (uint64_t)(uint32_t)k.v[i] * 0x9e3779b97f4a7c15ull + k.w[i];
            h *= 0x100000001b3ull;
            h ^= h >> 31;
        }
        return (size_t)h;
    }
};
} // namespace

int lagrange_dof_map(int d, int ntet, const int *tets, std::vector<int> &gd,
                     int &count)
{
    count = 0;
    if (d < 1 || d > 4) return 1;
    for (int64_t e = 0; e < ntet; ++e)
        for (int a = 0; a < 4; ++a) {
            if (tets[4 * e + a] < 1) return 1;
            for (int b = a + 1; b < 4; ++b)
                if (tets[4 * e + a] == tets[4 * e + b]) return 1;
        }
    std::vector<int> bary;
    std::vector<double> nodes;
    lagrange_indices(d, bary, nodes);
    const int n = lagrange_dof_count(d);
    gd.assign((size_t)n * ntet, -1);
    if (d == 1) {
        // a P1 dof is its vertex: first appearance by direct indexing
        int maxv = 0;
        for (int64_t i = 0; i < (int64_t)4 * ntet; ++i)
            maxv = std::max(maxv, tets[i]);
        std::vector<int> id(maxv + 1, -1);
        for (int64_t e = 0; e < ntet; ++e)
            for (int l = 0; l < n; ++l) {
                int vtx = -1;
                for (int c = 0; c < 4; ++c)
                    if (bary[4 * l + c]) vtx = tets[4 * e + c];
                if (id[vtx] < 0) id[vtx] = count++;
                gd[(size_t)n * e + l] = id[vtx];
            }
        return 0;
    }
    std::unordered_map<Key, int, KeyHash> map;
    map.reserve((size_t)ntet * n / 2 + 16);
    for (int64_t e = 0; e < ntet; ++e)
        for (int l = 0; l < n; ++l) {
            Key k;
            std::memset(&k, 0, sizeof k);
            int m = 0;
            for (int c = 0; c < 4; ++c) {
                if (!bary[4 * l + c]) continue;
                k.v[m] = tets[4 * e + c];
                k.w[m] = bary[4 * l + c];
                ++m;
            }
            // ascending vertex ids (all nonzero here), weights follow
            for (int a = 0; a < m; ++a)
                for (int b = a + 1; b < m; ++b)
                    if (k.v[b] < k.v[a]) {
                        std::swap(k.v[a], k.v[b]);
                        std::swap(k.w[a], k.w[b]);
                    }
            auto it = map.find(k);
            int id;
            if (it == map.end()) {
                id = count++;
                map.emplace(k, id);
            } else {
                id = it->second;
            }
            gd[(size_t)n * e + l] = id;
        }
    return 0;
}

// ------------------------------------------------------------- CSC pattern
// The pattern csc_from_triplet produces from the element triplets of
// assemble_tetra_lagrange_stiffness_csc (:322-333): column j holds every dof
// that shares a cell with j, rows ascending, duplicates merged, structural
// zeros kept.  Symmetric by construction.
void build_pattern(int n_global, int ncell, int nloc, const std::vector<int> &gd,
                   std::vector<int> &col_ptr, std::vector<int> &row_idx)
{
    // dof -> cells
    std::vector<int64_t> cptr(n_global + 1, 0);
    for (size_t i = 0; i < gd.size(); ++i) cptr[gd[i] + 1]++;
    for (int i = 0; i < n_global; ++i) cptr[i + 1] += cptr[i];
    std::vector<int> cells(gd.size());
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t e = 0; e < ncell; ++e)
            for (int l = 0; l < nloc; ++l) {
                int g = gd[(size_t)nloc * e + l];
                cells[fill[g]++] = (int)e;
            }
    }
    std::vector<int> counts(n_global, 0);
    auto column = [&](int j, std::vector<int> &buf) {
        buf.clear();
        for (int64_t k = cptr[j]; k < cptr[j + 1]; ++k) {
            const int *g = gd.data() + (size_t)nloc * cells[k];
            buf.insert(buf.end(), g, g + nloc);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    parallel_for(n_global, [&](int64_t lo, int64_t hi) {
        std::vector<int> buf;
        for (int64_t j = lo; j < hi; ++j) {
            column((int)j, buf);
            counts[j] = (int)buf.size();
        }
    });
    col_ptr.assign(n_global + 1, 0);
    int64_t total = 0;
    for (int j = 0; j < n_global; ++j) {
        col_ptr[j] = (int)total;
        total += counts[j];
    }
    FFH_REQUIRE(total < INT32_MAX, FFH_ERR_INVALID,
                "pattern exceeds int32 nonzeros");
    col_ptr[n_global] = (int)total;
    row_idx.resize(total);
    parallel_for(n_global, [&](int64_t lo, int64_t hi) {
        std::vector<int> buf;
        for (int64_t j = lo; j < hi; ++j) {
            column((int)j, buf);
            std::copy(buf.begin(), buf.end(), row_idx.begin() + col_ptr[j]);
        }
    });
}

// Greedy first-fit colouring in cell order: no two cells of one colour share
// a dof, so each colour scatters into the CSC values without atomics.
// Returns the colour count; perm lists cells colour by colour (ascending
// cell index inside a colour).
int color_cells(int n_global, int ncell, int nloc, const std::vector<int> &gd,
                std::vector<int> &color_start, std::vector<int> &perm)
{
    int words = 1;
    std::vector<uint64_t> used((size_t)n_global * words, 0);
    std::vector<int> color(ncell);
    int ncolors = 0;
    for (int64_t e = 0; e < ncell; ++e) {
        const int *g = gd.data() + (size_t)nloc * e;
        int c = -1;
        for (int w = 0; w < words && c < 0; ++w) {
            uint64_t mask = 0;
            for (int l = 0; l < nloc; ++l) mask |= used[(size_t)g[l] * words + w];
            if (~mask) c = w * 64 + __builtin_ctzll(~mask);
        }
        if (c < 0) {
            // widen the per-dof colour masks by one word
            int nw = words + 1;
            std::vector<uint64_t> wider((size_t)n_global * nw, 0);
            for (int64_t i = 0; i < n_global; ++i)
                for (int w = 0; w < words; ++w)
                    wider[(size_t)i * nw + w] = used[(size_t)i * words + w];
            used.swap(wider);
            c = words * 64;
            words = nw;
        }
        color[e] = c;
        ncolors = std::max(ncolors, c + 1);
        for (int l = 0; l < nloc; ++l)
            used[(size_t)g[l] * words + c / 64] |= (1ull << (c % 64));
    }
    color_start.assign(ncolors + 1, 0);
    for (int64_t e = 0; e < ncell; ++e) color_start[color[e] + 1]++;
    for (int c = 0; c < ncolors; ++c) color_start[c + 1] += color_start[c];
    perm.resize(ncell);
    std::vector<int> fill(color_start.begin(), color_start.end() - 1);
    for (int64_t e = 0; e < ncell; ++e) perm[fill[color[e]]++] = (int)e;
    return ncolors;
}

// Value slot of every element entry (r, c): the position of global row g_r in
// column g_c, stored [r*nloc + c][pos] in colour order so that a wavefront's
// lanes read consecutive slots.
void build_slots(int ncell, int nloc, const std::vector<int> &gd,
                 const std::vector<int> &perm, const std::vector<int> &col_ptr,
                 const std::vector<int> &row_idx, std::vector<int> &slots)
{
    slots.resize((size_t)nloc * nloc * ncell);
    parallel_for(ncell, [&](int64_t lo, int64_t hi) {
        for (int64_t pos = lo; pos < hi; ++pos) {
            const int *g = gd.data() + (size_t)nloc * perm[pos];
            for (int c = 0; c < nloc; ++c) {
                const int *b = row_idx.data() + col_ptr[g[c]];
                const int *e = row_idx.data() + col_ptr[g[c] + 1];
                for (int r = 0; r < nloc; ++r) {
                    const int *f = std::lower_bound(b, e, g[r]);
                    slots[(size_t)(r * nloc + c) * ncell + pos] =
                        (int)(f - row_idx.data());
                }
            }
        }
    });
}

// build_dirichlet_data (src/solvers/fortfem_tetra_lagrange_solver_3d.f90:
// 551-610): the local dofs with barycentric index 0 at the vertex opposite
// a boundary face (:598-606) are constrained.  bf[4 e + o] flags the
// boundary faces (a face contained in exactly one cell, :587-597), found on
// the device.
void lagrange_boundary(int d, int ntet, const std::vector<unsigned char> &bf,
                       const std::vector<int> &gd, int n_global,
                       std::vector<int> &constrained)
{
    std::vector<int> bary;
    std::vector<double> nodes;
    lagrange_indices(d, bary, nodes);
    const int n = lagrange_dof_count(d);
    constrained.assign(n_global, 0);
    for (int64_t e = 0; e < ntet; ++e)
        for (int o = 0; o < 4; ++o) {
            if (!bf[4 * e + o]) continue;
            for (int l = 0; l < n; ++l)
                if (bary[4 * l + o] == 0) constrained[gd[(size_t)n * e + l]] = 1;
        }
}

// Cell-major slots of every element entry: for colour position pos, entry
// t = c * nloc + r (column-major in the element) holds the value slot of
// (row g[r], col g[c]).  Lanes running over r then write rows of one column,
// which the CSC stores next to each other for dofs of one edge / face / cell.
// The first cell in colour order that reaches a slot stores it as ~slot
// (negative): it writes the value instead of adding, so the values need no
// clearing before an assembly.
void build_entry_slots(int ncell, int nloc, const std::vector<int> &gd,
                       const std::vector<int> &perm, const std::vector<int> &color_start,
                       const std::vector<int> &col_ptr, const std::vector<int> &row_idx,
                       std::vector<int> &slots)
{
    const size_t nn = (size_t)nloc * nloc;
    slots.resize(nn * ncell);
    parallel_for(ncell, [&](int64_t lo, int64_t hi) {
        for (int64_t pos = lo; pos < hi; ++pos) {
            const int *g = gd.data() + (size_t)nloc * perm[pos];
            int *o = slots.data() + (size_t)pos * nn;
            for (int c = 0; c < nloc; ++c) {
                const int *b = row_idx.data() + col_ptr[g[c]];
                const int *e = row_idx.data() + col_ptr[g[c] + 1];
                for (int r = 0; r < nloc; ++r) {
                    const int *f = std::lower_bound(b, e, g[r]);
                    FFH_REQUIRE(f != e && *f == g[r], FFH_ERR_INVALID,
                                "pattern misses an element entry");
                    o[(size_t)c * nloc + r] = (int)(f - row_idx.data());
                }
            }
        }
    }, 64);
    // first touch: colours in order; cells of one colour own disjoint slots
    const size_t nnz = row_idx.size();
    std::vector<std::atomic<uint64_t>> seen((nnz + 63) / 64);
    for (auto &w : seen) w.store(0, std::memory_order_relaxed);
    const int ncol = (int)color_start.size() - 1;
    for (int c = 0; c < ncol; ++c) {
        const int p0 = color_start[c], p1 = color_start[c + 1];
        parallel_for(p1 - p0, [&](int64_t lo, int64_t hi) {
            for (int64_t pos = p0 + lo; pos < p0 + hi; ++pos) {
                int *o = slots.data() + (size_t)pos * nn;
                for (size_t t = 0; t < nn; ++t) {
                    const int sl = o[t];
                    const uint64_t bit = 1ull << (sl & 63);
                    if (!(seen[sl >> 6].fetch_or(bit, std::memory_order_relaxed) & bit)) o[t] = ~sl;
                }
            }
        }, 64);
    }
}

} // namespace ffh

