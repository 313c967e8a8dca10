// ffh_capi_tet.cc -- the tetrahedral exports of include/fortfem_tet.h: the
// reference C API's conventions (void functions, status out arguments,
// zero-based indices, positive handles from 32-slot registries,
// src/bindings/fortfem_capi_solver.f90:10-124) over this library's C-ABI.
// Host code only: every device step goes through include/fortfem_hip.h.
#include <algorithm>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/fortfem_hip.h"
#include "../../include/fortfem_tet.h"

namespace {

constexpr int kSlots = 32; // fortfem_capi_solver.f90:10 (max_factors)

struct MeshSlot {
    bool used = false;
    int nv = 0, nt = 0;
    std::vector<double> vertices;
    std::vector<int> tets; // 1-based, the C-ABI's convention
};

struct OperatorSlot {
    bool used = false;
    ffh_problem *prob = nullptr; // assembled from a mesh, or
    ffh_matrix *mat = nullptr;   // uploaded from CSC, or
    ffh_zmatrix *zmat = nullptr; // a complex CSC
    int n = 0;
};

std::mutex g_lock;
ffh_context *g_ctx = nullptr;
MeshSlot g_mesh[kSlots];
OperatorSlot g_op[kSlots];

ffh_context *context(int *status)
{
    if (!g_ctx) {
        const int c = ffh_init(0, &g_ctx);
        if (c != FFH_OK) {
            g_ctx = nullptr;
            *status = c;
            return nullptr;
        }
    }
    return g_ctx;
}

MeshSlot *mesh_of(int handle)
{
    if (handle < 1 || handle > kSlots || !g_mesh[handle - 1].used) return nullptr;
    return &g_mesh[handle - 1];
}

OperatorSlot *op_of(int handle)
{
    if (handle < 1 || handle > kSlots || !g_op[handle - 1].used) return nullptr;
    return &g_op[handle - 1];
}

template <typename S>
int free_slot(S (&slots)[kSlots])
{
    for (int i = 0; i < kSlots; ++i)
        if (!slots[i].used) return i;
    return -1;
}

// a problem on the mesh with the family's pattern and operator assembled
int assemble(ffh_context *ctx, const MeshSlot &m, int family, int order, int qdeg,
             double coeff, double mass, ffh_problem **out)
{
    ffh_problem *p = nullptr;
    int c = ffh_problem_create(ctx, 3, m.nv, m.vertices.data(), m.nt, m.tets.data(), &p);
    if (c != FFH_OK) return c;
    int ng = 0;
    int64_t nnz = 0;
    c = ffh_pattern_build(p, family, order, &ng, &nnz);
    if (c == FFH_OK) c = ffh_assemble(p, qdeg, FFH_COEFF_SCALAR, &coeff, mass, FFH_FP64);
    if (c != FFH_OK) {
        ffh_problem_free(p);
        return c;
    }
    *out = p;
    return FFH_OK;
}

void export_csc(int mesh, int family, int order, int qdeg, double coeff, double mass,
                int nnz_capacity, int *n_dofs, int *nnz, int *col_ptr, int *row_ind,
                double *values, int *status)
{
    std::lock_guard<std::mutex> lk(g_lock);
    *status = -1;
    if (!n_dofs || !nnz) return;
    *n_dofs = 0;
    *nnz = 0;
    MeshSlot *m = mesh_of(mesh);
    if (!m) {
        *status = -3;
        return;
    }
    if (nnz_capacity < 0) return;
    ffh_context *ctx = context(status);
    if (!ctx) return;
    ffh_problem *p = nullptr;
    int c = assemble(ctx, *m, family, order, qdeg, coeff, mass, &p);
    if (c != FFH_OK) {
        *status = c;
        return;
    }
    ffh_matrix *a = nullptr;
    c = ffh_problem_matrix(p, FFH_FULL, &a);
    int n = 0;
    int64_t k = 0;
    if (c == FFH_OK)
        c = ffh_matrix_download(a, 0, nnz_capacity, col_ptr, row_ind, values, &n, &k);
    ffh_problem_free(p);
    if (c == FFH_OK || c == FFH_ERR_CAPACITY) {
        if (k > 0x7fffffff) {
            *status = -1;
            return;
        }
        *n_dofs = n;
        *nnz = (int)k;
    }
    *status = c;
}

void retain_problem(int mesh, int family, int order, int qdeg, double coeff, double mass,
                    int boundary, int *handle, int *n_dofs, int *status)
{
    std::lock_guard<std::mutex> lk(g_lock);
    *status = -1;
    if (!handle || !n_dofs) return;
    *handle = 0;
    *n_dofs = 0;
    MeshSlot *m = mesh_of(mesh);
    if (!m) {
        *status = -3;
        return;
    }
    if (boundary != 0 && boundary != 1) return;
    const int slot = free_slot(g_op);
    if (slot < 0) {
        *status = -3;
        return;
    }
    ffh_context *ctx = context(status);
    if (!ctx) return;
    ffh_problem *p = nullptr;
    int c = assemble(ctx, *m, family, order, qdeg, coeff, mass, &p);
    if (c == FFH_OK && boundary) c = ffh_dirichlet(p, 0, nullptr);
    int ng = 0;
    if (c == FFH_OK) c = ffh_problem_sizes(p, &ng, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (c != FFH_OK) {
        if (p) ffh_problem_free(p);
        *status = c;
        return;
    }
    g_op[slot] = OperatorSlot{true, p, nullptr, nullptr, ng};
    *handle = slot + 1;
    *n_dofs = ng;
    *status = 0;
}

// true when the complex CSC (zero-based, values as (re, im) pairs) equals its
// transpose entry for entry (duplicates summed first)
bool complex_symmetric(int n, int nnz, const int *col_ptr, const int *row_ind,
                       const double *values)
{
    using E = std::tuple<int, int, double, double>;
    std::vector<E> a, t;
    a.reserve(nnz);
    t.reserve(nnz);
    for (int j = 0; j < n; ++j)
        for (int k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            a.emplace_back(j, row_ind[k], values[2 * k], values[2 * k + 1]);
            t.emplace_back(row_ind[k], j, values[2 * k], values[2 * k + 1]);
        }
    auto merged = [](std::vector<E> &v) {
        std::sort(v.begin(), v.end(), [](const E &x, const E &y) {
            return std::tie(std::get<0>(x), std::get<1>(x)) <
                   std::tie(std::get<0>(y), std::get<1>(y));
        });
        std::vector<E> m;
        for (const E &e : v) {
            if (!m.empty() && std::get<0>(m.back()) == std::get<0>(e) &&
                std::get<1>(m.back()) == std::get<1>(e)) {
                std::get<2>(m.back()) += std::get<2>(e);
                std::get<3>(m.back()) += std::get<3>(e);
            } else {
                m.push_back(e);
            }
        }
        return m;
    };
    const std::vector<E> ma = merged(a), mt = merged(t);
    if (ma.size() != mt.size()) return false;
    double scale = 0.0;
    for (const E &e : ma) scale = std::max(scale, std::abs(std::get<2>(e)) + std::abs(std::get<3>(e)));
    const double tol = 1e-12 * scale;
    for (size_t i = 0; i < ma.size(); ++i) {
        if (std::get<0>(ma[i]) != std::get<0>(mt[i]) || std::get<1>(ma[i]) != std::get<1>(mt[i]))
            return false;
        if (std::abs(std::get<2>(ma[i]) - std::get<2>(mt[i])) > tol ||
            std::abs(std::get<3>(ma[i]) - std::get<3>(mt[i])) > tol)
            return false;
    }
    return true;
}

} // namespace

// Every export runs inside a function-try-block: an exception (a failed host
// allocation, say) must not unwind through the C boundary, where it would end
// the host program; it becomes status -4 and the registries stay as they were.
#define FFH_CAPI_CATCH(status)                                                 \
    catch (...)                                                              \
    {                                                                        \
        if (status) *status = -4;                                            \
    }

extern "C" {

void fortfem_tetra_mesh_create(int n_vertices, const double *vertices, int n_tetrahedra,
                               const int *tetrahedra, int *handle, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    *status = -1;
    if (!handle) return;
    *handle = 0;
    if (n_vertices < 4 || n_tetrahedra < 1 || !vertices || !tetrahedra) return;
    for (long long i = 0; i < 4ll * n_tetrahedra; ++i)
        if (tetrahedra[i] < 0 || tetrahedra[i] >= n_vertices) return;
    const int slot = free_slot(g_mesh);
    if (slot < 0) {
        *status = -3;
        return;
    }
    MeshSlot m;
    m.nv = n_vertices;
    m.nt = n_tetrahedra;
    m.vertices.assign(vertices, vertices + 3ll * n_vertices);
    m.tets.resize(4ll * n_tetrahedra);
    for (long long i = 0; i < 4ll * n_tetrahedra; ++i) m.tets[i] = tetrahedra[i] + 1;
    m.used = true;
    g_mesh[slot] = std::move(m);
    *handle = slot + 1;
    *status = 0;
}
FFH_CAPI_CATCH(status)

void fortfem_tetra_mesh_free(int handle, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    MeshSlot *m = mesh_of(handle);
    if (!m) {
        *status = -3;
        return;
    }
    *m = MeshSlot{};
    *status = 0;
}
FFH_CAPI_CATCH(status)

void fortfem_tetra_lagrange_stiffness_csc(int mesh, int degree, int quadrature_degree,
                                          double stiffness_coefficient,
                                          double mass_coefficient, int nnz_capacity,
                                          int *n_dofs, int *nnz, int *col_ptr, int *row_ind,
                                          double *values, int *status)
try
{
    if (!status) return;
    if (degree < 1 || degree > 4 || quadrature_degree < 0) {
        *status = -1;
        return;
    }
    export_csc(mesh, FFH_FAMILY_H1, degree, quadrature_degree, stiffness_coefficient,
               mass_coefficient, nnz_capacity, n_dofs, nnz, col_ptr, row_ind, values, status);
}
FFH_CAPI_CATCH(status)

void fortfem_tetra_nedelec_curl_mass_csc(int mesh, int order, double curl_coefficient,
                                         double mass_coefficient, int nnz_capacity,
                                         int *n_dofs, int *nnz, int *col_ptr, int *row_ind,
                                         double *values, int *status)
try
{
    if (!status) return;
    if (order < 1 || order > 6) {
        *status = -1;
        return;
    }
    export_csc(mesh, FFH_FAMILY_NEDELEC, order, -1, curl_coefficient, mass_coefficient,
               nnz_capacity, n_dofs, nnz, col_ptr, row_ind, values, status);
}
FFH_CAPI_CATCH(status)

void fortfem_tetra_lagrange_operator(int mesh, int degree, int quadrature_degree,
                                     double stiffness_coefficient, double mass_coefficient,
                                     int boundary, int *handle, int *n_dofs, int *status)
try
{
    if (!status) return;
    if (degree < 1 || degree > 4 || quadrature_degree < 0) {
        *status = -1;
        return;
    }
    retain_problem(mesh, FFH_FAMILY_H1, degree, quadrature_degree, stiffness_coefficient,
                   mass_coefficient, boundary, handle, n_dofs, status);
}
FFH_CAPI_CATCH(status)

void fortfem_tetra_nedelec_operator(int mesh, int order, double curl_coefficient,
                                    double mass_coefficient, int boundary, int *handle,
                                    int *n_dofs, int *status)
try
{
    if (!status) return;
    if (order < 1 || order > 6) {
        *status = -1;
        return;
    }
    retain_problem(mesh, FFH_FAMILY_NEDELEC, order, -1, curl_coefficient, mass_coefficient,
                   boundary, handle, n_dofs, status);
}
FFH_CAPI_CATCH(status)

void fortfem_real_operator_csc(int n, int nnz, const int *col_ptr, const int *row_ind,
                               const double *values, int *handle, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    *status = -1;
    if (!handle) return;
    *handle = 0;
    // valid_csc_input (fortfem_capi_solver.f90:126-142): zero-based,
    // monotone, rows in range
    if (n < 1 || nnz < 1 || !col_ptr || !row_ind || !values) return;
    if (col_ptr[0] != 0 || col_ptr[n] != nnz) return;
    for (int j = 0; j < n; ++j)
        if (col_ptr[j + 1] < col_ptr[j]) return;
    for (int k = 0; k < nnz; ++k)
        if (row_ind[k] < 0 || row_ind[k] >= n) return;
    const int slot = free_slot(g_op);
    if (slot < 0) {
        *status = -3;
        return;
    }
    ffh_context *ctx = context(status);
    if (!ctx) return;
    ffh_matrix *a = nullptr;
    const int c = ffh_matrix_upload(ctx, n, nnz, 0, col_ptr, row_ind, values, &a);
    if (c != FFH_OK) {
        *status = c;
        return;
    }
    g_op[slot] = OperatorSlot{true, nullptr, a, nullptr, n};
    *handle = slot + 1;
    *status = 0;
}
FFH_CAPI_CATCH(status)

void fortfem_operator_solve(int handle, int n, double tolerance, int max_iterations,
                            int precision, const double *rhs, double *solution,
                            int *iterations, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    OperatorSlot *op = op_of(handle);
    if (!op) {
        *status = -3;
        return;
    }
    *status = -1;
    if (op->zmat || n != op->n || !rhs || !solution || !(tolerance > 0.0) || max_iterations < 1 ||
        (precision != 0 && precision != 1))
        return;
    ffh_solver_options o;
    ffh_solver_options_default(&o);
    o.method = 1;
    o.preconditioner = 1;
    o.tolerance_type = 0;
    o.tolerance = tolerance;
    o.max_iterations = max_iterations;
    o.precision = precision == 1 ? FFH_MIXED : FFH_FP64;
    o.breakdown_mode = 1;
    ffh_solver_stats st;
    int c;
    if (op->prob) {
        c = ffh_set_rhs(op->prob, FFH_FULL, rhs);
        if (c == FFH_OK) c = ffh_problem_solve(op->prob, &o, solution, &st);
    } else {
        for (int i = 0; i < n; ++i) solution[i] = 0.0;
        c = ffh_solve_sparse(op->mat, rhs, solution, &o, &st);
    }
    if (c != FFH_OK) {
        *status = c;
        return;
    }
    if (iterations) *iterations = st.iterations;
    *status = st.converged ? 0 : 2;
}
FFH_CAPI_CATCH(status)

void fortfem_complex_operator_csc(int n, int nnz, const int *col_ptr, const int *row_ind,
                                  const double *values, int *handle, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    *status = -1;
    if (!handle) return;
    *handle = 0;
    if (n < 1 || nnz < 1 || !col_ptr || !row_ind || !values) return;
    if (col_ptr[0] != 0 || col_ptr[n] != nnz) return;
    for (int j = 0; j < n; ++j)
        if (col_ptr[j + 1] < col_ptr[j]) return;
    for (int k = 0; k < nnz; ++k)
        if (row_ind[k] < 0 || row_ind[k] >= n) return;
    // COCG needs A = A^T (the PML curl-curl operators are complex-symmetric)
    if (!complex_symmetric(n, nnz, col_ptr, row_ind, values)) return;
    const int slot = free_slot(g_op);
    if (slot < 0) {
        *status = -3;
        return;
    }
    ffh_context *ctx = context(status);
    if (!ctx) return;
    ffh_zmatrix *z = nullptr;
    const int c = ffh_zmatrix_upload(ctx, n, nnz, 0, col_ptr, row_ind, values, &z);
    if (c != FFH_OK) {
        *status = c;
        return;
    }
    OperatorSlot o;
    o.used = true;
    o.zmat = z;
    o.n = n;
    g_op[slot] = o;
    *handle = slot + 1;
    *status = 0;
}
FFH_CAPI_CATCH(status)

void fortfem_complex_operator_solve(int handle, int n, double tolerance, int max_iterations,
                                    const double *rhs, double *solution, int *iterations,
                                    int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    OperatorSlot *op = op_of(handle);
    if (!op) {
        *status = -3;
        return;
    }
    *status = -1;
    if (!op->zmat || n != op->n || !rhs || !solution || !(tolerance > 0.0) ||
        max_iterations < 0)
        return;
    int it = 0, conv = 0;
    double rel = 0.0;
    const int c = ffh_zmatrix_solve(op->zmat, rhs, solution, tolerance, max_iterations, &it,
                                    &rel, &conv);
    if (iterations) *iterations = it;
    if (c != FFH_OK) {
        *status = c;
        return;
    }
    *status = conv ? 0 : 2;
}
FFH_CAPI_CATCH(status)

void fortfem_operator_free(int handle, int *status)
try
{
    if (!status) return;
    std::lock_guard<std::mutex> lk(g_lock);
    OperatorSlot *op = op_of(handle);
    if (!op) {
        *status = -3;
        return;
    }
    if (op->prob) ffh_problem_free(op->prob);
    if (op->mat) ffh_matrix_free(op->mat);
    if (op->zmat) ffh_zmatrix_free(op->zmat);
    *op = OperatorSlot{};
    *status = 0;
}
FFH_CAPI_CATCH(status)

} // extern "C"
