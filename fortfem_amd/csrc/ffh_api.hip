// ffh_api.hip -- the extern "C" entry points of libfortfem_hip.so.  Each one
// validates its arguments, converts exceptions into the reference C-API's
// status codes and records a message for ffh_last_error().
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ffh_internal.hpp"
#include "ffh_krylov.hpp"

using namespace ffh;

struct ffh_problem : Problem {};
struct ffh_matrix : Matrix {};
struct ffh_zmatrix {
    Context *ctx = nullptr;
    ZMatrix *m = nullptr;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guarded(Context *ctx, F &&f)
{
    try {
        f();
        return FFH_OK;
    } catch (const Error &e) {
        if (ctx) ctx->err = e.msg;
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        g_err = "host allocation failed";
        return FFH_ERR_DEVICE;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        g_err = e.what();
        return FFH_ERR_INVALID;
    } catch (...) {
        if (ctx) ctx->err = "unexpected error";
        g_err = "unexpected error";
        return FFH_ERR_DEVICE;
    }
}

// CSC (base) -> 0-based device CSC; transposed copy if not exactly symmetric
void load_matrix(Context *ctx, DevMatrix &A, int n, int64_t nnz, int base,
                 const int *col_ptr, const int *row_idx, const double *val)
{
    FFH_REQUIRE(n >= 0 && nnz >= 0 && nnz < INT32_MAX, FFH_ERR_INVALID, "bad matrix sizes");
    FFH_REQUIRE(base == 0 || base == 1, FFH_ERR_INVALID, "base must be 0 or 1");
    std::vector<int> cp(n + 1), ri(nnz);
    for (int j = 0; j <= n; ++j) cp[j] = col_ptr[j] - base;
    FFH_REQUIRE(cp[0] == 0 && cp[n] == nnz, FFH_ERR_INVALID, "col_ptr does not span nnz");
    for (int j = 0; j < n; ++j)
        FFH_REQUIRE(cp[j + 1] >= cp[j], FFH_ERR_INVALID, "col_ptr not monotone");
    for (int64_t k = 0; k < nnz; ++k) {
        ri[k] = row_idx[k] - base;
        FFH_REQUIRE(ri[k] >= 0 && ri[k] < n, FFH_ERR_INVALID, "row index out of range");
    }
    // exact symmetry test: sorted (row, value) of each column against the
    // transpose
    std::vector<int> tp(n + 1, 0), ti(nnz);
    std::vector<double> tv(nnz);
    for (int64_t k = 0; k < nnz; ++k) tp[ri[k] + 1]++;
    for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
    {
        std::vector<int> fill(tp.begin(), tp.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int k = cp[j]; k < cp[j + 1]; ++k) {
                int o = fill[ri[k]]++;
                ti[o] = j;
                tv[o] = val[k];
            }
    }
    bool sym = true;
    for (int j = 0; j < n && sym; ++j) {
        if (cp[j + 1] - cp[j] != tp[j + 1] - tp[j]) {
            sym = false;
            break;
        }
        for (int k = cp[j], o = tp[j]; k < cp[j + 1]; ++k, ++o)
            if (ri[k] != ti[o] || val[k] != tv[o]) {
                sym = false;
                break;
            }
    }
    hipStream_t s = ctx->stream;
    A.n = n;
    A.nnz = nnz;
    A.col_ptr.upload(cp, s);
    A.row_idx.upload(ri, s);
    A.val.upload(val, nnz, s);
    A.symmetric = sym;
    A.values_changed();
    if (!sym) {
        A.t_ptr.upload(tp, s);
        A.t_idx.upload(ti, s);
        A.t_val.upload(tv, s);
    }
    build_rowblocks(ctx, A, sym ? cp : tp);
    FFH_HIP(hipStreamSynchronize(s));
}

void check_options(const ffh_solver_options *o)
{
    FFH_REQUIRE(o, FFH_ERR_INVALID, "options must not be NULL");
    FFH_REQUIRE(o->method == 0 || o->method == 1, FFH_ERR_INVALID, "method must be cg or pcg");
    FFH_REQUIRE(o->preconditioner == 0 || o->preconditioner == 1, FFH_ERR_INVALID,
                "preconditioner must be none or jacobi");
    FFH_REQUIRE(o->tolerance_type == 0 || o->tolerance_type == 1, FFH_ERR_INVALID,
                "tolerance_type must be relative or absolute");
    FFH_REQUIRE(o->precision == FFH_FP64 || o->precision == FFH_MIXED, FFH_ERR_INVALID,
                "precision must be FP64 or MIXED");
    FFH_REQUIRE(o->breakdown_mode == 0 || o->breakdown_mode == 1, FFH_ERR_INVALID,
                "breakdown_mode must be 0 or 1");
    FFH_REQUIRE(o->check_interval >= 0, FFH_ERR_INVALID, "check_interval must be >= 0");
}
} // namespace

void Context::prof_begin()
{
    prof_used[0] = prof_used[1] = 0;
    prof_seen[0] = prof_seen[1] = 0;
}

void Context::prof_mark(int which, bool start)
{
    if (start) prof_on[which] = prof_seen[which]++ % prof_stride == 0;
    if (!prof_on[which]) return;
    auto &v = prof_ev[which];
    int &u = prof_used[which];
    if ((int)v.size() <= u) {
        hipEvent_t e;
        FFH_HIP(hipEventCreate(&e));
        v.push_back(e);
    }
    FFH_HIP(hipEventRecord(v[u], stream));
    ++u;
}

void Context::prof_end()
{
    for (int w = 0; w < 2; ++w) {
        double total = 0.0;
        int count = prof_used[w] / 2;
        for (int i = 0; i < count; ++i) {
            float ms = 0.0f;
            FFH_HIP(hipEventElapsedTime(&ms, prof_ev[w][2 * i], prof_ev[w][2 * i + 1]));
            total += ms;
        }
        prof_count[w] = count;
        prof_avg[w] = count ? total / count : 0.0;
    }
}

extern "C" {

int ffh_init(int device, ffh_context **out)
{
    if (!out) return FFH_ERR_INVALID;
    *out = nullptr;
    ffh_context *ctx = nullptr;
    int rc = guarded(nullptr, [&] {
        int count = 0;
        FFH_HIP(hipGetDeviceCount(&count));
        FFH_REQUIRE(device >= 0 && device < count, FFH_ERR_DEVICE, "no such HIP device");
        FFH_HIP(hipSetDevice(device));
        ctx = new ffh_context();
        ctx->device = device;
        hipDeviceProp_t prop;
        FFH_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        FFH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        FFH_HIP(hipEventCreate(&ctx->ev_start));
        FFH_HIP(hipEventCreate(&ctx->ev_stop));
        scratch_setup(device); // device address of the allocation-free solve's array
    });
    if (rc != FFH_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return FFH_OK;
}

int ffh_finalize(ffh_context *ctx)
{
    if (!ctx) return FFH_ERR_HANDLE;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &v : ctx->prof_ev)
        for (auto e : v) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->ev_start);
    (void)hipEventDestroy(ctx->ev_stop);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FFH_OK;
}

const char *ffh_last_error(const ffh_context *ctx)
{
    if (ctx) return ctx->err.c_str();
    return g_err.c_str();
}

int ffh_solver_options_default(ffh_solver_options *o)
{
    if (!o) return FFH_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->method = 1; // 'auto' -> 'pcg'
    o->preconditioner = 0;
    o->tolerance_type = 0;
    o->max_iterations = 1000;
    o->tolerance = 1.0e-6;
    o->precision = FFH_FP64;
    o->breakdown_mode = 0;
    o->inner_tolerance = 1.0e-4;
    o->max_refinements = 50;
    o->check_interval = 0;
    o->use_graph = 0;
    return FFH_OK;
}

int ffh_box_mesh_size(const int counts[3], int *nvert, int *ntet)
{
    if (!counts || !nvert || !ntet) return FFH_ERR_INVALID;
    for (int c = 0; c < 3; ++c)
        if (counts[c] < 1) return FFH_ERR_INVALID;
    int64_t nv = (int64_t)(counts[0] + 1) * (counts[1] + 1) * (counts[2] + 1);
    int64_t nt = 6 * (int64_t)counts[0] * counts[1] * counts[2];
    if (nv >= INT32_MAX || nt >= INT32_MAX) return FFH_ERR_INVALID;
    *nvert = (int)nv;
    *ntet = (int)nt;
    return FFH_OK;
}

int ffh_box_mesh(const double bounds[6], const int counts[3], double *vertices,
                 int *tetrahedra)
{
    int nv, nt;
    if (!bounds || !vertices || !tetrahedra) return FFH_ERR_INVALID;
    int rc = ffh_box_mesh_size(counts, &nv, &nt);
    if (rc) return rc;
    for (int c = 0; c < 3; ++c)
        if (!(bounds[3 + c] > bounds[c])) return FFH_ERR_INVALID;
    return guarded(nullptr, [&] { box_mesh(bounds, counts, vertices, tetrahedra); });
}

int ffh_tetra_duffy(int p, int *nq, double *x, double *y, double *z, double *w)
{
    if (p < 0 || !nq) return FFH_ERR_INVALID;
    const int m = (p + 4) / 2;
    *nq = m * m * m;
    if (!x) return FFH_OK;
    if (!y || !z || !w) return FFH_ERR_INVALID;
    return guarded(nullptr, [&] {
        std::vector<double> a, b, c, d;
        tetra_duffy(p, a, b, c, d);
        std::copy(a.begin(), a.end(), x);
        std::copy(b.begin(), b.end(), y);
        std::copy(c.begin(), c.end(), z);
        std::copy(d.begin(), d.end(), w);
    });
}

int ffh_unit_square_mesh(int n, double *vertices, int *triangles, int *is_boundary_vertex)
{
    if (n < 2 || !vertices || !triangles) return FFH_ERR_INVALID;
    return guarded(nullptr,
                   [&] { unit_square_mesh(n, vertices, triangles, is_boundary_vertex); });
}

int ffh_nedelec_dof_map(int k, int ntet, const int *tets, int base, int *global_dofs,
                        int *edge_orient, int *face_perm, int *ndofs)
{
    if (k < 1 || k > 6 || ntet < 1 || !tets || (base != 0 && base != 1)) return FFH_ERR_INVALID;
    return guarded(nullptr, [&] {
        std::vector<int> gd;
        std::vector<unsigned char> cls;
        int nd = 0, ne = 0, nf = 0;
        FFH_REQUIRE(nedelec_dof_map(k, ntet, tets, gd, cls, nd, ne, nf) == 0, FFH_ERR_INVALID,
                    "invalid tetrahedra for the Nedelec dof map");
        if (global_dofs)
            for (size_t i = 0; i < gd.size(); ++i) global_dofs[i] = gd[i] + base;
        for (int64_t e = 0; e < ntet; ++e) {
            int eo[6], fp[12];
            nedelec_class_orientation(cls[e], eo, fp);
            if (edge_orient) std::copy(eo, eo + 6, edge_orient + 6 * e);
            if (face_perm) std::copy(fp, fp + 12, face_perm + 12 * e);
        }
        if (ndofs) *ndofs = nd;
    });
}

int ffh_nedelec_basis(int k, int npts, const double *points, double *values, double *curls)
{
    if (k < 1 || k > 6 || npts < 0 || !points || !values || !curls) return FFH_ERR_INVALID;
    return guarded(nullptr, [&] {
        const int n = nedelec_dof_count(k);
        for (int q = 0; q < npts; ++q)
            FFH_REQUIRE(nedelec_basis_eval(k, points + 3 * q, values + (size_t)3 * n * q,
                                           curls + (size_t)3 * n * q) == 0,
                        FFH_ERR_INVALID, "Nedelec basis construction failed");
    });
}

int ffh_nedelec_load_points(int k, int *nq, double *points)
{
    if (k < 1 || k > 6 || !nq) return FFH_ERR_INVALID;
    return guarded(nullptr, [&] {
        std::vector<double> pts;
        nedelec_load_rule(k, pts);
        *nq = (int)(pts.size() / 4);
        if (points) std::copy(pts.begin(), pts.end(), points);
    });
}

int ffh_problem_create(ffh_context *ctx, int dim, int nvert, const double *vertices,
                       int ncell, const int *cells, ffh_problem **out)
{
    if (!ctx) return FFH_ERR_HANDLE;
    if (!out) return FFH_ERR_INVALID;
    *out = nullptr;
    ffh_problem *P = nullptr;
    int rc = guarded(ctx, [&] {
        FFH_REQUIRE(dim == 2 || dim == 3, FFH_ERR_INVALID, "dim must be 2 or 3");
        FFH_REQUIRE(nvert > 0 && ncell > 0 && vertices && cells, FFH_ERR_INVALID,
                    "empty mesh");
        const int vpc = dim == 3 ? 4 : 3;
        P = new ffh_problem();
        P->ctx = ctx;
        P->dim = dim;
        P->nvert = nvert;
        P->ncell = ncell;
        P->vertices.assign(vertices, vertices + (size_t)dim * nvert);
        P->cells.assign(cells, cells + (size_t)vpc * ncell);
        for (size_t i = 0; i < P->cells.size(); ++i)
            FFH_REQUIRE(P->cells[i] >= 1 && P->cells[i] <= nvert, FFH_ERR_INVALID,
                        "cell vertex id out of range (ids are 1-based)");
        P->d_vertices.upload(P->vertices, ctx->stream);
        FFH_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (rc != FFH_OK) {
        delete P;
        return rc;
    }
    *out = P;
    return FFH_OK;
}

int ffh_problem_free(ffh_problem *P)
{
    if (!P) return FFH_ERR_HANDLE;
    (void)hipStreamSynchronize(P->ctx->stream);
    delete P;
    return FFH_OK;
}

int ffh_pattern_build(ffh_problem *P, int family, int order, int *n_global, int64_t *nnz)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        problem_pattern(P, family, order);
        if (n_global) *n_global = P->n_global;
        if (nnz) *nnz = P->full.m.nnz;
    });
}

int ffh_problem_sizes(const ffh_problem *P, int *n_global, int64_t *nnz, int *n_free,
                      int64_t *nnz_free, int *n_local, int *n_colors)
{
    if (!P) return FFH_ERR_HANDLE;
    if (n_global) *n_global = P->n_global;
    if (nnz) *nnz = P->full.m.nnz;
    if (n_free) *n_free = P->have_dirichlet ? P->reduced.m.n : P->n_global;
    if (nnz_free) *nnz_free = P->have_dirichlet ? P->reduced.m.nnz : P->full.m.nnz;
    if (n_local) *n_local = P->n_local;
    if (n_colors) *n_colors = P->n_colors;
    return FFH_OK;
}

int ffh_global_dofs_download(ffh_problem *P, int base, int *gd)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!gd || (base != 0 && base != 1) || P->n_global == 0) return FFH_ERR_INVALID;
    for (size_t i = 0; i < P->global_dofs.size(); ++i) gd[i] = P->global_dofs[i] + base;
    return FFH_OK;
}

int ffh_problem_colouring(ffh_problem *P, int base, int *color_start, int *cells)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!color_start || !cells || (base != 0 && base != 1) || P->n_global == 0)
        return FFH_ERR_INVALID;
    for (size_t c = 0; c < P->color_start.size(); ++c) color_start[c] = P->color_start[c];
    for (size_t i = 0; i < P->perm.size(); ++i) cells[i] = P->perm[i] + base;
    return FFH_OK;
}

int ffh_assemble(ffh_problem *P, int qdeg, int coeff_kind, const double *coeff,
                 double mass, int precision)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        FFH_REQUIRE(precision == FFH_FP64 || precision == FFH_FP32, FFH_ERR_INVALID,
                    "precision must be FP64 or FP32");
        problem_assemble(P, qdeg, coeff_kind, coeff, mass, precision);
    });
}

int ffh_assemble_jvp(ffh_problem *P, int qdeg, int coeff_kind, const double *coeff,
                     double mass, const double *vertices_dot, const double *coeff_dot,
                     double mass_dot, double *values_dot)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!values_dot) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        problem_assemble_jvp(P, qdeg, coeff_kind, coeff, mass, vertices_dot, coeff_dot,
                             mass_dot, values_dot);
    });
}

int ffh_assemble_vjp(ffh_problem *P, int qdeg, int coeff_kind, const double *coeff,
                     double mass, const double *values_bar, double *vertices_bar,
                     double *coeff_bar, double *mass_bar)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!values_bar || !vertices_bar || !coeff_bar || !mass_bar) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        problem_assemble_vjp(P, qdeg, coeff_kind, coeff, mass, values_bar, vertices_bar,
                             coeff_bar, mass_bar);
    });
}

int ffh_nedelec_pml_values(ffh_problem *P, int kind, const double *stretch, double wave_number,
                           double *values)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!stretch || !values) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] { problem_nedelec_pml(P, kind, stretch, wave_number, values); });
}

int ffh_zmatrix_upload(ffh_context *ctx, int n, int64_t nnz, int base, const int *col_ptr,
                       const int *row_idx, const double *values, ffh_zmatrix **out)
{
    if (!ctx || !out) return FFH_ERR_HANDLE;
    *out = nullptr;
    ffh_zmatrix *z = nullptr;
    const int code = guarded(ctx, [&] {
        ZMatrix *m = zmatrix_create(ctx, n, nnz, base, col_ptr, row_idx, values);
        z = new ffh_zmatrix{ctx, m};
    });
    *out = z;
    return code;
}

int ffh_zmatrix_free(ffh_zmatrix *z)
{
    if (!z) return FFH_ERR_HANDLE;
    zmatrix_destroy(z->m);
    delete z;
    return FFH_OK;
}

int ffh_zmatrix_solve(ffh_zmatrix *z, const double *b, double *x, double tolerance,
                      int max_iterations, int *iterations, double *relative_residual,
                      int *converged)
{
    if (!z) return FFH_ERR_HANDLE;
    return guarded(z->ctx, [&] {
        zmatrix_solve(z->m, b, x, tolerance, max_iterations, iterations, relative_residual,
                      converged);
    });
}

int ffh_nedelec_pml_values_vjp(ffh_problem *P, int kind, const double *stretch,
                               double wave_number, const double *values_bar,
                               double *vertices_bar, double *stretch_bar,
                               double *wave_number_bar)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!stretch || !values_bar || !vertices_bar || !stretch_bar || !wave_number_bar)
        return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        problem_nedelec_pml_vjp(P, kind, stretch, wave_number, values_bar, vertices_bar,
                                stretch_bar, wave_number_bar);
    });
}

int ffh_nedelec_pml_values_jvp(ffh_problem *P, int kind, const double *stretch,
                               const double *stretch_dot, double wave_number,
                               double wave_number_dot, const double *vertices_dot,
                               double *values_dot)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!stretch || !stretch_dot || !values_dot) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        problem_nedelec_pml_jvp(P, kind, stretch, stretch_dot, wave_number, wave_number_dot,
                                vertices_dot, values_dot);
    });
}

int ffh_load(ffh_problem *P, int qdeg, int kind, const double *params, int64_t np)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] { problem_load(P, qdeg, kind, params, np); });
}

int ffh_dirichlet(ffh_problem *P, int mode, const double *values)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        FFH_REQUIRE(mode == 0 || mode == 1, FFH_ERR_INVALID, "mode must be 0 or 1");
        problem_dirichlet(P, mode, values);
        problem_reduce(P);
        FFH_HIP(hipStreamSynchronize(P->ctx->stream));
    });
}

int ffh_constrained_download(ffh_problem *P, int *constrained)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!constrained || !P->have_dirichlet) return FFH_ERR_INVALID;
    std::copy(P->constrained.begin(), P->constrained.end(), constrained);
    return FFH_OK;
}

int ffh_set_rhs(ffh_problem *P, int which, const double *rhs)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        FFH_REQUIRE(rhs, FFH_ERR_INVALID, "rhs must not be NULL");
        if (which == FFH_FULL) {
            FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
            P->rhs_full.upload(rhs, P->n_global, P->ctx->stream);
            P->red_dirty = true;
            P->loaded = true;
        } else {
            FFH_REQUIRE(P->have_dirichlet, FFH_ERR_INVALID, "no reduced system");
            problem_reduce(P);
            P->rhs_red.upload(rhs, P->reduced.m.n, P->ctx->stream);
        }
        FFH_HIP(hipStreamSynchronize(P->ctx->stream));
    });
}

int ffh_rhs_download(ffh_problem *P, int which, double *rhs)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        FFH_REQUIRE(rhs, FFH_ERR_INVALID, "rhs must not be NULL");
        if (which == FFH_FULL) {
            P->rhs_full.download(rhs, P->n_global, P->ctx->stream);
        } else {
            FFH_REQUIRE(P->have_dirichlet, FFH_ERR_INVALID, "no reduced system");
            problem_reduce(P);
            P->rhs_red.download(rhs, P->reduced.m.n, P->ctx->stream);
        }
        FFH_HIP(hipStreamSynchronize(P->ctx->stream));
    });
}

int ffh_problem_matrix(ffh_problem *P, int which, ffh_matrix **mat)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!mat) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        if (which == FFH_FULL) {
            FFH_REQUIRE(P->n_global > 0, FFH_ERR_INVALID, "pattern_build must come first");
            *mat = reinterpret_cast<ffh_matrix *>(&P->full);
        } else {
            FFH_REQUIRE(P->have_dirichlet, FFH_ERR_INVALID, "no reduced system");
            problem_reduce(P);
            *mat = reinterpret_cast<ffh_matrix *>(&P->reduced);
        }
    });
}

int ffh_problem_solve(ffh_problem *P, const ffh_solver_options *o, double *solution,
                      ffh_solver_stats *st)
{
    if (!P) return FFH_ERR_HANDLE;
    return guarded(P->ctx, [&] {
        check_options(o);
        FFH_REQUIRE(st, FFH_ERR_INVALID, "stats must not be NULL");
        std::memset(st, 0, sizeof(*st));
        problem_solve(P, *o, solution, *st);
    });
}

int ffh_problem_solution(ffh_problem *P, double *solution)
{
    if (!P) return FFH_ERR_HANDLE;
    if (!solution) return FFH_ERR_INVALID;
    return guarded(P->ctx, [&] {
        FFH_REQUIRE(P->x_full.p && (int)P->x_full.n >= P->n_global, FFH_ERR_INVALID,
                    "no solution yet: solve first");
        P->x_full.download(solution, P->n_global, P->ctx->stream);
        FFH_HIP(hipStreamSynchronize(P->ctx->stream));
    });
}

int ffh_matrix_upload(ffh_context *ctx, int n, int64_t nnz, int base, const int *col_ptr,
                      const int *row_idx, const double *val, ffh_matrix **out)
{
    if (!ctx) return FFH_ERR_HANDLE;
    if (!out || !col_ptr || (nnz > 0 && (!row_idx || !val))) return FFH_ERR_INVALID;
    *out = nullptr;
    ffh_matrix *M = new ffh_matrix();
    M->ctx = ctx;
    int rc = guarded(ctx, [&] { load_matrix(ctx, M->m, n, nnz, base, col_ptr, row_idx, val); });
    if (rc != FFH_OK) {
        delete M;
        return rc;
    }
    *out = M;
    return FFH_OK;
}

int ffh_matrix_free(ffh_matrix *M)
{
    if (!M) return FFH_ERR_HANDLE;
    if (M->borrowed) return FFH_ERR_HANDLE; // owned by its problem
    (void)hipStreamSynchronize(M->ctx->stream);
    delete M;
    return FFH_OK;
}

int ffh_matrix_download(ffh_matrix *M, int base, int64_t capacity, int *col_ptr,
                        int *row_idx, double *val, int *n, int64_t *nnz)
{
    if (!M) return FFH_ERR_HANDLE;
    if (base != 0 && base != 1) return FFH_ERR_INVALID;
    if (n) *n = M->m.n;
    if (nnz) *nnz = M->m.nnz;
    if (capacity < M->m.nnz) return FFH_ERR_CAPACITY;
    return guarded(M->ctx, [&] {
        hipStream_t s = M->ctx->stream;
        ensure_val64(M->ctx, M->m);
        if (col_ptr) M->m.col_ptr.download(col_ptr, M->m.n + 1, s);
        if (row_idx) M->m.row_idx.download(row_idx, M->m.nnz, s);
        if (val) M->m.val.download(val, M->m.nnz, s);
        FFH_HIP(hipStreamSynchronize(s));
        if (base) {
            if (col_ptr)
                for (int j = 0; j <= M->m.n; ++j) col_ptr[j] += 1;
            if (row_idx)
                for (int64_t k = 0; k < M->m.nnz; ++k) row_idx[k] += 1;
        }
    });
}

int ffh_spmv(ffh_matrix *M, const double *x, double *y)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!x || !y) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        Context *ctx = M->ctx;
        ensure_workspace(ctx, *M);
        ensure_val64(ctx, M->m);
        // the solver's b and Ap vectors serve as scratch (sizes unchanged)
        FFH_HIP(hipMemcpyAsync(M->ws.b.p, x, sizeof(double) * M->m.n, hipMemcpyHostToDevice,
                               ctx->stream));
        spmv64(ctx, M->m, M->ws.b.p, M->ws.ap.p);
        M->ws.ap.download(y, M->m.n, ctx->stream);
        FFH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ffh_solve_sparse(ffh_matrix *M, const double *b, double *x, const ffh_solver_options *o,
                     ffh_solver_stats *st)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!b || !x || !st) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        check_options(o);
        std::memset(st, 0, sizeof(*st));
        Context *ctx = M->ctx;
        const int n = M->m.n;
        DBuf<double> db, dx;
        db.upload(b, n, ctx->stream);
        dx.upload(x, n, ctx->stream);
        if (o->precision == FFH_MIXED)
            solve_mixed(ctx, *M, db.p, dx.p, *o, *st);
        else
            solve_fp64(ctx, *M, db.p, dx.p, *o, *st);
        dx.download(x, n, ctx->stream);
        FFH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int ffh_solve_sparse_device_noalloc(ffh_context *ctx, int n, int64_t nnz, const int *col_ptr,
                                    const int *row_idx, const double *val, const double *b,
                                    double *x, const ffh_solver_options *o, ffh_solver_stats *st)
{
    if (!ctx) return FFH_ERR_HANDLE;
    if (!col_ptr || !row_idx || !val || !b || !x || !st) return FFH_ERR_INVALID;
    return guarded(ctx, [&] {
        check_options(o);
        FFH_REQUIRE(o->precision == FFH_FP64, FFH_ERR_INVALID,
                    "allocation-free solve: precision must be FP64");
        std::memset(st, 0, sizeof(*st));
        solve_fp64_noalloc(ctx, n, nnz, col_ptr, row_idx, val, b, x, *o, *st);
    });
}

int ffh_noalloc_spmv_variant(int n, int64_t nnz) { return noalloc_variant(n, nnz); }

int ffh_spmv_jvp(ffh_matrix *M, const double *x, const double *values_dot, const double *x_dot,
                 double *y_dot)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!x || !values_dot || !x_dot || !y_dot) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        Context *ctx = M->ctx;
        const int n = M->m.n;
        hipStream_t s = ctx->stream;
        ensure_val64(ctx, M->m);
        DBuf<double> dx, dvd, dxd, dy;
        dx.upload(x, n, s);
        dvd.upload(values_dot, M->m.nnz, s);
        dxd.upload(x_dot, n, s);
        dy.alloc(n);
        csc_jvp(ctx, M->m, dvd.p, dx.p, dxd.p, dy.p);
        dy.download(y_dot, n, s);
        FFH_HIP(hipStreamSynchronize(s));
    });
}

int ffh_spmv_vjp(ffh_matrix *M, const double *x, const double *y_bar, double *values_bar,
                 double *x_bar)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!x || !y_bar || !values_bar || !x_bar) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        Context *ctx = M->ctx;
        const int n = M->m.n;
        hipStream_t s = ctx->stream;
        ensure_val64(ctx, M->m);
        DBuf<double> dx, dyb, dvb, dxb;
        dx.upload(x, n, s);
        dyb.upload(y_bar, n, s);
        dvb.alloc(std::max<int64_t>(M->m.nnz, 1));
        dxb.alloc(n);
        csc_vjp(ctx, M->m, dx.p, dyb.p, 1.0, dvb.p, dxb.p);
        dvb.download(values_bar, M->m.nnz, s);
        dxb.download(x_bar, n, s);
        FFH_HIP(hipStreamSynchronize(s));
    });
}

int ffh_solve_sparse_jvp(ffh_matrix *M, const double *x, const double *values_dot,
                         const double *b_dot, const ffh_solver_options *o, double *x_dot,
                         ffh_solver_stats *st)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!x || !values_dot || !b_dot || !x_dot || !st) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        check_options(o);
        FFH_REQUIRE(M->m.symmetric, FFH_ERR_INVALID, "CG tangents need a symmetric matrix");
        FFH_REQUIRE(o->max_iterations >= 1 && o->tolerance > 0.0, FFH_ERR_INVALID,
                    "max_iterations >= 1 and tolerance > 0");
        std::memset(st, 0, sizeof(*st));
        Context *ctx = M->ctx;
        const int n = M->m.n;
        hipStream_t s = ctx->stream;
        ensure_val64(ctx, M->m);
        DBuf<double> dx, dvd, rhs, xd;
        dx.upload(x, n, s);
        dvd.upload(values_dot, M->m.nnz, s);
        rhs.alloc(n);
        // A_dot x, then tangent_rhs = b_dot - A_dot x (:410-411)
        csc_jvp(ctx, M->m, dvd.p, dx.p, nullptr, rhs.p);
        std::vector<double> h(n);
        rhs.download(h.data(), n, s);
        FFH_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) h[i] = b_dot[i] - h[i];
        rhs.upload(h, s);
        xd.alloc(n);
        FFH_HIP(hipMemsetAsync(xd.p, 0, sizeof(double) * n, s));
        if (o->precision == FFH_MIXED)
            solve_mixed(ctx, *M, rhs.p, xd.p, *o, *st);
        else
            solve_fp64(ctx, *M, rhs.p, xd.p, *o, *st);
        xd.download(x_dot, n, s);
        FFH_HIP(hipStreamSynchronize(s));
    });
}

int ffh_solve_sparse_vjp(ffh_matrix *M, const double *x, const double *x_bar,
                         const ffh_solver_options *o, double *values_bar, double *b_bar,
                         ffh_solver_stats *st)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!x || !x_bar || !values_bar || !b_bar || !st) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        check_options(o);
        FFH_REQUIRE(M->m.symmetric, FFH_ERR_INVALID, "CG adjoints need a symmetric matrix");
        FFH_REQUIRE(o->max_iterations >= 1 && o->tolerance > 0.0, FFH_ERR_INVALID,
                    "max_iterations >= 1 and tolerance > 0");
        std::memset(st, 0, sizeof(*st));
        Context *ctx = M->ctx;
        const int n = M->m.n;
        hipStream_t s = ctx->stream;
        ensure_val64(ctx, M->m);
        DBuf<double> dx, dxb, lam, dvb;
        dx.upload(x, n, s);
        dxb.upload(x_bar, n, s);
        lam.alloc(n);
        FFH_HIP(hipMemsetAsync(lam.p, 0, sizeof(double) * n, s));
        // A^T lambda = x_bar (A = A^T here), b_bar = lambda,
        // A_bar = -lambda x^T on the stored entries (:441-448)
        if (o->precision == FFH_MIXED)
            solve_mixed(ctx, *M, dxb.p, lam.p, *o, *st);
        else
            solve_fp64(ctx, *M, dxb.p, lam.p, *o, *st);
        dvb.alloc(std::max<int64_t>(M->m.nnz, 1));
        csc_vjp(ctx, M->m, dx.p, lam.p, -1.0, dvb.p, nullptr);
        dvb.download(values_bar, M->m.nnz, s);
        lam.download(b_bar, n, s);
        FFH_HIP(hipStreamSynchronize(s));
    });
}

int ffh_bench_spmv(ffh_matrix *M, int precision, int reps, double *avg_ms, double *bytes)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!avg_ms || !bytes || reps < 1) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] { *avg_ms = bench_spmv(M->ctx, *M, precision, reps, bytes); });
}

int ffh_matrix_spmv_bytes(ffh_matrix *M, int precision, double *bytes)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!bytes || (precision != FFH_FP64 && precision != FFH_FP32)) return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] { *bytes = spmv_launch_bytes(M->ctx, *M, precision == FFH_FP32); });
}

int ffh_pcg_launches_per_iteration(int64_t n, int precision)
{
    if (n < 0 || (precision != FFH_FP64 && precision != FFH_FP32)) return FFH_ERR_INVALID;
    return three_launch(n, precision == FFH_FP32) ? 3 : 2;
}

int ffh_matrix_spmv_variant(ffh_matrix *M, int precision, int request, int *chosen)
{
    if (!M) return FFH_ERR_HANDLE;
    if (!chosen || (precision != FFH_FP64 && precision != FFH_FP32) ||
        request >= FFH_SPMV_VARIANTS)
        return FFH_ERR_INVALID;
    return guarded(M->ctx, [&] {
        ensure_workspace(M->ctx, *M);
        *chosen = select_spmv_variant(M->ctx, *M, precision == FFH_FP32, request);
    });
}

int ffh_set_profiling(ffh_context *ctx, int enable)
{
    if (!ctx) return FFH_ERR_HANDLE;
    ctx->profiling = enable != 0;
    return FFH_OK;
}

int ffh_last_solve_profile(ffh_context *ctx, int which, int *count, double *avg_ms)
{
    if (!ctx) return FFH_ERR_HANDLE;
    if (which < 0 || which > 1 || !count || !avg_ms) return FFH_ERR_INVALID;
    *count = ctx->prof_count[which];
    *avg_ms = ctx->prof_avg[which];
    return FFH_OK;
}

} // extern "C"
