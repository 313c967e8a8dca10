// ffh_nedelec.cc -- host side of the tetrahedral first-kind Nedelec family.
//
// Reference: src/elements/tetra_nedelec_arbitrary_order.f90 (moment-dual
// basis, build_runtime_coefficients :203-315, runtime candidates :317-397,
// face / volume moment polynomials :642-669), tetra_nedelec_first_order.f90
// (:11-51, Whitney functions for k = 1), tetra_edge_dof_map.f90 (:10-77),
// tetra_nedelec_global_dof_map.f90 (numbering :15-105, basis transform
// :107-155), tetra_face_moment_transforms.f90 (runtime face transform
// :296-381), assembly_tetra_nedelec_3d.f90 (element rules :584-643,
// :2357-2404, :2406-2514), fortfem_tetra_nedelec_solver_3d.f90
// (zero tangential mask :431-470).
//
// Design: on an affine cell the reference's element matrix
//   M_rc = sum_q det w_q [ (J c^_r)^T K (J c^_c) / det^2 + beta (J^-T v^_r).(J^-T v^_c) ]
// is a 12-term combination of reference tensors, with the cell entering only
// through the symmetric 3x3 matrices J^T K J / det and beta det J^-1 J^-T.
// The reference's orientation transform T depends only on the rank pattern of
// the cell's four vertex ids, so T^T M T is the same combination of the
// transformed tensors T^T R_s T of the cell's class (at most 24).  This file
// builds those tables once per order; the device only forms 12-term sums.
// The moment-dual coefficients are solved in IEEE quad precision (the scaled
// moment matrix loses about 12 digits in double at k = 4); basis values and
// tensors are accumulated in 80-bit long double and rounded once.
#include "ffh_nedelec.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

namespace ffh {

typedef __float128 qd;
typedef long double xd;

int nedelec_dof_count(int k) { return k * (k + 2) * (k + 3) / 2; }

static const int kEdgeV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int kFaceV[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// ------------------------------------------------------------ numbering --
namespace {

struct Key3Hash {
    size_t operator()(const uint64_t &k) const
    {
        uint64_t h = k * 0x9e3779b97f4a7c15ull;
        return (size_t)(h ^ (h >> 29));
    }
};

// first-appearance ids for keys of up to three vertex ids
class FirstAppearance {
  public:
    explicit FirstAppearance(size_t n) { map_.reserve(n); }
    int get(int a, int b, int c)
    {
        if (a < (1 << 21) && b < (1 << 21) && c < (1 << 21)) {
            const uint64_t key = (uint64_t)a | ((uint64_t)b << 21) | ((uint64_t)c << 42);
            auto it = map_.emplace(key, count_);
            if (it.second) ++count_;
            return it.first->second;
        }
        auto it = wide_.emplace(std::make_tuple(a, b, c), count_);
        if (it.second) ++count_;
        return it.first->second;
    }
    int count() const { return count_; }

  private:
    std::unordered_map<uint64_t, int, Key3Hash> map_;
    std::map<std::tuple<int, int, int>, int> wide_; // vertex ids >= 2^21
    int count_ = 0;
};

// class id of a rank pattern (ranks of vertices 0..3 among the cell's ids)
struct ClassTable {
    int code_to_class[256];
    int ranks[kNedelecClasses][4];
    ClassTable()
    {
        std::fill(code_to_class, code_to_class + 256, -1);
        int p[4] = {0, 1, 2, 3}, c = 0;
        do {
            int code = p[0] | (p[1] << 2) | (p[2] << 4) | (p[3] << 6);
            code_to_class[code] = c;
            for (int i = 0; i < 4; ++i) ranks[c][i] = p[i];
            ++c;
        } while (std::next_permutation(p, p + 4));
    }
};

const ClassTable &class_table()
{
    static ClassTable t;
    return t;
}

} // namespace

void nedelec_cell_classes(int ntet, const int *tets, std::vector<unsigned char> &cls)
{
    const ClassTable &ct = class_table();
    cls.assign(ntet, 0);
    parallel_for(ntet, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            const int *t = tets + 4 * e;
            int rk[4];
            for (int i = 0; i < 4; ++i) {
                rk[i] = 0;
                for (int j = 0; j < 4; ++j) rk[i] += t[j] < t[i];
            }
            cls[e] = (unsigned char)ct.code_to_class[rk[0] | (rk[1] << 2) | (rk[2] << 4) |
                                                     (rk[3] << 6)];
        }
    });
}

int nedelec_dof_map(int k, int ntet, const int *tets, std::vector<int> &gd,
                    std::vector<unsigned char> &cls, int &ndofs, int &nedges, int &nfaces)
{
    if (k < 1) return 1;
    for (int64_t e = 0; e < ntet; ++e)
        for (int a = 0; a < 4; ++a) {
            if (tets[4 * e + a] < 1) return 1;
            for (int b = a + 1; b < 4; ++b)
                if (tets[4 * e + a] == tets[4 * e + b]) return 1;
        }
    const int n = nedelec_dof_count(k), fdc = k * (k - 1), cdc = k * (k - 1) * (k - 2) / 2;
    gd.assign((size_t)n * ntet, 0);
    cls.assign(ntet, 0);
    const ClassTable &ct = class_table();
    // edges: (min, max) by first appearance over (cell, local edge)
    FirstAppearance edges((size_t)2 * ntet + 16);
    std::vector<int> eid((size_t)6 * ntet);
    for (int64_t e = 0; e < ntet; ++e) {
        const int *t = tets + 4 * e;
        int rk[4];
        for (int i = 0; i < 4; ++i) {
            rk[i] = 0;
            for (int j = 0; j < 4; ++j) rk[i] += t[j] < t[i];
        }
        cls[e] = (unsigned char)ct.code_to_class[rk[0] | (rk[1] << 2) | (rk[2] << 4) | (rk[3] << 6)];
        for (int l = 0; l < 6; ++l) {
            int a = t[kEdgeV[l][0]], b = t[kEdgeV[l][1]];
            eid[6 * e + l] = edges.get(std::min(a, b), std::max(a, b), 0);
        }
    }
    nedges = edges.count();
    const int edc = k * nedges;
    FirstAppearance faces((size_t)2 * ntet + 16);
    for (int64_t e = 0; e < ntet; ++e) {
        const int *t = tets + 4 * e;
        int *g = gd.data() + (size_t)n * e;
        for (int l = 0; l < 6; ++l)
            for (int m = 0; m < k; ++m) g[l * k + m] = eid[6 * e + l] * k + m;
        for (int f = 0; f < 4; ++f) {
            int v[3] = {t[kFaceV[f][0]], t[kFaceV[f][1]], t[kFaceV[f][2]]};
            std::sort(v, v + 3);
            const int face = faces.get(v[0], v[1], v[2]);
            for (int d = 0; d < fdc; ++d) g[6 * k + f * fdc + d] = edc + face * fdc + d;
        }
    }
    nfaces = faces.count();
    for (int64_t e = 0; e < ntet; ++e) {
        int *g = gd.data() + (size_t)n * e;
        for (int d = 0; d < cdc; ++d)
            g[6 * k + 4 * fdc + d] = edc + nfaces * fdc + (int)e * cdc + d;
    }
    ndofs = edc + nfaces * fdc + ntet * cdc;
    return 0;
}

void nedelec_boundary_dofs(int k, int ntet, const std::vector<unsigned char> &bf,
                           const std::vector<int> &gd, std::vector<unsigned char> &constrained)
{
    const int n = nedelec_dof_count(k), fdc = k * (k - 1);
    int nd = 0;
    for (int v : gd) nd = std::max(nd, v + 1);
    constrained.assign(nd, 0);
    // local edges of each local face (positions in kEdgeV); local face f
    // (kFaceV) is the face opposite local vertex 3 - f
    static const int kFaceEdges[4][3] = {{0, 1, 3}, {0, 2, 4}, {1, 2, 5}, {3, 4, 5}};
    for (int64_t e = 0; e < ntet; ++e) {
        const int *g = gd.data() + (size_t)n * e;
        for (int f = 0; f < 4; ++f) {
            if (!bf[4 * e + (3 - f)]) continue;
            for (int j = 0; j < 3; ++j)
                for (int m = 0; m < k; ++m) constrained[g[kFaceEdges[f][j] * k + m]] = 1;
            for (int d = 0; d < fdc; ++d) constrained[g[6 * k + f * fdc + d]] = 1;
        }
    }
}

// ---------------------------------------------------- quad-precision basis --
namespace {

qd qabs(qd x) { return x < 0 ? -x : x; }

qd qpow(qd x, int e)
{
    qd r = 1;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

// Gauss-Legendre on [0, 1], ascending nodes (Newton on the recurrence)
void gauss01(int n, std::vector<qd> &x, std::vector<qd> &w)
{
    x.assign(n, 0);
    w.assign(n, 0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        qd z = (qd)std::cos((long double)M_PI * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; ++it) {
            qd p0 = 1, p1 = z;
            for (int j = 2; j <= n; ++j) {
                qd p2 = ((2 * j - 1) * z * p1 - (qd)(j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            qd dp = n * (z * p1 - p0) / (z * z - 1);
            qd dz = p1 / dp;
            z -= dz;
            if (qabs(dz) < (qd)1e-32L) break;
        }
        qd p0 = 1, p1 = z;
        for (int j = 2; j <= n; ++j) {
            qd p2 = ((2 * j - 1) * z * p1 - (qd)(j - 1) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        qd dp = n * (z * p1 - p0) / (z * z - 1);
        qd wt = 2 / ((1 - z * z) * dp * dp);
        const int lo = i, hi = n - 1 - i;
        x[lo] = (1 - z) / 2;
        x[hi] = (1 + z) / 2;
        w[lo] = w[hi] = wt / 2;
        if (lo == hi) x[lo] = (qd)0.5L;
    }
}

// collapsed Gauss rules (tetra_duffy_quadrature / triangle_duffy_quadrature)
void tri_rule(int p, std::vector<qd> &x, std::vector<qd> &y, std::vector<qd> &w)
{
    const int m = (p + 3) / 2;
    std::vector<qd> s, sw;
    gauss01(m, s, sw);
    x.clear(), y.clear(), w.clear();
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
            x.push_back(s[a]);
            y.push_back((1 - s[a]) * s[b]);
            w.push_back(sw[a] * sw[b] * (1 - s[a]));
        }
}

void tet_rule(int p, std::vector<qd> &x, std::vector<qd> &y, std::vector<qd> &z,
              std::vector<qd> &w)
{
    const int m = (p + 4) / 2;
    std::vector<qd> s, sw;
    gauss01(m, s, sw);
    x.clear(), y.clear(), z.clear(), w.clear();
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            for (int c = 0; c < m; ++c) {
                const qd c1 = 1 - s[a], c2 = 1 - s[b];
                x.push_back(s[a]);
                y.push_back(c1 * s[b]);
                z.push_back(c1 * c2 * s[c]);
                w.push_back(sw[a] * sw[b] * sw[c] * c1 * c1 * c2);
            }
}

// Jacobi P_n^(a,0)
qd jacobi(int n, int a, qd x)
{
    if (n == 0) return 1;
    qd p0 = 1, p1 = (qd)a / 2 + (qd)(a + 2) / 2 * x;
    for (int m = 2; m <= n; ++m) {
        const qd c = 2 * m + a;
        const qd a1 = 2 * (qd)m * (m + a) * (c - 2);
        const qd a2 = (c - 1) * (c * (c - 2) * x + (qd)a * a);
        const qd a3 = 2 * (qd)(m + a - 1) * (m - 1) * c;
        const qd p2 = (a2 * p1 - a3 * p0) / a1;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// face / volume moment polynomials: monomials up to k = 5, orthogonal
// collapsed-coordinate (Dubiner / Koornwinder) polynomials from k = 6
qd face_poly(int k, int a, int b, qd x, qd y)
{
    if (k <= 5) return qpow(x, a) * qpow(y, b);
    const qd s = 1 - y, e = s > 0 ? 2 * x / s - 1 : (qd)-1;
    return jacobi(a, 0, e) * qpow(s, a) * jacobi(b, 2 * a + 1, 2 * y - 1);
}

qd volume_poly(int k, int a, int b, int g, qd x, qd y, qd z)
{
    if (k <= 5) return qpow(x, a) * qpow(y, b) * qpow(z, g);
    const qd s1 = 1 - y - z, s2 = 1 - z;
    const qd e1 = s1 > 0 ? 2 * x / s1 - 1 : (qd)-1, e2 = s2 > 0 ? 2 * y / s2 - 1 : (qd)-1;
    return jacobi(a, 0, e1) * qpow(s1, a) * jacobi(b, 2 * a + 1, e2) * qpow(s2, b) *
           jacobi(g, 2 * a + 2 * b + 2, 2 * z - 1);
}

qd legendre01(int m, qd t)
{
    const qd x = 2 * t - 1;
    if (m == 0) return 1;
    qd p0 = 1, p1 = x;
    for (int d = 2; d <= m; ++d) {
        const qd p2 = ((2 * d - 1) * x * p1 - (qd)(d - 1) * p0) / d;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// One monomial term coef x^a y^b z^g in component comp: value and curl.
template <typename R>
void add_term(R *v, R *c, int comp, R coef, int a, int b, int g, const R p[3])
{
    auto pw = [](R x, int e) {
        R r = 1;
        for (int i = 0; i < e; ++i) r *= x;
        return r;
    };
    const R x = p[0], y = p[1], z = p[2];
    v[comp] += coef * pw(x, a) * pw(y, b) * pw(z, g);
    const R dx = a ? coef * a * pw(x, a - 1) * pw(y, b) * pw(z, g) : R(0);
    const R dy = b ? coef * b * pw(x, a) * pw(y, b - 1) * pw(z, g) : R(0);
    const R dz = g ? coef * g * pw(x, a) * pw(y, b) * pw(z, g - 1) : R(0);
    if (comp == 0) {
        c[1] += dz;
        c[2] -= dy;
    } else if (comp == 1) {
        c[0] -= dz;
        c[2] += dx;
    } else {
        c[0] += dy;
        c[1] -= dx;
    }
}

// Candidate spanning set of the runtime basis (:317-397): P_{k-1}^3, then the
// homogeneous degree-k fields of the two families built on x (y e_x - x e_y
// style pairs), then the z family.  vals / curls: 3 per candidate.
template <typename R>
void candidates(int k, const R p[3], R *vals, R *curls)
{
    const int n = nedelec_dof_count(k);
    std::fill(vals, vals + 3 * n, R(0));
    std::fill(curls, curls + 3 * n, R(0));
    int j = 0;
    for (int comp = 0; comp < 3; ++comp)
        for (int td = 0; td <= k - 1; ++td)
            for (int a = 0; a <= td; ++a)
                for (int b = 0; b <= td - a; ++b, ++j)
                    add_term<R>(vals + 3 * j, curls + 3 * j, comp, R(1), a, b, td - a - b, p);
    const int td = k - 1;
    for (int fam = 0; fam < 2; ++fam)
        for (int a = 0; a <= td; ++a)
            for (int b = 0; b <= td - a; ++b, ++j) {
                const int g = td - a - b;
                if (fam == 0) {
                    add_term<R>(vals + 3 * j, curls + 3 * j, 0, R(-1), a, b + 1, g, p);
                    add_term<R>(vals + 3 * j, curls + 3 * j, 1, R(1), a + 1, b, g, p);
                } else {
                    add_term<R>(vals + 3 * j, curls + 3 * j, 0, R(-1), a, b, g + 1, p);
                    add_term<R>(vals + 3 * j, curls + 3 * j, 2, R(1), a + 1, b, g, p);
                }
            }
    for (int b = 0; b <= td; ++b, ++j) {
        const int g = td - b;
        add_term<R>(vals + 3 * j, curls + 3 * j, 1, R(-1), 0, b, g + 1, p);
        add_term<R>(vals + 3 * j, curls + 3 * j, 2, R(1), 0, b + 1, g, p);
    }
}

// Gaussian elimination with partial pivoting: A (n x n, row-major) X = B
// (n x m), B overwritten with X.
int solve_dense(int n, std::vector<qd> &A, int m, std::vector<qd> &B)
{
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (qabs(A[(size_t)r * n + c]) > qabs(A[(size_t)p * n + c])) p = r;
        if (A[(size_t)p * n + c] == 0) return 1;
        if (p != c) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)c * n + j], A[(size_t)p * n + j]);
            for (int j = 0; j < m; ++j) std::swap(B[(size_t)c * m + j], B[(size_t)p * m + j]);
        }
        const qd inv = 1 / A[(size_t)c * n + c];
        for (int r = c + 1; r < n; ++r) {
            const qd f = A[(size_t)r * n + c] * inv;
            if (f == 0) continue;
            for (int j = c; j < n; ++j) A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
            for (int j = 0; j < m; ++j) B[(size_t)r * m + j] -= f * B[(size_t)c * m + j];
        }
    }
    for (int c = n - 1; c >= 0; --c)
        for (int j = 0; j < m; ++j) {
            qd s = B[(size_t)c * m + j];
            for (int i = c + 1; i < n; ++i) s -= A[(size_t)c * n + i] * B[(size_t)i * m + j];
            B[(size_t)c * m + j] = s / A[(size_t)c * n + c];
        }
    return 0;
}

const qd kRefV[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

// build_runtime_coefficients (:203-315): the moment matrix of the degrees of
// freedom applied to the candidates, row / column max-scaled, inverted.
int build_coefficients(int k, std::vector<qd> &coeff)
{
    const int n = nedelec_dof_count(k);
    std::vector<qd> M((size_t)n * n, 0), cv(3 * n), cc(3 * n);
    int row = 0;
    std::vector<qd> en, ew;
    gauss01(k + 2, en, ew);
    for (int e = 0; e < 6; ++e) {
        qd tan[3], p[3];
        for (int c = 0; c < 3; ++c) tan[c] = kRefV[kEdgeV[e][1]][c] - kRefV[kEdgeV[e][0]][c];
        for (int m = 0; m < k; ++m, ++row)
            for (int q = 0; q < k + 2; ++q) {
                for (int c = 0; c < 3; ++c) p[c] = kRefV[kEdgeV[e][0]][c] + en[q] * tan[c];
                candidates<qd>(k, p, cv.data(), cc.data());
                const qd f = ew[q] * legendre01(m, en[q]);
                for (int j = 0; j < n; ++j)
                    M[(size_t)row * n + j] +=
                        f * (tan[0] * cv[3 * j] + tan[1] * cv[3 * j + 1] + tan[2] * cv[3 * j + 2]);
            }
    }
    // face moments (candidates evaluated once per point, rows (comp, td, a))
    const int fdc = k * (k - 1);
    std::vector<qd> tx, ty, tw;
    tri_rule(2 * k + 4, tx, ty, tw);
    for (int f = 0; f < 4 && fdc > 0; ++f) {
        qd t[2][3], p[3];
        for (int c = 0; c < 3; ++c) {
            t[0][c] = kRefV[kFaceV[f][1]][c] - kRefV[kFaceV[f][0]][c];
            t[1][c] = kRefV[kFaceV[f][2]][c] - kRefV[kFaceV[f][0]][c];
        }
        for (size_t q = 0; q < tw.size(); ++q) {
            for (int c = 0; c < 3; ++c)
                p[c] = kRefV[kFaceV[f][0]][c] + tx[q] * t[0][c] + ty[q] * t[1][c];
            candidates<qd>(k, p, cv.data(), cc.data());
            int r = 6 * k + f * fdc;
            for (int comp = 0; comp < 2; ++comp)
                for (int td = 0; td <= k - 2; ++td)
                    for (int a = 0; a <= td; ++a, ++r) {
                        const qd w = tw[q] * face_poly(k, a, td - a, tx[q], ty[q]);
                        for (int j = 0; j < n; ++j)
                            M[(size_t)r * n + j] +=
                                w * (t[comp][0] * cv[3 * j] + t[comp][1] * cv[3 * j + 1] +
                                     t[comp][2] * cv[3 * j + 2]);
                    }
        }
    }
    row += 4 * fdc;
    // volume moments, rows (comp, td, a, b)
    const int cdc = k * (k - 1) * (k - 2) / 2;
    std::vector<qd> vx, vy, vz, vw;
    tet_rule(2 * k + 4, vx, vy, vz, vw);
    if (cdc > 0) {
        std::vector<qd> polys;
        for (size_t q = 0; q < vw.size(); ++q) {
            const qd p[3] = {vx[q], vy[q], vz[q]};
            candidates<qd>(k, p, cv.data(), cc.data());
            polys.clear();
            for (int td = 0; td <= k - 3; ++td)
                for (int a = 0; a <= td; ++a)
                    for (int b = 0; b <= td - a; ++b)
                        polys.push_back(vw[q] * volume_poly(k, a, b, td - a - b, p[0], p[1], p[2]));
            int r = row;
            for (int comp = 0; comp < 3; ++comp)
                for (size_t i = 0; i < polys.size(); ++i, ++r)
                    for (int j = 0; j < n; ++j) M[(size_t)r * n + j] += polys[i] * cv[3 * j + comp];
        }
    }
    row += cdc;
    if (row != n) return 1;
    std::vector<qd> rs(n), cs(n);
    for (int i = 0; i < n; ++i) {
        qd mx = 0;
        for (int j = 0; j < n; ++j) mx = std::max(mx, qabs(M[(size_t)i * n + j]));
        rs[i] = 1 / mx;
        for (int j = 0; j < n; ++j) M[(size_t)i * n + j] *= rs[i];
    }
    for (int j = 0; j < n; ++j) {
        qd mx = 0;
        for (int i = 0; i < n; ++i) mx = std::max(mx, qabs(M[(size_t)i * n + j]));
        cs[j] = 1 / mx;
        for (int i = 0; i < n; ++i) M[(size_t)i * n + j] *= cs[j];
    }
    std::vector<qd> X((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) X[(size_t)i * n + i] = rs[i];
    if (solve_dense(n, M, n, X)) return 1;
    coeff.resize((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) coeff[(size_t)i * n + j] = cs[i] * X[(size_t)i * n + j];
    return 0;
}

std::mutex g_cache_mutex;

const std::vector<qd> *coefficients_cached(int k)
{
    static std::vector<qd> cache[9];
    static bool done[9] = {false};
    if (k < 2 || k > 8) return nullptr;
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    if (!done[k]) {
        if (build_coefficients(k, cache[k])) return nullptr;
        done[k] = true;
    }
    return &cache[k];
}

// Face moment matrix of a face vertex permutation (:335-381), s = k(k-1)
void face_moment_matrix(int k, const int perm[3], std::vector<qd> &M)
{
    static const int V2[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    const int off[2] = {V2[perm[0] - 1][0], V2[perm[0] - 1][1]};
    int aff[2][2];
    for (int i = 0; i < 2; ++i) {
        aff[i][0] = V2[perm[1] - 1][i] - off[i];
        aff[i][1] = V2[perm[2] - 1][i] - off[i];
    }
    const int s = k * (k - 1);
    M.assign((size_t)s * s, 0);
    std::vector<qd> u, v, w;
    tri_rule(2 * k + 2, u, v, w);
    int col = 0;
    for (int bc = 0; bc < 2; ++bc)
        for (int bd = 0; bd <= k - 2; ++bd)
            for (int bx = 0; bx <= bd; ++bx, ++col) {
                const int by = bd - bx;
                int row = 0;
                for (int mc = 0; mc < 2; ++mc)
                    for (int md = 0; md <= k - 2; ++md)
                        for (int mx = 0; mx <= md; ++mx, ++row) {
                            const int my = md - mx;
                            qd sum = 0;
                            for (size_t q = 0; q < w.size(); ++q) {
                                const qd m0 = off[0] + aff[0][0] * u[q] + aff[0][1] * v[q];
                                const qd m1 = off[1] + aff[1][0] * u[q] + aff[1][1] * v[q];
                                sum += w[q] * (qd)aff[bc][mc] * face_poly(k, bx, by, m0, m1) *
                                       face_poly(k, mx, my, u[q], v[q]);
                            }
                            M[(size_t)row * s + col] = sum;
                        }
            }
}

// transform = local * canonical^-1 (runtime route :296-333), row-major s x s
int face_transform_build(int k, const int perm[3], std::vector<xd> &T)
{
    const int s = k * (k - 1);
    std::vector<qd> C, L;
    static const int ident[3] = {1, 2, 3};
    face_moment_matrix(k, ident, C);
    face_moment_matrix(k, perm, L);
    std::vector<qd> Ct((size_t)s * s), X((size_t)s * s);
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) {
            Ct[(size_t)i * s + j] = C[(size_t)j * s + i];
            X[(size_t)i * s + j] = L[(size_t)j * s + i];
        }
    if (solve_dense(s, Ct, s, X)) return 1;
    T.resize((size_t)s * s);
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) T[(size_t)i * s + j] = (xd)X[(size_t)j * s + i];
    return 0;
}

// face transforms cached per (order, permutation)
int face_transform(int k, const int perm[3], std::vector<xd> &T)
{
    static std::map<std::tuple<int, int, int, int>, std::vector<xd>> cache;
    static std::mutex m;
    const auto key = std::make_tuple(k, perm[0], perm[1], perm[2]);
    {
        std::lock_guard<std::mutex> lock(m);
        auto it = cache.find(key);
        if (it != cache.end()) {
            T = it->second;
            return 0;
        }
    }
    if (face_transform_build(k, perm, T)) return 1;
    std::lock_guard<std::mutex> lock(m);
    cache[key] = T;
    return 0;
}

// Whitney functions (tetra_nedelec_first_order.f90:11-51)
void whitney(const xd p[3], xd *vals, xd *curls)
{
    static const xd G[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const xd lam[4] = {1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
    for (int e = 0; e < 6; ++e) {
        const int a = kEdgeV[e][0], b = kEdgeV[e][1];
        for (int c = 0; c < 3; ++c) vals[3 * e + c] = lam[a] * G[b][c] - lam[b] * G[a][c];
        curls[3 * e + 0] = 2 * (G[a][1] * G[b][2] - G[a][2] * G[b][1]);
        curls[3 * e + 1] = 2 * (G[a][2] * G[b][0] - G[a][0] * G[b][2]);
        curls[3 * e + 2] = 2 * (G[a][0] * G[b][1] - G[a][1] * G[b][0]);
    }
}

// The element rule: the 4-point degree-2 rule at k = 1, else the Duffy rule
// of degree 2k (curl-mass element) or 2k + 2 (weighted element and load).
void element_rule(int k, bool weighted_rule, std::vector<xd> &px, std::vector<xd> &py,
                  std::vector<xd> &pz, std::vector<xd> &pw)
{
    px.clear(), py.clear(), pz.clear(), pw.clear();
    if (k == 1) {
        // 4-point degree-2 rule of assemble_tetra_nedelec_element (:2364-2366)
        const xd a = (5 + 3 * std::sqrt((xd)5)) / 20, b = (5 - std::sqrt((xd)5)) / 20;
        const xd pts[4][3] = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
        for (int q = 0; q < 4; ++q) {
            px.push_back(pts[q][0]);
            py.push_back(pts[q][1]);
            pz.push_back(pts[q][2]);
            pw.push_back((xd)1 / 24);
        }
    } else {
        std::vector<qd> x, y, z, w;
        tet_rule(weighted_rule ? 2 * k + 2 : 2 * k, x, y, z, w);
        for (size_t q = 0; q < w.size(); ++q) {
            px.push_back((xd)x[q]);
            py.push_back((xd)y[q]);
            pz.push_back((xd)z[q]);
            pw.push_back((xd)w[q]);
        }
    }
}

// basis curls and values at points: B[q][j][0..2] curl, [3..5] value
int basis_at_points(int k, const std::vector<xd> &px, const std::vector<xd> &py,
                    const std::vector<xd> &pz, std::vector<xd> &B)
{
    const int n = nedelec_dof_count(k);
    const int nq = (int)px.size();
    B.assign((size_t)nq * n * 6, 0);
    if (k == 1) {
        for (int q = 0; q < nq; ++q) {
            const xd p[3] = {px[q], py[q], pz[q]};
            xd v[18], c[18];
            whitney(p, v, c);
            for (int j = 0; j < n; ++j)
                for (int d = 0; d < 3; ++d) {
                    B[((size_t)q * n + j) * 6 + d] = c[3 * j + d];
                    B[((size_t)q * n + j) * 6 + 3 + d] = v[3 * j + d];
                }
        }
    } else {
        const std::vector<qd> *C = coefficients_cached(k);
        if (!C) return 1;
        std::vector<xd> Cx(C->size());
        for (size_t i = 0; i < C->size(); ++i) Cx[i] = (xd)(*C)[i];
        parallel_for(nq, [&](int64_t lo, int64_t hi) {
            std::vector<qd> cv(3 * n), cc(3 * n);
            for (int64_t q = lo; q < hi; ++q) {
                const qd p[3] = {(qd)px[q], (qd)py[q], (qd)pz[q]};
                candidates<qd>(k, p, cv.data(), cc.data());
                std::vector<xd> cvx(3 * n), ccx(3 * n);
                for (int i = 0; i < 3 * n; ++i) cvx[i] = (xd)cv[i], ccx[i] = (xd)cc[i];
                for (int j = 0; j < n; ++j)
                    for (int d = 0; d < 3; ++d) {
                        xd sv = 0, sc = 0;
                        for (int i = 0; i < n; ++i) {
                            const xd a = Cx[(size_t)i * n + j];
                            sv += cvx[3 * i + d] * a;
                            sc += ccx[3 * i + d] * a;
                        }
                        B[((size_t)q * n + j) * 6 + d] = sc;
                        B[((size_t)q * n + j) * 6 + 3 + d] = sv;
                    }
            }
        }, 1);
    }
    return 0;
}

// 12 reference tensors (full n x n, row-major) at the element rule
int reference_tensors(int k, bool weighted_rule, std::vector<xd> &R)
{
    const int n = nedelec_dof_count(k);
    std::vector<xd> px, py, pz, pw;
    element_rule(k, weighted_rule, px, py, pz, pw);
    const int nq = (int)pw.size();
    std::vector<xd> B;
    if (basis_at_points(k, px, py, pz, B)) return 1;
    static const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    const size_t NN = (size_t)n * n;
    R.assign(12 * NN, 0);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t r = lo; r < hi; ++r)
            for (int c = (int)r; c < n; ++c) {
                xd acc[12] = {0};
                for (int q = 0; q < nq; ++q) {
                    const xd *br = &B[((size_t)q * n + r) * 6], *bc = &B[((size_t)q * n + c) * 6];
                    const xd w = pw[q];
                    for (int part = 0; part < 2; ++part)
                        for (int s = 0; s < 6; ++s) {
                            const int a = pa[s] + 3 * part, b = pb[s] + 3 * part;
                            xd t = br[a] * bc[b];
                            if (a != b) t += br[b] * bc[a];
                            acc[6 * part + s] += w * t;
                        }
                }
                for (int s = 0; s < 12; ++s) {
                    R[s * NN + (size_t)r * n + c] = acc[s];
                    R[s * NN + (size_t)c * n + r] = acc[s];
                }
            }
    }, 1);
    return 0;
}

// basis transform of a class (build_tetra_nedelec_basis_transform :107-155),
// dense row-major
int class_transform(int k, int cls, std::vector<xd> &T)
{
    const int n = nedelec_dof_count(k), fdc = k * (k - 1);
    const int *rk = class_table().ranks[cls];
    T.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) T[(size_t)i * n + i] = 1;
    for (int e = 0; e < 6; ++e) {
        if (rk[kEdgeV[e][0]] < rk[kEdgeV[e][1]]) continue; // positive orientation
        for (int d = 0; d < k; d += 2) T[(size_t)(e * k + d) * n + e * k + d] = -1;
    }
    if (fdc == 0) return 0;
    for (int f = 0; f < 4; ++f) {
        int perm[3];
        for (int i = 0; i < 3; ++i) {
            int pos = 1;
            for (int j = 0; j < 3; ++j) pos += rk[kFaceV[f][j]] < rk[kFaceV[f][i]];
            perm[i] = pos;
        }
        std::vector<xd> F;
        if (face_transform(k, perm, F)) return 1;
        const int start = 6 * k + f * fdc;
        for (int i = 0; i < fdc; ++i)
            for (int j = 0; j < fdc; ++j) T[(size_t)(start + i) * n + start + j] = F[(size_t)i * fdc + j];
    }
    return 0;
}

} // namespace

void nedelec_class_orientation(int cls, int edge_orient[6], int face_perm[12])
{
    const int *rk = class_table().ranks[cls];
    for (int e = 0; e < 6; ++e) edge_orient[e] = rk[kEdgeV[e][0]] < rk[kEdgeV[e][1]] ? 1 : -1;
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < 3; ++i) {
            int pos = 1;
            for (int j = 0; j < 3; ++j) pos += rk[kFaceV[f][j]] < rk[kFaceV[f][i]];
            face_perm[3 * f + i] = pos;
        }
}

int nedelec_basis_eval(int k, const double p[3], double *values, double *curls)
{
    const int n = nedelec_dof_count(k);
    if (k < 1 || k > 8) return 1;
    if (k == 1) {
        const xd px[3] = {p[0], p[1], p[2]};
        xd v[18], c[18];
        whitney(px, v, c);
        for (int i = 0; i < 18; ++i) values[i] = (double)v[i], curls[i] = (double)c[i];
        return 0;
    }
    const std::vector<qd> *C = coefficients_cached(k);
    if (!C) return 1;
    const qd pq[3] = {p[0], p[1], p[2]};
    std::vector<qd> cv(3 * n), cc(3 * n);
    candidates<qd>(k, pq, cv.data(), cc.data());
    for (int j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) {
            xd sv = 0, sc = 0;
            for (int i = 0; i < n; ++i) {
                const xd a = (xd)(*C)[(size_t)i * n + j];
                sv += (xd)cv[3 * i + d] * a;
                sc += (xd)cc[3 * i + d] * a;
            }
            values[3 * j + d] = (double)sv;
            curls[3 * j + d] = (double)sc;
        }
    return 0;
}

int nedelec_coefficients(int k, std::vector<double> &coeff)
{
    const int n = nedelec_dof_count(k);
    if (k == 1) {
        coeff.assign((size_t)n * n, 0.0);
        return 1;
    }
    const std::vector<qd> *C = coefficients_cached(k);
    if (!C) return 1;
    coeff.resize(C->size());
    for (size_t i = 0; i < C->size(); ++i) coeff[i] = (double)(*C)[i];
    (void)n;
    return 0;
}

void nedelec_class_tables(int k, bool weighted_rule, const std::vector<int> &classes,
                          std::vector<double> &out)
{
    const int n = nedelec_dof_count(k);
    const size_t np = (size_t)n * (n + 1) / 2, NN = (size_t)n * n;
    std::vector<xd> R;
    if (reference_tensors(k, weighted_rule, R)) throw std::runtime_error("Nedelec basis construction failed");
    out.assign(classes.size() * 12 * np, 0.0);
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        std::vector<xd> T;
        if (class_transform(k, classes[ci], T)) throw std::runtime_error("Nedelec face transform failed");
        // nonzeros of each column of T
        std::vector<std::vector<std::pair<int, xd>>> tcol(n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (T[(size_t)i * n + j] != 0) tcol[j].push_back({i, T[(size_t)i * n + j]});
        parallel_for(12, [&](int64_t lo, int64_t hi) {
            std::vector<xd> RT(NN);
            for (int64_t s = lo; s < hi; ++s) {
                const xd *Rs = R.data() + s * NN;
                // RT = R T
                for (int r = 0; r < n; ++r)
                    for (int c = 0; c < n; ++c) {
                        xd acc = 0;
                        for (const auto &t : tcol[c]) acc += Rs[(size_t)r * n + t.first] * t.second;
                        RT[(size_t)r * n + c] = acc;
                    }
                // T^T (R T), upper triangle
                double *o = out.data() + (ci * 12 + s) * np;
                size_t u = 0;
                for (int r = 0; r < n; ++r)
                    for (int c = r; c < n; ++c, ++u) {
                        xd acc = 0;
                        for (const auto &t : tcol[r]) acc += t.second * RT[(size_t)t.first * n + c];
                        o[u] = (double)acc;
                    }
            }
        }, 1);
    }
}

} // namespace ffh

namespace ffh {

void nedelec_load_rule(int k, std::vector<double> &pts)
{
    std::vector<xd> px, py, pz, pw;
    element_rule(k, true, px, py, pz, pw);
    pts.resize(4 * px.size());
    for (size_t q = 0; q < px.size(); ++q) {
        pts[4 * q + 0] = (double)px[q];
        pts[4 * q + 1] = (double)py[q];
        pts[4 * q + 2] = (double)pz[q];
        pts[4 * q + 3] = (double)pw[q];
    }
}

// Oriented reference values (part 3) or curls (part 0) of every class slot
// at the weighted element's rule: tab[((i * nq + q) * 3 + c) * ldn + r] =
// sum_j B_j,part+c(point q) T_jr, zero for r >= n.
static void point_tables(int k, const std::vector<int> &classes, int part, int ldn,
                         std::vector<double> &tab, std::vector<double> &pts)
{
    const int n = nedelec_dof_count(k);
    std::vector<xd> px, py, pz, pw, B;
    element_rule(k, true, px, py, pz, pw);
    const int nq = (int)pw.size();
    if (basis_at_points(k, px, py, pz, B)) throw std::runtime_error("Nedelec basis construction failed");
    pts.resize((size_t)4 * nq);
    for (int q = 0; q < nq; ++q) {
        pts[4 * q + 0] = (double)px[q];
        pts[4 * q + 1] = (double)py[q];
        pts[4 * q + 2] = (double)pz[q];
        pts[4 * q + 3] = (double)pw[q];
    }
    const size_t per = (size_t)nq * 3 * ldn;
    tab.assign(classes.size() * per, 0.0);
    parallel_for((int64_t)classes.size(), [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            std::vector<xd> T;
            if (class_transform(k, classes[i], T)) throw std::runtime_error("Nedelec face transform failed");
            double *out = tab.data() + (size_t)i * per;
            for (int q = 0; q < nq; ++q)
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < n; ++r) {
                        xd acc = 0;
                        for (int j = 0; j < n; ++j) {
                            const xd t = T[(size_t)j * n + r];
                            if (t != 0) acc += B[((size_t)q * n + j) * 6 + part + c] * t;
                        }
                        out[((size_t)q * 3 + c) * ldn + r] = (double)acc;
                    }
        }
    }, 1);
}

void nedelec_load_tables(int k, const std::vector<int> &classes, std::vector<double> &tab,
                         std::vector<double> &pts)
{
    point_tables(k, classes, 3, nedelec_dof_count(k), tab, pts);
}

void nedelec_curl_point_tables(int k, const std::vector<int> &classes, int ldn,
                               std::vector<double> &tab, std::vector<double> &pts)
{
    point_tables(k, classes, 0, ldn, tab, pts);
}

} // namespace ffh
