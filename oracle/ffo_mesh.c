/*
 * ffo_mesh.c -- oracle restatement of the reference meshes and the 3x3
 * determinant / inverse helpers.  Test infrastructure only (see ffo.h).
 */
#include "ffo.h"

#include <math.h>
#include <string.h>

/* fortnum det3 (called at src/mesh/structured_tetra_box_mesh.f90:65 and
 * src/assembly/assembly_tetra_lagrange_arbitrary_order_3d.f90:64):
 * cofactor expansion along the first row.  a is column-major a(r,c)=a[r+3c]. */
double ffo_det3(const double a[9])
{
#define A(r, c) a[(r) + 3 * (c)]
    return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) -
           A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
           A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
}

/* fortnum inv3 (assembly_tetra_lagrange_arbitrary_order_3d.f90:67):
 * adjugate / determinant; status 1 for a singular matrix. */
int ffo_inv3(const double a[9], double inv[9])
{
    double det = ffo_det3(a);
    if (det == 0.0) return 1;
    double r = 1.0 / det;
#define I(r_, c_) inv[(r_) + 3 * (c_)]
    I(0, 0) = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * r;
    I(0, 1) = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * r;
    I(0, 2) = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * r;
    I(1, 0) = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * r;
    I(1, 1) = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * r;
    I(1, 2) = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * r;
    I(2, 0) = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * r;
    I(2, 1) = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * r;
    I(2, 2) = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * r;
#undef I
#undef A
    return 0;
}

static int vertex_index(int i, int j, int k, const int counts[3])
{
    /* src/mesh/structured_tetra_box_mesh.f90:72-76 */
    return 1 + i + (counts[0] + 1) * (j + (counts[1] + 1) * k);
}

/* generate_structured_tetra_box_mesh
 * (src/mesh/structured_tetra_box_mesh.f90:12-70, add_brick_tetrahedra
 * :78-94).  bounds is bounds(3,2) column-major: lower xyz then upper xyz. */
int ffo_box_mesh(const double bounds[6], const int counts[3],
                 double *vertices, int *tetrahedra)
{
    for (int c = 0; c < 3; ++c) {
        if (counts[c] < 1) return 1;
        if (bounds[3 + c] <= bounds[c]) return 1;
    }
    for (int k = 0; k <= counts[2]; ++k)
        for (int j = 0; j <= counts[1]; ++j)
            for (int i = 0; i <= counts[0]; ++i) {
                int v = vertex_index(i, j, k, counts) - 1;
                int ijk[3] = {i, j, k};
                for (int c = 0; c < 3; ++c) {
                    double coord = (double)ijk[c] / (double)counts[c];
                    vertices[3 * v + c] =
                        bounds[c] + coord * (bounds[3 + c] - bounds[c]);
                }
            }
    static const int local[6][4] = {{1, 2, 4, 8}, {1, 2, 6, 8}, {1, 3, 4, 8},
                                     {1, 3, 7, 8}, {1, 5, 6, 8}, {1, 5, 7, 8}};
    long cell = 0;
    for (int k = 0; k < counts[2]; ++k)
        for (int j = 0; j < counts[1]; ++j)
            for (int i = 0; i < counts[0]; ++i) {
                int brick[8] = {vertex_index(i, j, k, counts),
                                vertex_index(i + 1, j, k, counts),
                                vertex_index(i, j + 1, k, counts),
                                vertex_index(i + 1, j + 1, k, counts),
                                vertex_index(i, j, k + 1, counts),
                                vertex_index(i + 1, j, k + 1, counts),
                                vertex_index(i, j + 1, k + 1, counts),
                                vertex_index(i + 1, j + 1, k + 1, counts)};
                for (int t = 0; t < 6; ++t, ++cell)
                    for (int n = 0; n < 4; ++n)
                        tetrahedra[4 * cell + n] = brick[local[t][n] - 1];
            }
    /* orientation fix: swap local vertices 3 and 4 when det < 0 (:56-68) */
    for (long e = 0; e < cell; ++e) {
        int *t = tetrahedra + 4 * e;
        double jac[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r)
                jac[r + 3 * c] = vertices[3 * (t[c + 1] - 1) + r] -
                                 vertices[3 * (t[0] - 1) + r];
        if (ffo_det3(jac) < 0.0) {
            int tmp = t[2];
            t[2] = t[3];
            t[3] = tmp;
        }
    }
    return 0;
}

/* unit_square_mesh(n) (src/fortfem_api_mesh.f90:128-139) ->
 * create_rectangular (src/mesh/mesh_2d.f90:90-148) and find_boundary
 * (:553-603).  On the rectangular grid an edge belongs to exactly one triangle
 * iff it lies on the outer boundary, so a vertex is a boundary vertex iff its
 * grid index is on the first/last row or column. */
int ffo_unit_square_mesh(int n, double *vertices, int *triangles,
                         int *is_boundary_vertex)
{
    if (n < 2) return 1;
    double dx = 1.0 / (double)(n - 1), dy = 1.0 / (double)(n - 1);
    int k = 0;
    for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= n; ++i, ++k) {
            vertices[2 * k + 0] = 0.0 + (i - 1) * dx;
            vertices[2 * k + 1] = 0.0 + (j - 1) * dy;
            if (is_boundary_vertex)
                is_boundary_vertex[k] = (i == 1 || i == n || j == 1 || j == n);
        }
    k = 0;
    for (int j = 1; j <= n - 1; ++j)
        for (int i = 1; i <= n - 1; ++i) {
            int v1 = (j - 1) * n + i, v2 = (j - 1) * n + i + 1;
            int v3 = j * n + i + 1, v4 = j * n + i;
            triangles[3 * k + 0] = v1;
            triangles[3 * k + 1] = v2;
            triangles[3 * k + 2] = v3;
            ++k;
            triangles[3 * k + 0] = v1;
            triangles[3 * k + 1] = v3;
            triangles[3 * k + 2] = v4;
            ++k;
        }
    return 0;
}
