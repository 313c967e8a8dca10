// ffh_nedelec.hpp -- host side of the tetrahedral first-kind Nedelec family:
// global numbering, orientation classes and the reference tensors of the
// curl-mass element in every orientation class.  Plain C++ (no HIP types):
// the implementation is compiled by the host compiler because it uses IEEE
// quad precision.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace ffh {

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &f,
                  int64_t grain = 4096);

int nedelec_dof_count(int k);

// Orientation class of a cell: the rank pattern of its four vertex ids
// (0..23).  Edge signs and face permutations of the reference's basis
// transform depend on nothing else.
constexpr int kNedelecClasses = 24;

// orientation class of every cell (vertex-id ranks), as nedelec_dof_map
void nedelec_cell_classes(int ntet, const int *tets, std::vector<unsigned char> &cls);
// build_tetra_nedelec_dof_map (src/elements/tetra_nedelec_global_dof_map.f90:
// 15-105) with O(1) hashed lookups: gd (ntet x n, 0-based), cls (ntet).
// Returns 0, or 1 for invalid cells (ids < 1 or repeated in a cell).
int nedelec_dof_map(int k, int ntet, const int *tets, std::vector<int> &gd,
                    std::vector<unsigned char> &cls, int &ndofs, int &nedges, int &nfaces);

// Constrained dofs of the tangential trace on the boundary
// (build_zero_tangential_mask, src/solvers/fortfem_tetra_nedelec_solver_3d.f90:
// 431-470): every dof of an edge or face that lies on a boundary face.
// bf[4 e + o] flags the boundary face of cell e opposite its local vertex o
// (a face of one cell, found on the device).
void nedelec_boundary_dofs(int k, int ntet, const std::vector<unsigned char> &bf,
                           const std::vector<int> &gd, std::vector<unsigned char> &constrained);

// Reference tensors of the curl-mass element, transformed into orientation
// classes.  For class slot i (classes[i]) and term s, the packed upper
// triangle (r <= c, row-major) of T^T R_s T is stored at
// out[((size_t)i * 12 + s) * npairs + u].  Terms 0..5: curl tensors
// C^00, C^11, C^22, C^01+C^10, C^02+C^20, C^12+C^21 with
// C^ab_rc = sum_q w_q curl^_ra curl^_cb; terms 6..11 the same for values.
// weighted_rule selects the quadrature of the weighted element (2k+2)
// instead of the curl-mass element (2k); k = 1 uses the 4-point rule.
void nedelec_class_tables(int k, bool weighted_rule, const std::vector<int> &classes,
                          std::vector<double> &out);

// Load tables of assemble_tetra_nedelec_vector_load(_order)
// (src/assembly/assembly_tetra_nedelec_3d.f90:71-127, 525-582, 2002-2040,
// 2516-2553): the reference's load rule (4-point degree-2 rule at k = 1,
// Duffy 2k + 2 otherwise) as pts (x, y, z, w per point) and, for class slot i,
// the oriented reference values tab[((i * nq + q) * 3 + c) * n + r] =
// sum_j value_j,c(point q) T_jr, with T the class's basis transform, so that
// the oriented element load is T^T b = sum_(q,c) g_qc tab[q][c][r] with
// g_q = det w_q J^-1 f(x_q).
void nedelec_load_tables(int k, const std::vector<int> &classes, std::vector<double> &tab,
                         std::vector<double> &pts);
// Oriented reference curls at the same rule (the weighted element's rule of
// assemble_tetra_nedelec_weighted_element(_order), assembly_tetra_nedelec_3d.f90:
// 2406-2514): tab[((i * nq + q) * 3 + a) * ldn + r] = sum_j curl_j,a(point q)
// T_jr for class slot i, zero for n <= r < ldn.
void nedelec_curl_point_tables(int k, const std::vector<int> &classes, int ldn,
                               std::vector<double> &tab, std::vector<double> &pts);
// the load points alone (x, y, z, w per point)
void nedelec_load_rule(int k, std::vector<double> &pts);

// Edge orientations (+1 / -1 per local edge) and face permutations (position
// of each local face vertex in the sorted triple, 1..3) of a class.
void nedelec_class_orientation(int cls, int edge_orient[6], int face_perm[12]);

// evaluate_tetra_nedelec_first_kind (tetra_nedelec_arbitrary_order.f90:75-106)
// at a reference point: values / curls (3 x n, component fastest).
int nedelec_basis_eval(int k, const double p[3], double *values, double *curls);

// Moment-dual coefficients (n x n, row-major, candidate i -> basis j),
// rounded to double; for tests of the construction.
int nedelec_coefficients(int k, std::vector<double> &coeff);

} // namespace ffh
