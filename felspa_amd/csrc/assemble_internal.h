// The generated Stokes system as the host generator (assemble.cpp) and the
// device assembler (assemble_gpu.hip) both see it, plus the reference-element
// tables and field description they share so that the two assemble from the
// same numbers in the same order.
#pragma once
#include <cstdint>
#include <vector>

#include "fs_assemble.h"
#include "fs_common.h"

struct fs_system {
  int dim = 0, N = 0;
  int64_t n_u = 0, n_p = 0;
  fs_problem prob{};
  fs::HostCsr A, Bt, B, K11, Mp;
  std::vector<double> rhs;
  std::vector<int8_t> kind;
  std::vector<int64_t> master;
  std::vector<double> cval;
  std::vector<int64_t> vnode, pnode;
  std::vector<int8_t> vcomp;
  std::vector<float> log_eta;
};

namespace fs {

// Q2 velocity / Q1 pressure reference element on a cell of size h with the
// 3^d Gauss points (QGauss(degree_v + 1), stokes_common.implement.h:168).
//   N[q][n], dN[q][n][k] (physical gradient), P[q][j], JxW[q], xq[q][k]
//   Bl[j][(n,c)] = -sum_q psi_j dN_n/dx_c JxW   (the viscosity-free B block)
// Node n has lattice offsets (n%3, n/3%3, n/9%3); pressure node j has
// (j%2, j/2%2, j/4%2); quadrature point q has Gauss indices (q%3, q/3%3, q/9%3).
struct RefTables {
  int d = 0, nv = 0, np = 0, nq = 0;
  std::vector<double> N, dN, P, JxW, xq, Bl;
};
RefTables ref_tables(int d, double h);

// Seeded ellipse / ellipsoid inclusions of FS_FIELD_ELLIPSES with the factor
// 10^logc already evaluated on the host (std::pow), so that device and host
// viscosities agree bit for bit.
struct FieldEllipsoid {
  double c[3], ax[3], R[3][3];
  double logc, scale;
};
std::vector<FieldEllipsoid> field_ellipsoids(const fs_problem& p);

// The layered field's interface height 0.5 + amp cos(2 pi x0).
double layer_interface(const fs_problem& p, double x0);

// Row chunk of the deterministic two-level sum behind mean |A_ii|: the sum
// over each chunk of free velocity rows, in row order, then over the chunks.
constexpr int64_t DIAG_CHUNK = 256;

}  // namespace fs
