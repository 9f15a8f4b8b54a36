// Q2-Q1 Stokes system generator on uniform meshes of the unit square / cube.
//
// Restates FELSPA's assembly of the blocks that StokesLinearSystem::solve
// consumes (src/pde/stokes_ilu.cc:739-821 local assembly, 260-292 and
// include/felspa/pde/src/linear_systems.implement.h:185-220 sparsity,
// src/pde/stokes_ilu.cc:310-403 constraints) for structured grids, so the
// solver and the oracle see identical matrices. deal.II's own
// AffineConstraints condensation, DoFRenumbering tie-breaking and periodic
// bookkeeping are third-party internals; the choices made here are:
//   * constrained rows keep only their diagonal (keep_constrained_dofs=false);
//     the diagonal is the value the dof's own cells would have put there and
//     the rhs is diagonal*value, so the solve returns the constrained value;
//   * Dirichlet columns are eliminated into the rhs;
//   * periodic slave dofs (x = 1) are merged into their masters (x = 0);
//   * the pinned pressure dof (first pressure dof) and pressure slaves get
//     K11 diagonal = mean |A_ii| over free velocity rows;
//   * Cuthill-McKee on the node lattice graph from node 0, neighbours sorted
//     by (degree, lattice index); all components of a node are consecutive.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>

#include "assemble_internal.h"
#include "fs_assemble.h"
#include "fs_common.h"
#include "fs_stokes.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using fs::HostCsr;

// ---------------------------------------------------------------------------
// PRNG: SplitMix64-seeded xoshiro256** (published by Blackman & Vigna).
struct Xoshiro256ss {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

}  // namespace

namespace fs {

std::vector<FieldEllipsoid> field_ellipsoids(const fs_problem& p) {
  std::vector<FieldEllipsoid> ell;
  if (p.field != FS_FIELD_ELLIPSES) return ell;
  Xoshiro256ss rng(p.seed);
  int nmin = std::max(1, p.n_ellipses_min), nmax = std::max(nmin, p.n_ellipses_max);
  int n = nmin + (int)(rng.next() % (uint64_t)(nmax - nmin + 1));
  for (int e = 0; e < n; ++e) {
    FieldEllipsoid E{};
    for (int k = 0; k < p.dim; ++k) E.c[k] = rng.uniform();
    for (int k = 0; k < p.dim; ++k) E.ax[k] = rng.uniform(p.axis_min, p.axis_max);
    if (p.dim == 2) {
      double th = rng.uniform(0.0, M_PI);
      double cs = std::cos(th), sn = std::sin(th);
      E.R[0][0] = cs; E.R[0][1] = sn; E.R[0][2] = 0;
      E.R[1][0] = -sn; E.R[1][1] = cs; E.R[1][2] = 0;
      E.R[2][0] = 0; E.R[2][1] = 0; E.R[2][2] = 1;
    } else {
      // uniform random rotation from a uniform unit quaternion (Shoemake 1992)
      double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
      double a = std::sqrt(1 - u1), b = std::sqrt(u1);
      double w = a * std::sin(2 * M_PI * u2), x = a * std::cos(2 * M_PI * u2);
      double y = b * std::sin(2 * M_PI * u3), z = b * std::cos(2 * M_PI * u3);
      E.R[0][0] = 1 - 2 * (y * y + z * z); E.R[0][1] = 2 * (x * y - z * w); E.R[0][2] = 2 * (x * z + y * w);
      E.R[1][0] = 2 * (x * y + z * w); E.R[1][1] = 1 - 2 * (x * x + z * z); E.R[1][2] = 2 * (y * z - x * w);
      E.R[2][0] = 2 * (x * z - y * w); E.R[2][1] = 2 * (y * z + x * w); E.R[2][2] = 1 - 2 * (x * x + y * y);
    }
    E.logc = rng.uniform(p.logc_min, p.logc_max);
    E.scale = std::pow(10.0, E.logc);
    ell.push_back(E);
  }
  return ell;
}

double layer_interface(const fs_problem& p, double x0) { return 0.5 + p.amp * std::cos(2 * M_PI * x0); }

}  // namespace fs

namespace {

struct Field {
  fs_problem p;
  std::vector<fs::FieldEllipsoid> ell;

  void init() { ell = fs::field_ellipsoids(p); }

  // viscosity and density at a point
  void eval(const double* x, double& eta, double& rho) const {
    const int d = p.dim;
    switch (p.field) {
      case FS_FIELD_INCLUSION: {
        // SphericalInclusion::scalar_values (viscous_flow.implement.h:42-72):
        // inside the sphere -> inclusion value; else if |x| > r -> matrix
        // value; else the average (the reference compares |x| with r).
        double dist2 = 0, norm2 = 0;
        for (int k = 0; k < d; ++k) {
          dist2 += (x[k] - 0.5) * (x[k] - 0.5);
          norm2 += x[k] * x[k];
        }
        double dist = std::sqrt(dist2), nrm = std::sqrt(norm2);
        if (dist < p.radius) {
          eta = p.eta_incl; rho = p.rho_incl;
        } else if (nrm > p.radius) {
          eta = p.eta_matrix; rho = p.rho_matrix;
        } else {
          eta = 0.5 * (p.eta_matrix + p.eta_incl);
          rho = 0.5 * (p.rho_matrix + p.rho_incl);
        }
        return;
      }
      case FS_FIELD_LAYERS: {
        double yif = fs::layer_interface(p, x[0]);
        if (x[d - 1] > yif) { eta = p.eta_incl; rho = p.rho_incl; }
        else { eta = p.eta_matrix; rho = p.rho_matrix; }
        return;
      }
      case FS_FIELD_ELLIPSES: {
        double best = -1.0, scale = 1.0;
        for (const auto& E : ell) {
          double r2 = 0;
          for (int a = 0; a < d; ++a) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += E.R[a][k] * (x[k] - E.c[k]);
            r2 += (s / E.ax[a]) * (s / E.ax[a]);
          }
          if (r2 <= 1.0 && E.logc > best) { best = E.logc; scale = E.scale; }
        }
        eta = best >= 0 ? p.eta_matrix * scale : p.eta_matrix;
        rho = p.rho_matrix;
        return;
      }
      default:
        eta = p.eta_matrix; rho = p.rho_matrix;
        return;
    }
  }
};

// ---------------------------------------------------------------------------
// Lattice helpers
struct Lattice {
  int d;
  int64_t n1;  // nodes per direction
  int64_t size() const { int64_t s = 1; for (int k = 0; k < d; ++k) s *= n1; return s; }
  void coords(int64_t idx, int64_t* c) const {
    for (int k = 0; k < d; ++k) { c[k] = idx % n1; idx /= n1; }
  }
  int64_t index(const int64_t* c) const {
    int64_t idx = 0;
    for (int k = d - 1; k >= 0; --k) idx = idx * n1 + c[k];
    return idx;
  }
};

// neighbour range of a lattice coordinate for the node graph (nodes sharing a cell)
inline void nbr_range(int64_t i, int64_t n1, bool q2, int64_t& lo, int64_t& hi) {
  if (q2) {
    int64_t r = (i & 1) ? 1 : 2;
    lo = std::max<int64_t>(0, i - r);
    hi = std::min<int64_t>(n1 - 1, i + r);
  } else {
    lo = std::max<int64_t>(0, i - 1);
    hi = std::min<int64_t>(n1 - 1, i + 1);
  }
}

template <class F>
inline void for_each_nbr(const Lattice& L, int64_t node, bool q2, F&& f) {
  int64_t c[3], lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  L.coords(node, c);
  for (int k = 0; k < L.d; ++k) nbr_range(c[k], L.n1, q2, lo[k], hi[k]);
  int64_t q[3];
  for (q[2] = lo[2]; q[2] <= hi[2]; ++q[2])
    for (q[1] = lo[1]; q[1] <= hi[1]; ++q[1])
      for (q[0] = lo[0]; q[0] <= hi[0]; ++q[0]) f(L.index(q));
}

inline int64_t nbr_degree(const Lattice& L, int64_t node, bool q2) {
  int64_t c[3], deg = 1;
  L.coords(node, c);
  for (int k = 0; k < L.d; ++k) {
    int64_t lo, hi;
    nbr_range(c[k], L.n1, q2, lo, hi);
    deg *= (hi - lo + 1);
  }
  return deg - 1;
}

// Cuthill-McKee on the lattice node graph: BFS from node 0, unvisited
// neighbours appended in order of increasing degree (ties: lattice index).
std::vector<int64_t> cuthill_mckee(const Lattice& L, bool q2) {
  const int64_t n = L.size();
  std::vector<int64_t> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<std::pair<int64_t, int64_t>> nb;
  order.push_back(0);
  seen[0] = 1;
  for (int64_t head = 0; head < (int64_t)order.size(); ++head) {
    int64_t u = order[head];
    nb.clear();
    for_each_nbr(L, u, q2, [&](int64_t v) {
      if (!seen[v]) { seen[v] = 1; nb.emplace_back(nbr_degree(L, v, q2), v); }
    });
    std::sort(nb.begin(), nb.end());
    for (auto& e : nb) order.push_back(e.second);
  }
  return order;
}

}  // namespace

// ---------------------------------------------------------------------------
// Reference-element data on the unit cell, scaled to cell size h.
namespace fs {

RefTables ref_tables(int d, double h) {
  RefTables R;
  R.d = d;
  const double g = std::sqrt(3.0 / 5.0) / 2.0;
  const double qp[3] = {0.5 - g, 0.5, 0.5 + g};
  const double qw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
  int nv = 1, np = 1, nq = 1;
  for (int k = 0; k < d; ++k) { nv *= 3; np *= 2; nq *= 3; }
  R.nv = nv; R.np = np; R.nq = nq;
  R.N.resize((size_t)nq * nv);
  R.dN.resize((size_t)nq * nv * d);
  R.P.resize((size_t)nq * np);
  R.JxW.resize(nq);
  R.xq.resize((size_t)nq * d);
  auto L2 = [](int a, double x) {
    return a == 0 ? 2 * (x - 0.5) * (x - 1) : a == 1 ? -4 * x * (x - 1) : 2 * x * (x - 0.5);
  };
  auto dL2 = [](int a, double x) { return a == 0 ? 4 * x - 3 : a == 1 ? -8 * x + 4 : 4 * x - 1; };
  auto L1 = [](int b, double x) { return b == 0 ? 1 - x : x; };
  double hd = 1;
  for (int k = 0; k < d; ++k) hd *= h;
  for (int q = 0; q < nq; ++q) {
    int qi[3] = {q % 3, (q / 3) % 3, (q / 9) % 3};
    double x[3] = {0, 0, 0}, w = 1;
    for (int k = 0; k < d; ++k) { x[k] = qp[qi[k]]; w *= qw[qi[k]]; R.xq[(size_t)q * d + k] = x[k]; }
    R.JxW[q] = w * hd;
    for (int n = 0; n < nv; ++n) {
      int a[3] = {n % 3, (n / 3) % 3, (n / 9) % 3};
      double val = 1;
      for (int k = 0; k < d; ++k) val *= L2(a[k], x[k]);
      R.N[(size_t)q * nv + n] = val;
      for (int c = 0; c < d; ++c) {
        double gv = 1;
        for (int k = 0; k < d; ++k) gv *= (k == c) ? dL2(a[k], x[k]) / h : L2(a[k], x[k]);
        R.dN[((size_t)q * nv + n) * d + c] = gv;
      }
    }
    for (int m = 0; m < np; ++m) {
      int b[3] = {m % 2, (m / 2) % 2, (m / 4) % 2};
      double val = 1;
      for (int k = 0; k < d; ++k) val *= L1(b[k], x[k]);
      R.P[(size_t)q * np + m] = val;
    }
  }
  // Bl[j][(n,c)] = -sum_q psi_j dN_n/dx_c JxW  (viscosity independent)
  const int nvl = nv * d;
  R.Bl.assign((size_t)np * nvl, 0.0);
  for (int q = 0; q < nq; ++q)
    for (int j = 0; j < np; ++j)
      for (int n = 0; n < nv; ++n)
        for (int c = 0; c < d; ++c)
          R.Bl[(size_t)j * nvl + n * d + c] -=
            R.P[(size_t)q * np + j] * R.dN[((size_t)q * nv + n) * d + c] * R.JxW[q];
  return R;
}

}  // namespace fs

namespace {

void build(const fs_problem& prob, fs_system& S) {
  const int d = prob.dim;
  const int64_t Nc = prob.n_cells;
  const double h = 1.0 / (double)Nc;
  Field F;
  F.p = prob;
  F.init();

  Lattice LV{d, 2 * Nc + 1}, LP{d, Nc + 1};
  const int64_t nvn = LV.size(), npn = LP.size();
  S.dim = d;
  S.N = (int)Nc;
  S.prob = prob;
  S.n_u = nvn * d;
  S.n_p = npn;
  const int64_t n_u = S.n_u, n_p = S.n_p, ntot = n_u + n_p;

  // dof numbering
  std::vector<int64_t> vorder = cuthill_mckee(LV, true), porder = cuthill_mckee(LP, false);
  std::vector<int64_t> vpos(nvn), ppos(npn);
  for (int64_t i = 0; i < nvn; ++i) vpos[vorder[i]] = i;
  for (int64_t i = 0; i < npn; ++i) ppos[porder[i]] = i;
  S.vnode.resize(n_u);
  S.vcomp.resize(n_u);
  for (int64_t i = 0; i < nvn; ++i)
    for (int c = 0; c < d; ++c) {
      S.vnode[i * d + c] = vorder[i];
      S.vcomp[i * d + c] = (int8_t)c;
    }
  S.pnode = porder;
  auto vdof = [&](int64_t node, int c) { return vpos[node] * d + c; };
  auto pdof = [&](int64_t node) { return n_u + ppos[node]; };  // global index in [n_u, ntot)

  // constraints
  S.kind.assign(ntot, FS_DOF_FREE);
  S.master.resize(ntot);
  S.cval.assign(ntot, 0.0);
  for (int64_t i = 0; i < ntot; ++i) S.master[i] = i;
  const int64_t VE = 2 * Nc, PE = Nc;
  const bool periodic = prob.bc == FS_BC_TOPBOTTOM_SHEAR_PERIODIC;
  if (periodic && (Nc % 2)) throw std::runtime_error("periodic meshes need an even cell count");
  for (int64_t node = 0; node < nvn; ++node) {
    int64_t c[3];
    LV.coords(node, c);
    double x[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) x[k] = (double)c[k] / (double)VE;
    bool onb = false;
    for (int k = 0; k < d; ++k) onb |= (c[k] == 0 || c[k] == VE);
    switch (prob.bc) {
      case FS_BC_TOPBOTTOM_SHEAR_PERIODIC: {
        if (c[d - 1] == 0 || c[d - 1] == VE) {
          for (int k = 0; k < d; ++k) {
            int64_t g = vdof(node, k);
            S.kind[g] = FS_DOF_DIRICHLET;
            S.cval[g] = k == 0 ? (c[d - 1] == 0 ? -prob.shear_velocity : prob.shear_velocity) : 0.0;
          }
        } else if (c[0] == VE) {
          int64_t cm[3] = {0, c[1], c[2]};
          int64_t mnode = LV.index(cm);
          for (int k = 0; k < d; ++k) {
            int64_t g = vdof(node, k);
            S.kind[g] = FS_DOF_PERIODIC;
            S.master[g] = vdof(mnode, k);
          }
        }
        break;
      }
      case FS_BC_FREE_SLIP: {
        for (int k = 0; k < d; ++k)
          if (c[k] == 0 || c[k] == VE) {
            int64_t g = vdof(node, k);
            S.kind[g] = FS_DOF_DIRICHLET;
            S.cval[g] = 0.0;
          }
        break;
      }
      case FS_BC_LINEAR_SHEAR: {
        if (onb)
          for (int k = 0; k < d; ++k) {
            int64_t g = vdof(node, k);
            S.kind[g] = FS_DOF_DIRICHLET;
            S.cval[g] = k == 0 ? (x[d - 1] - 0.5) / 0.5 * prob.shear_velocity : 0.0;
          }
        break;
      }
      default:
        throw std::runtime_error("unknown boundary condition");
    }
  }
  if (periodic)
    for (int64_t node = 0; node < npn; ++node) {
      int64_t c[3];
      LP.coords(node, c);
      if (c[0] == PE) {
        int64_t cm[3] = {0, c[1], c[2]};
        int64_t g = pdof(node);
        S.kind[g] = FS_DOF_PERIODIC;
        S.master[g] = pdof(LP.index(cm));
      }
    }
  {
    // pin the first pressure dof (StokesLinearSystem::setup_bc_constraints)
    int64_t g = n_u;  // pressure dof 0
    if (S.kind[g] == FS_DOF_FREE) { S.kind[g] = FS_DOF_PINNED; S.cval[g] = 0.0; }
  }
  // equation index of each dof: -1 if Dirichlet/pinned (eliminated), master if periodic
  std::vector<int64_t> eq(ntot);
  for (int64_t g = 0; g < ntot; ++g) {
    int8_t k = S.kind[g];
    eq[g] = (k == FS_DOF_DIRICHLET || k == FS_DOF_PINNED) ? -1 : g;
  }
  // a periodic slave follows its master's equation (-1 when the master is
  // itself eliminated, e.g. the pinned pressure dof)
  for (int64_t g = 0; g < ntot; ++g)
    if (S.kind[g] == FS_DOF_PERIODIC) eq[g] = eq[S.master[g]];

  // master node maps (node-level), slave lists
  auto vmaster_node = [&](int64_t node) -> int64_t {
    int64_t c[3];
    LV.coords(node, c);
    if (periodic && c[0] == VE && !(c[d - 1] == 0 || c[d - 1] == VE)) {
      int64_t cm[3] = {0, c[1], c[2]};
      return LV.index(cm);
    }
    return node;
  };
  auto pmaster_node = [&](int64_t node) -> int64_t {
    int64_t c[3];
    LP.coords(node, c);
    if (periodic && c[0] == PE) { int64_t cm[3] = {0, c[1], c[2]}; return LP.index(cm); }
    return node;
  };
  // nodes whose equations are merged into `node` (itself + periodic slave)
  auto vsources = [&](int64_t node, int64_t* out) -> int {
    int n = 0;
    out[n++] = node;
    int64_t c[3];
    LV.coords(node, c);
    if (periodic && c[0] == 0 && !(c[d - 1] == 0 || c[d - 1] == VE)) {
      int64_t cs[3] = {VE, c[1], c[2]};
      out[n++] = LV.index(cs);
    }
    return n;
  };
  auto psources = [&](int64_t node, int64_t* out) -> int {
    int n = 0;
    out[n++] = node;
    int64_t c[3];
    LP.coords(node, c);
    if (periodic && c[0] == 0) { int64_t cs[3] = {PE, c[1], c[2]}; out[n++] = LP.index(cs); }
    return n;
  };

  // ---------------- sparsity patterns ----------------
  // A: per velocity node, the sorted list of master neighbour nodes (by vpos)
  auto build_rows = [&](int64_t nrows, auto&& row_cols, HostCsr& M, int64_t ncols) {
    M.n_rows = nrows;
    M.n_cols = ncols;
    M.rp.assign(nrows + 1, 0);
    std::vector<std::vector<int32_t>> tmp(nrows);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < nrows; ++r) {
      std::vector<int32_t> cols;
      row_cols(r, cols);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      tmp[r].swap(cols);
    }
    for (int64_t r = 0; r < nrows; ++r) M.rp[r + 1] = M.rp[r] + (int64_t)tmp[r].size();
    M.ci.resize(M.rp[nrows]);
    M.v.assign(M.rp[nrows], 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrows; ++r)
      std::copy(tmp[r].begin(), tmp[r].end(), M.ci.begin() + M.rp[r]);
  };

  // velocity rows
  build_rows(n_u, [&](int64_t r, std::vector<int32_t>& cols) {
    if (eq[r] != r) { cols.push_back((int32_t)r); return; }
    int64_t node = S.vnode[r];
    int64_t src[2];
    int ns = vsources(node, src);
    for (int s = 0; s < ns; ++s)
      for_each_nbr(LV, src[s], true, [&](int64_t m) {
        int64_t mm = vmaster_node(m);
        for (int e = 0; e < d; ++e) {
          int64_t g = eq[vdof(mm, e)];
          if (g >= 0) cols.push_back((int32_t)g);
        }
      });
  }, S.A, n_u);
  // B rows (pressure rows, velocity columns)
  build_rows(n_p, [&](int64_t q, std::vector<int32_t>& cols) {
    int64_t g = n_u + q;
    if (eq[g] != g) return;
    int64_t node = S.pnode[q];
    int64_t src[2];
    int ns = psources(node, src);
    for (int s = 0; s < ns; ++s) {
      int64_t c[3];
      LP.coords(src[s], c);
      // cells containing this pressure node; their velocity nodes
      int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      for (int k = 0; k < d; ++k) {
        lo[k] = std::max<int64_t>(0, 2 * (c[k] - 1));
        hi[k] = std::min<int64_t>(VE, 2 * (c[k] + 1));
      }
      int64_t qv[3];
      for (qv[2] = lo[2]; qv[2] <= hi[2]; ++qv[2])
        for (qv[1] = lo[1]; qv[1] <= hi[1]; ++qv[1])
          for (qv[0] = lo[0]; qv[0] <= hi[0]; ++qv[0]) {
            int64_t mm = vmaster_node(LV.index(qv));
            for (int e = 0; e < d; ++e) {
              int64_t gg = eq[vdof(mm, e)];
              if (gg >= 0) cols.push_back((int32_t)gg);
            }
          }
    }
  }, S.B, n_u);
  // Mp rows
  build_rows(n_p, [&](int64_t q, std::vector<int32_t>& cols) {
    int64_t g = n_u + q;
    if (eq[g] != g) { cols.push_back((int32_t)q); return; }
    int64_t src[2];
    int ns = psources(S.pnode[q], src);
    for (int s = 0; s < ns; ++s)
      for_each_nbr(LP, src[s], false, [&](int64_t m) {
        int64_t gg = eq[pdof(pmaster_node(m))];
        if (gg >= 0) cols.push_back((int32_t)(gg - n_u));
      });
  }, S.Mp, n_p);
  // K11: diagonal-only pattern
  S.K11.n_rows = S.K11.n_cols = n_p;
  S.K11.rp.resize(n_p + 1);
  S.K11.ci.resize(n_p);
  S.K11.v.assign(n_p, 0.0);
  for (int64_t q = 0; q <= n_p; ++q) S.K11.rp[q] = q;
  for (int64_t q = 0; q < n_p; ++q) S.K11.ci[q] = (int32_t)q;

  // ---------------- values ----------------
  const fs::RefTables R = fs::ref_tables(d, h);
  const int nvl = R.nv * d;  // local velocity dofs
  // G_q[(n,c),(m,e)] = (delta_ce gradN_n.gradN_m + dN_n/dx_e dN_m/dx_c) JxW
  std::vector<double> G((size_t)R.nq * nvl * nvl);
  for (int q = 0; q < R.nq; ++q)
    for (int n = 0; n < R.nv; ++n)
      for (int c = 0; c < d; ++c)
        for (int m = 0; m < R.nv; ++m)
          for (int e = 0; e < d; ++e) {
            const double* gn = &R.dN[((size_t)q * R.nv + n) * d];
            const double* gm = &R.dN[((size_t)q * R.nv + m) * d];
            double dot = 0;
            for (int k = 0; k < d; ++k) dot += gn[k] * gm[k];
            double v = (c == e ? dot : 0.0) + gn[e] * gm[c];
            G[((size_t)q * nvl + n * d + c) * nvl + m * d + e] = v * R.JxW[q];
          }
  const std::vector<double>& Bl = R.Bl;

  std::vector<double> rhs(ntot, 0.0), own_diag(ntot, 0.0);
  auto find = [](const HostCsr& M, int64_t r, int64_t c) -> int64_t {
    auto b = M.ci.begin() + M.rp[r], e = M.ci.begin() + M.rp[r + 1];
    auto it = std::lower_bound(b, e, (int32_t)c);
    if (it == e || *it != (int32_t)c) return -1;
    return it - M.ci.begin();
  };
  double gvec[3] = {0, 0, 0};
  gvec[d - 1] = -prob.gravity;

  // colour cells by parity so same-colour cells share no nodes (also across the
  // periodic seam, since Nc is even there)
  const int ncolor = 1 << d;
  int64_t ncell = 1;
  for (int k = 0; k < d; ++k) ncell *= Nc;
  for (int color = 0; color < ncolor; ++color) {
#pragma omp parallel
    {
      std::vector<double> Aloc((size_t)nvl * nvl), Mloc((size_t)R.np * R.np), floc(nvl);
      std::vector<double> eta_q(R.nq), rho_q(R.nq);
      int64_t vg[81], pg[8], vdofs[81], pdofs_[8];
#pragma omp for schedule(dynamic, 16)
      for (int64_t cell = 0; cell < ncell; ++cell) {
        int64_t cc[3] = {0, 0, 0}, t = cell;
        for (int k = 0; k < d; ++k) { cc[k] = t % Nc; t /= Nc; }
        int col = 0;
        for (int k = 0; k < d; ++k) col |= (int)(cc[k] & 1) << k;
        if (col != color) continue;
        // quadrature point values
        for (int q = 0; q < R.nq; ++q) {
          double x[3] = {0, 0, 0};
          for (int k = 0; k < d; ++k) x[k] = (cc[k] + R.xq[(size_t)q * d + k]) * h;
          F.eval(x, eta_q[q], rho_q[q]);
        }
        std::fill(Aloc.begin(), Aloc.end(), 0.0);
        std::fill(Mloc.begin(), Mloc.end(), 0.0);
        std::fill(floc.begin(), floc.end(), 0.0);
        for (int q = 0; q < R.nq; ++q) {
          const double* Gq = &G[(size_t)q * nvl * nvl];
          const double eq_ = eta_q[q];
          for (size_t i = 0; i < (size_t)nvl * nvl; ++i) Aloc[i] += eq_ * Gq[i];
          const double ie = 1.0 / eta_q[q];
          for (int i = 0; i < R.np; ++i)
            for (int j = 0; j < R.np; ++j)
              Mloc[i * R.np + j] += ie * R.P[(size_t)q * R.np + i] * R.P[(size_t)q * R.np + j] * R.JxW[q];
          if (prob.gravity != 0.0)
            for (int n = 0; n < R.nv; ++n)
              for (int c = 0; c < d; ++c)
                floc[n * d + c] += rho_q[q] * gvec[c] * R.N[(size_t)q * R.nv + n] * R.JxW[q];
        }
        // local -> global dofs
        for (int n = 0; n < R.nv; ++n) {
          int a[3] = {n % 3, (n / 3) % 3, (n / 9) % 3};
          int64_t qv[3] = {0, 0, 0};
          for (int k = 0; k < d; ++k) qv[k] = 2 * cc[k] + a[k];
          int64_t node = LV.index(qv);
          for (int c = 0; c < d; ++c) {
            vdofs[n * d + c] = vdof(node, c);
            vg[n * d + c] = eq[vdofs[n * d + c]];
          }
        }
        for (int m = 0; m < R.np; ++m) {
          int b[3] = {m % 2, (m / 2) % 2, (m / 4) % 2};
          int64_t qp[3] = {0, 0, 0};
          for (int k = 0; k < d; ++k) qp[k] = cc[k] + b[k];
          pdofs_[m] = pdof(LP.index(qp));
          pg[m] = eq[pdofs_[m]];
        }
        // scatter A, rhs
        for (int i = 0; i < nvl; ++i) {
          own_diag[vdofs[i]] += Aloc[(size_t)i * nvl + i];
          const int64_t r = vg[i];
          if (r < 0) continue;
          rhs[r] += floc[i];
          for (int j = 0; j < nvl; ++j) {
            const double val = Aloc[(size_t)i * nvl + j];
            if (vg[j] >= 0) {
              int64_t pos = find(S.A, r, vg[j]);
              S.A.v[pos] += val;
            } else {
              rhs[r] -= val * S.cval[vdofs[j]];
            }
          }
        }
        // scatter B (pressure rows) and its rhs contribution
        for (int i = 0; i < R.np; ++i) {
          own_diag[pdofs_[i]] += Mloc[i * R.np + i];
          const int64_t r = pg[i];
          if (r < 0) continue;
          for (int j = 0; j < nvl; ++j) {
            const double val = Bl[(size_t)i * nvl + j];
            if (vg[j] >= 0) {
              int64_t pos = find(S.B, r - n_u, vg[j]);
              S.B.v[pos] += val;
            } else {
              rhs[r] -= val * S.cval[vdofs[j]];
            }
          }
          for (int j = 0; j < R.np; ++j) {
            if (pg[j] < 0) continue;  // pinned pressure has value 0
            int64_t pos = find(S.Mp, r - n_u, pg[j] - n_u);
            S.Mp.v[pos] += Mloc[i * R.np + j];
          }
        }
        // velocity rows also get B^T p for Dirichlet-eliminated pressure: the
        // only eliminated pressure dof is pinned to 0, so nothing to add.
      }
    }
  }
  // constrained rows
  // mean |A_ii| over free velocity rows, summed by row chunks (the device
  // assembler sums in the same order)
  double diag_sum = 0;
  int64_t diag_cnt = 0;
  for (int64_t r0 = 0; r0 < n_u; r0 += fs::DIAG_CHUNK) {
    double part = 0;
    for (int64_t r = r0; r < std::min(n_u, r0 + fs::DIAG_CHUNK); ++r)
      if (eq[r] == r) { part += std::fabs(S.A.v[find(S.A, r, r)]); ++diag_cnt; }
    diag_sum += part;
  }
  const double Dp = diag_cnt ? diag_sum / (double)diag_cnt : 1.0;
  for (int64_t r = 0; r < n_u; ++r) {
    if (eq[r] == r) continue;
    double D = own_diag[r];
    if (D == 0.0) D = Dp;
    S.A.v[S.A.rp[r]] = D;  // only entry
    rhs[r] = (S.kind[r] == FS_DOF_DIRICHLET) ? D * S.cval[r] : 0.0;
  }
  for (int64_t q = 0; q < n_p; ++q) {
    int64_t g = n_u + q;
    if (eq[g] == g) continue;
    S.K11.v[q] = Dp;
    double Dm = own_diag[g];
    S.Mp.v[S.Mp.rp[q]] = Dm != 0.0 ? Dm : 1.0;
    rhs[g] = 0.0;
  }
  S.Bt = fs::transpose(S.B);
  S.rhs.swap(rhs);

  // log10 viscosity on the pressure-node lattice
  S.log_eta.resize(npn);
  for (int64_t node = 0; node < npn; ++node) {
    int64_t c[3];
    LP.coords(node, c);
    double x[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) x[k] = (double)c[k] * h;
    double eta, rho;
    F.eval(x, eta, rho);
    S.log_eta[node] = (float)std::log10(eta);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int fs_system_create(const fs_problem* prob, fs_system** out) {
  if (!prob || !out) { fs::set_error("fs_system_create: null argument"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  if (prob->dim != 2 && prob->dim != 3) { fs::set_error("fs_system_create: dim must be 2 or 3"); return FS_ERR_BAD_INPUT; }
  if (prob->n_cells < 1 || prob->n_cells > 4096) { fs::set_error("fs_system_create: bad n_cells %d", prob->n_cells); return FS_ERR_BAD_INPUT; }
  int64_t nu = 1;
  for (int k = 0; k < prob->dim; ++k) nu *= (2 * (int64_t)prob->n_cells + 1);
  if (nu * prob->dim >= (int64_t)INT32_MAX) { fs::set_error("fs_system_create: problem too large for 32-bit columns"); return FS_ERR_BAD_INPUT; }
  try {
    std::unique_ptr<fs_system> S(new fs_system());
    build(*prob, *S);
    *out = S.release();
    return FS_OK;
  } catch (const std::bad_alloc&) {
    fs::set_error("fs_system_create: out of host memory");
    return FS_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    fs::set_error("fs_system_create: %s", e.what());
    return FS_ERR_BAD_INPUT;
  }
}

void fs_system_destroy(fs_system* sys) { delete sys; }

int fs_system_get_info(const fs_system* S, fs_system_info* info) {
  if (!S || !info) { fs::set_error("fs_system_get_info: null argument"); return FS_ERR_BAD_INPUT; }
  info->dim = S->dim;
  info->n_cells = S->N;
  info->n_u = S->n_u;
  info->n_p = S->n_p;
  info->nnz_A = S->A.nnz();
  info->nnz_Bt = S->Bt.nnz();
  info->nnz_B = S->B.nnz();
  info->nnz_K11 = S->K11.nnz();
  info->nnz_Mp = S->Mp.nnz();
  info->n_vnodes_1d = 2 * (int64_t)S->N + 1;
  info->n_pnodes_1d = (int64_t)S->N + 1;
  return FS_OK;
}

int fs_system_csr(const fs_system* S, int which, const int64_t** rp, const int32_t** ci, const double** v) {
  if (!S || !rp || !ci || !v) { fs::set_error("fs_system_csr: null argument"); return FS_ERR_BAD_INPUT; }
  const fs::HostCsr* M = nullptr;
  switch (which) {
    case 0: M = &S->A; break;
    case 1: M = &S->Bt; break;
    case 2: M = &S->B; break;
    case 3: M = &S->K11; break;
    case 4: M = &S->Mp; break;
    default: fs::set_error("fs_system_csr: bad block %d", which); return FS_ERR_BAD_INPUT;
  }
  *rp = M->rp.data();
  *ci = M->ci.data();
  *v = M->v.data();
  return FS_OK;
}

const double* fs_system_rhs(const fs_system* S) { return S ? S->rhs.data() : nullptr; }

int fs_system_constraints(const fs_system* S, const int8_t** kind, const int64_t** master, const double** value) {
  if (!S) { fs::set_error("fs_system_constraints: null system"); return FS_ERR_BAD_INPUT; }
  if (kind) *kind = S->kind.data();
  if (master) *master = S->master.data();
  if (value) *value = S->cval.data();
  return FS_OK;
}

int fs_system_distribute(const fs_system* S, double* x) {
  if (!S || !x) { fs::set_error("fs_system_distribute: null argument"); return FS_ERR_BAD_INPUT; }
  const int64_t n = S->n_u + S->n_p;
  for (int64_t i = 0; i < n; ++i) {
    switch (S->kind[i]) {
      case FS_DOF_DIRICHLET: case FS_DOF_PINNED: x[i] = S->cval[i]; break;
      case FS_DOF_PERIODIC: x[i] = x[S->master[i]]; break;
      default: break;
    }
  }
  return FS_OK;
}

int fs_system_dof_maps(const fs_system* S, const int64_t** vnode, const int8_t** vcomp, const int64_t** pnode) {
  if (!S) { fs::set_error("fs_system_dof_maps: null system"); return FS_ERR_BAD_INPUT; }
  if (vnode) *vnode = S->vnode.data();
  if (vcomp) *vcomp = S->vcomp.data();
  if (pnode) *pnode = S->pnode.data();
  return FS_OK;
}

const float* fs_system_log_eta_grid(const fs_system* S) { return S ? S->log_eta.data() : nullptr; }

int fs_system_quadrature_fields(const fs_system* S, double* eta_q, double* rho_q) {
  if (!S || !eta_q || !rho_q) { fs::set_error("fs_system_quadrature_fields: null argument"); return FS_ERR_BAD_INPUT; }
  const int d = S->dim;
  const int64_t Nc = S->N;
  const double h = 1.0 / (double)Nc;
  const fs::RefTables R = fs::ref_tables(d, h);
  Field F;
  F.p = S->prob;
  F.init();
  int64_t ncell = 1;
  for (int k = 0; k < d; ++k) ncell *= Nc;
#pragma omp parallel for schedule(static)
  for (int64_t cell = 0; cell < ncell; ++cell) {
    int64_t cc[3] = {0, 0, 0}, t = cell;
    for (int k = 0; k < d; ++k) { cc[k] = t % Nc; t /= Nc; }
    for (int q = 0; q < R.nq; ++q) {
      double x[3] = {0, 0, 0};
      for (int k = 0; k < d; ++k) x[k] = (cc[k] + R.xq[(size_t)q * d + k]) * h;
      F.eval(x, eta_q[cell * R.nq + q], rho_q[cell * R.nq + q]);
    }
  }
  return FS_OK;
}

int fs_system_block_hint(const fs_system* S, int32_t nodes_per_block, int32_t* out) {
  if (!S || !out || nodes_per_block < 1) { fs::set_error("fs_system_block_hint: bad argument"); return FS_ERR_BAD_INPUT; }
  const int d = S->dim;
  const int64_t n1 = 2 * (int64_t)S->N + 1;
  const int64_t nb1 = (n1 + nodes_per_block - 1) / nodes_per_block;
  for (int64_t i = 0; i < S->n_u; ++i) {
    int64_t node = S->vnode[i], b = 0, mul = 1;
    for (int k = 0; k < d; ++k) {
      int64_t c = node % n1;
      node /= n1;
      b += (c / nodes_per_block) * mul;
      mul *= nb1;
    }
    out[i] = (int32_t)b;
  }
  return FS_OK;
}

}  // extern "C"
