// CPU restatement of FELSPA's Stokes saddle-point solve -- TEST INFRASTRUCTURE.
//
// This file is the checker for the GPU library. It is used only by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg. Each function cites
// the reference file:line it restates. deal.II (SolverGMRES, SolverCG,
// ReductionControl, SparseMatrix) is not vendored by the reference; its
// documented semantics are restated and marked as such.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <functional>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// ---------------------------------------------------------------------------
// deal.II SparseMatrix layout: square matrices keep the diagonal first.
struct DMat {
  int64_t m = 0, n = 0;
  bool square = false;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
};

DMat from_csr(const or_csr* M) {
  DMat D;
  D.m = M->n_rows;
  D.n = M->n_cols;
  D.square = (D.m == D.n);
  D.rp.assign(D.m + 1, 0);
  for (int64_t r = 0; r < D.m; ++r) {
    int64_t len = M->rp[r + 1] - M->rp[r];
    if (D.square) {
      bool has_diag = false;
      for (int64_t k = M->rp[r]; k < M->rp[r + 1]; ++k) has_diag |= (M->ci[k] == r);
      if (!has_diag) ++len;  // square deal.II patterns always store the diagonal
    }
    D.rp[r + 1] = D.rp[r] + len;
  }
  D.ci.resize(D.rp[D.m]);
  D.v.assign(D.rp[D.m], 0.0);
  for (int64_t r = 0; r < D.m; ++r) {
    int64_t p = D.rp[r];
    if (D.square) {
      D.ci[p] = (int32_t)r;
      for (int64_t k = M->rp[r]; k < M->rp[r + 1]; ++k)
        if (M->ci[k] == r) D.v[p] = M->v ? M->v[k] : 0.0;
      ++p;
    }
    for (int64_t k = M->rp[r]; k < M->rp[r + 1]; ++k) {
      if (D.square && M->ci[k] == r) continue;
      D.ci[p] = M->ci[k];
      D.v[p] = M->v ? M->v[k] : 0.0;
      ++p;
    }
  }
  return D;
}

// SparseMatrix::vmult: row-parallel, each row summed in storage order.
void vmult(const DMat& M, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < M.m; ++r) {
    double s = 0;
    for (int64_t k = M.rp[r]; k < M.rp[r + 1]; ++k) s += M.v[k] * x[M.ci[k]];
    y[r] = s;
  }
}

// dst = g - M u  (SparseMatrix::residual)
void residual(const DMat& M, double* dst, const double* u, const double* g) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < M.m; ++r) {
    double s = 0;
    for (int64_t k = M.rp[r]; k < M.rp[r + 1]; ++k) s += M.v[k] * u[M.ci[k]];
    dst[r] = g[r] - s;
  }
}

// deterministic parallel dot: fixed static partition, partials summed in order
double dot(const double* a, const double* b, int64_t n) {
#ifdef _OPENMP
  const int nt = omp_get_max_threads();
#else
  const int nt = 1;
#endif
  std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    double s = 0;
    for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[t] = s;
  }
  double s = 0;
  for (double p : part) s += p;
  return s;
}
double norm(const double* a, int64_t n) { return std::sqrt(dot(a, a, n)); }

// ---------------------------------------------------------------------------
// dealii::ReductionControl::check (documented semantics): success when the
// value falls below reduce * initial value or below tol; failure at max steps
// or NaN.
struct Control {
  int max_steps = 100;
  double tol = 1e-10, reduce = 1e-6, reduced_tol = 0;
  int last_step = 0;
  double last_value = 0;
  enum State { ITERATE, SUCCESS, FAILURE };
  State check(int step, double value) {
    if (step == 0) reduced_tol = value * reduce;
    last_step = step;
    last_value = value;
    if (value <= reduced_tol) return SUCCESS;
    if (value <= tol) return SUCCESS;
    if (step >= max_steps || std::isnan(value)) return FAILURE;
    return ITERATE;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// SparseILU<double> (include/felspa/linear_algebra/ilu.h:25-238)
struct or_ilu {
  int k = 0;
  bool force_generic = false;
  int64_t N = 0;
  std::vector<int64_t> rp;     // deal.II layout: diagonal first
  std::vector<int32_t> ci;
  std::vector<double> v;
  std::vector<int64_t> one_after_diagonal;  // absolute index of first column > row
};

namespace {

using Rows = std::vector<std::vector<int32_t>>;  // sorted columns, diagonal in place

// sorted columns of row r of a deal.II-layout square matrix (diagonal placed in order)
std::vector<int32_t> sorted_row(const DMat& A, int64_t r) {
  std::vector<int32_t> c(A.ci.begin() + A.rp[r], A.ci.begin() + A.rp[r + 1]);
  std::sort(c.begin(), c.end());
  return c;
}

// symbolic_factorize_levelwise + row_factorization (ilu.implement.h:90-242):
// row i = A(i,:) U {i} U  union over k in A(i,:), k < i, of {j in A(k,:) : j > k}.
Rows symbolic_levelwise(const DMat& A) {
  const int64_t N = A.m;
  Rows rows(N);
#pragma omp parallel for schedule(dynamic, 1000)
  for (int64_t i = 0; i < N; ++i) {
    std::vector<int32_t> c = sorted_row(A, i);
    std::vector<int32_t> acc = c;
    for (int32_t kk : c) {
      if (kk >= i) break;
      for (int64_t p = A.rp[kk] + 1; p < A.rp[kk + 1]; ++p)
        if (A.ci[p] > kk) acc.push_back(A.ci[p]);
    }
    std::sort(acc.begin(), acc.end());
    acc.erase(std::unique(acc.begin(), acc.end()), acc.end());
    rows[i].swap(acc);
  }
  return rows;
}

// symbolic_factorize_generic (ilu.implement.h:245-435). Level of a new fill
// (i,j) through pivot row k is min(lev(k,j) + lev(i,k) + 1, lev(i,j)); the fill
// is kept when <= max level. An entry already in row i keeps the level it was
// first given (the reference only writes the level when inserting, :367-392).
Rows symbolic_generic(const DMat& A, unsigned kmax) {
  const int64_t N = A.m;
  Rows rows(N);
  if (kmax == 0) {  // :259 -- ILU(0) uses the matrix pattern
    for (int64_t i = 0; i < N; ++i) rows[i] = sorted_row(A, i);
    return rows;
  }
  const unsigned invalid = UINT_MAX;
  std::vector<std::vector<unsigned>> lofs(N);
  std::vector<size_t> rdiag(N);
  std::vector<unsigned> lof(N, invalid);
  for (int64_t i = 0; i < N; ++i) {
    std::set<int32_t> list;
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) list.insert(A.ci[p]);
    list.insert((int32_t)i);
    for (int32_t c : list) lof[c] = 0;
    for (auto it = list.begin(); *it < i; ++it) {
      const int32_t kk = *it;
      if (lof[kk] < kmax) {
        const auto& rk = rows[kk];
        for (size_t pos = rdiag[kk] + 1; pos < rk.size(); ++pos) {
          const int32_t j = rk[pos];
          unsigned lev = std::min(lofs[kk][pos] + lof[kk] + 1, lof[j]);
          if (lev <= kmax && lof[j] == invalid) {
            list.insert(j);
            lof[j] = lev;
          }
        }
      }
    }
    rows[i].assign(list.begin(), list.end());
    lofs[i].resize(rows[i].size());
    for (size_t p = 0; p < rows[i].size(); ++p) {
      lofs[i][p] = lof[rows[i][p]];
      lof[rows[i][p]] = invalid;
      if (rows[i][p] == i) rdiag[i] = p;
    }
  }
  return rows;
}

void ilu_set_pattern(or_ilu& F, const Rows& rows) {
  const int64_t N = (int64_t)rows.size();
  F.N = N;
  F.rp.assign(N + 1, 0);
  for (int64_t i = 0; i < N; ++i) F.rp[i + 1] = F.rp[i] + (int64_t)rows[i].size();
  F.ci.resize(F.rp[N]);
  F.v.assign(F.rp[N], 0.0);
  F.one_after_diagonal.resize(N);
  for (int64_t i = 0; i < N; ++i) {
    int64_t p = F.rp[i];
    F.ci[p++] = (int32_t)i;
    for (int32_t c : rows[i])
      if (c != i) F.ci[p++] = c;
    // cache_diagonal_iterators (ilu.implement.h:750-776)
    int64_t q = F.rp[i] + 1;
    while (q < F.rp[i + 1] && F.ci[q] < i) ++q;
    F.one_after_diagonal[i] = q;
  }
}

// numeric_factorize (ilu.implement.h:438-541): row-wise IKJ on the fixed pattern.
void ilu_numeric(or_ilu& F, const DMat& A) {
  const int64_t N = F.N;
  std::vector<char> filter(N, 0);
  std::vector<double> w(N, 0.0);
  for (int64_t i = 0; i < N; ++i) {
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) w[A.ci[p]] = A.v[p];
    for (int64_t p = F.rp[i]; p < F.rp[i + 1]; ++p) filter[F.ci[p]] = 1;
    for (int64_t p = F.rp[i] + 1; p < F.one_after_diagonal[i]; ++p) {
      const int32_t kc = F.ci[p];
      const double mult = w[kc] / F.v[F.rp[kc]];
      w[kc] = mult;
      for (int64_t q = F.one_after_diagonal[kc]; q < F.rp[kc + 1]; ++q) {
        const int32_t j = F.ci[q];
        w[j] -= mult * F.v[q] * (double)filter[j];
      }
    }
    for (int64_t p = F.rp[i]; p < F.rp[i + 1]; ++p) {
      F.v[p] = w[F.ci[p]];
      filter[F.ci[p]] = 0;
      w[F.ci[p]] = 0.0;
    }
  }
}

// SparseILU::vmult, serial (USE_SERIAL_VMULT, ilu.implement.h:544-595)
void ilu_vmult(const or_ilu& F, const double* src, double* dst) {
  const int64_t N = F.N;
  if (dst != src) std::memcpy(dst, src, sizeof(double) * N);
  for (int64_t i = 0; i < N; ++i) {
    double s = dst[i];
    for (int64_t p = F.rp[i] + 1; p < F.one_after_diagonal[i]; ++p) s -= F.v[p] * dst[F.ci[p]];
    dst[i] = s;
  }
  for (int64_t i1 = N; i1 > 0; --i1) {
    const int64_t i = i1 - 1;
    double s = dst[i];
    for (int64_t p = F.one_after_diagonal[i]; p < F.rp[i + 1]; ++p) s -= F.v[p] * dst[F.ci[p]];
    dst[i] = s / F.v[F.rp[i]];
  }
}

// SparseILU::Tvmult (ilu.implement.h:626-680)
void ilu_Tvmult(const or_ilu& F, const double* src, double* dst) {
  const int64_t N = F.N;
  if (dst != src) std::memcpy(dst, src, sizeof(double) * N);
  std::vector<double> t(N, 0.0);
  for (int64_t i = 0; i < N; ++i) {
    dst[i] = (dst[i] - t[i]) / F.v[F.rp[i]];
    const double d = dst[i];
    for (int64_t p = F.one_after_diagonal[i]; p < F.rp[i + 1]; ++p) t[F.ci[p]] += F.v[p] * d;
  }
  std::fill(t.begin(), t.end(), 0.0);
  for (int64_t i1 = N; i1 > 0; --i1) {
    const int64_t i = i1 - 1;
    dst[i] -= t[i];
    const double d = dst[i];
    for (int64_t p = F.one_after_diagonal[i] - 1; p > F.rp[i]; --p) t[F.ci[p]] += d * F.v[p];
  }
}

// SparseILU::initialize (ilu.implement.h:29-55)
void ilu_initialize(or_ilu& F, const DMat& A, int k, bool force_generic, bool use_previous) {
  if (!use_previous) {
    Rows rows = (k == 1 && !force_generic) ? symbolic_levelwise(A) : symbolic_generic(A, (unsigned)k);
    ilu_set_pattern(F, rows);
  }
  ilu_numeric(F, A);
}

// dealii::SolverCG (documented algorithm), preconditioned by an ILU.
// Returns last step; throws nothing, reports non-convergence via ok=false.
int solve_cg(const std::function<void(const double*, double*)>& Aop,
             const std::function<void(const double*, double*)>& Pop, int64_t n, double* x,
             const double* b, Control& ctrl, bool& ok) {
  std::vector<double> g(n), h(n), d(n);
  bool xzero = true;
  for (int64_t i = 0; i < n && xzero; ++i) xzero = (x[i] == 0.0);
  if (!xzero) {
    Aop(x, g.data());
    for (int64_t i = 0; i < n; ++i) g[i] -= b[i];
  } else {
    for (int64_t i = 0; i < n; ++i) g[i] = -b[i];
  }
  double res = norm(g.data(), n);
  Control::State st = ctrl.check(0, res);
  int it = 0;
  if (st == Control::ITERATE) {
    Pop(g.data(), h.data());
    for (int64_t i = 0; i < n; ++i) d[i] = -h[i];
    double gh = dot(g.data(), h.data(), n);
    while (st == Control::ITERATE) {
      ++it;
      Aop(d.data(), h.data());
      const double alpha = gh / dot(d.data(), h.data(), n);
      // deal.II applies a non-finite alpha and then throws NoConvergence on the
      // NaN residual; the oracle keeps the last finite iterate and reports
      // non-convergence instead (round-off breakdown of tolerances like 1e-15).
      if (!std::isfinite(alpha)) { st = Control::FAILURE; break; }
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * d[i];
        g[i] += alpha * h[i];
      }
      res = norm(g.data(), n);
      st = ctrl.check(it, res);
      if (st != Control::ITERATE) break;
      Pop(g.data(), h.data());
      double beta = gh;
      gh = dot(g.data(), h.data(), n);
      beta = gh / beta;
      for (int64_t i = 0; i < n; ++i) d[i] = beta * d[i] - h[i];
    }
  }
  ok = (st == Control::SUCCESS);
  return it;
}

// InverseMatrixCG::vmult (linear_operators.implement.h:64-95):
// SolverCG with ReductionControl(tol, reduce), max_steps = n, from dst.
struct InverseCG {
  const DMat* M;
  const or_ilu* P;
  double tol, reduce;
  mutable long long iters = 0;
  mutable bool ok = true;
  void vmult(double* dst, const double* src) const {
    Control c;
    c.tol = tol;
    c.reduce = reduce;
    c.max_steps = (int)M->m;
    bool conv = true;
    iters += solve_cg([&](const double* a, double* b) { vmult_op(a, b); },
                      [&](const double* a, double* b) { ilu_vmult(*P, a, b); }, M->m, dst, src, c, conv);
    ok = ok && conv;
  }
  void vmult_op(const double* a, double* b) const { ::vmult(*M, a, b); }
};

// Stokes block operator K = [[A, Bt], [B, K11]]
struct Stokes {
  DMat A, Bt, B, K11, Mp;
  int64_t nu = 0, np = 0;
  void vmult(const double* x, double* y) const {
    std::vector<double> t(std::max(nu, np));
    ::vmult(A, x, y);
    ::vmult(Bt, x + nu, t.data());
    for (int64_t i = 0; i < nu; ++i) y[i] += t[i];
    ::vmult(B, x, y + nu);
    ::vmult(K11, x + nu, t.data());
    for (int64_t i = 0; i < np; ++i) y[nu + i] += t[i];
  }
};

// BlockSchurPreconditionerBase::vmult (linear_operators.implement.h:200-209)
struct BlockSchur {
  const Stokes* S;
  const or_ilu* ILU_A;
  InverseCG inverse_S;
  mutable std::vector<double> tmp;
  void vmult(double* dst, const double* src) const {
    const int64_t nu = S->nu, np = S->np;
    ilu_vmult(*ILU_A, src, dst);
    tmp.resize(np);
    residual(S->B, tmp.data(), dst, src + nu);
    for (int64_t i = 0; i < np; ++i) tmp[i] *= -1.0;
    // The inner CG starts from dst's block 1. deal.II's GMRES passes a pooled
    // vector there whose old contents are an internal detail; start from zero.
    std::fill(dst + nu, dst + nu + np, 0.0);
    inverse_S.vmult(dst + nu, tmp.data());
  }
};

// Givens step of dealii::SolverGMRES (documented algorithm)
void givens(std::vector<double>& h, std::vector<double>& g, std::vector<double>& c,
            std::vector<double>& s, int col) {
  for (int i = 0; i < col; ++i) {
    const double hi = h[i];
    h[i] = c[i] * hi + s[i] * h[i + 1];
    h[i + 1] = -s[i] * hi + c[i] * h[i + 1];
  }
  const double r = 1.0 / std::sqrt(h[col] * h[col] + h[col + 1] * h[col + 1]);
  s[col] = h[col + 1] * r;
  c[col] = h[col] * r;
  h[col] = c[col] * h[col] + s[col] * h[col + 1];
  g[col + 1] = -s[col] * g[col];
  g[col] *= c[col];
}

// dealii::SolverGMRES with left preconditioning, modified Gram-Schmidt,
// restart = max_n_tmp_vectors - 2 (documented algorithm), as called from
// StokesLinearSystem::solve_gmres (stokes_ilu.cc:533-542).
template <class Op, class Prec>
void gmres_left(const Op& K, const Prec& P, int64_t n, double* x, const double* b, int restart,
                Control& ctrl) {
  std::vector<std::vector<double>> V(restart + 1, std::vector<double>(n));
  std::vector<double> p(n);
  std::vector<double> H((size_t)(restart + 1) * restart, 0.0), h(restart + 1), gam(restart + 1),
    cs(restart), sn(restart);
  int acc = 0;
  Control::State st = Control::ITERATE;
  do {
    K.vmult(x, p.data());
    for (int64_t i = 0; i < n; ++i) p[i] = b[i] - p[i];
    P.vmult(V[0].data(), p.data());
    double rho = norm(V[0].data(), n);
    st = ctrl.check(acc, rho);
    if (st != Control::ITERATE) break;
    gam[0] = rho;
    for (int64_t i = 0; i < n; ++i) V[0][i] /= rho;
    int dim = 0;
    for (int j = 0; j < restart && st == Control::ITERATE; ++j) {
      ++acc;
      K.vmult(V[j].data(), p.data());
      P.vmult(V[j + 1].data(), p.data());
      dim = j + 1;
      double* vv = V[j + 1].data();
      for (int i = 0; i < dim; ++i) {
        h[i] = dot(vv, V[i].data(), n);
        const double hi = h[i];
        const double* vi = V[i].data();
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) vv[t] -= hi * vi[t];
      }
      const double s = norm(vv, n);
      h[j + 1] = s;
      if (s != 0)
        for (int64_t t = 0; t < n; ++t) vv[t] /= s;
      givens(h, gam, cs, sn, j);
      for (int i = 0; i < dim; ++i) H[(size_t)i * restart + j] = h[i];
      rho = std::fabs(gam[dim]);
      st = ctrl.check(acc, rho);
    }
    // back substitution H y = gamma, x += V y
    std::vector<double> y(dim);
    for (int i = dim - 1; i >= 0; --i) {
      double s = gam[i];
      for (int k = i + 1; k < dim; ++k) s -= H[(size_t)i * restart + k] * y[k];
      y[i] = s / H[(size_t)i * restart + i];
    }
    for (int i = 0; i < dim; ++i)
      for (int64_t t = 0; t < n; ++t) x[t] += y[i] * V[i][t];
  } while (st == Control::ITERATE);
}

// Inner PCG on A (fixed number of steps from zero), ILU_A preconditioned:
// the A^{-1} of the block preconditioner in FS_METHOD_FGMRES.
// The flexible method's A^{-1}: a fixed number of ILU(k)-preconditioned CG
// steps from x = 0 (no reference counterpart; the reference applies ILU_A once,
// linear_operators.implement.h:200-209), in the Chronopoulos-Gear form that
// the GPU runs: z = M^-1 r, s = A z, gamma = r.z, delta = z.s;
// beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old);
// p = z + beta p, q = s + beta q, x += alpha p, r -= alpha q. These are the
// textbook PCG iterates (q = A p by recurrence, p.Ap = delta - beta gamma /
// alpha_old), which tests/test_cpu_oracle.py checks against a textbook PCG.
struct InnerPcgA {
  const DMat* A;
  const or_ilu* P;
  int iters;
  mutable long long applies = 0;
  void vmult(double* x, const double* b) const {
    const int64_t n = A->m;
    std::vector<double> r(b, b + n), z(n), s(n), p(n), q(n);
    std::fill(x, x + n, 0.0);
    double alpha = 0, beta = 0, gamma_old = 0;
    for (int it = 0; it < iters; ++it) {
      if (it > 0) {
        const bool first = it == 1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
          p[i] = first ? z[i] : z[i] + beta * p[i];
          q[i] = first ? s[i] : s[i] + beta * q[i];
          x[i] = first ? alpha * p[i] : x[i] + alpha * p[i];
          r[i] -= alpha * q[i];
        }
      }
      ilu_vmult(*P, r.data(), z.data());
      ++applies;
      ::vmult(*A, z.data(), s.data());
      const double gamma = dot(r.data(), z.data(), n), delta = dot(z.data(), s.data(), n);
      double den = delta;
      beta = 0;
      if (it > 0 && gamma_old != 0.0 && alpha != 0.0) {
        beta = gamma / gamma_old;
        den = delta - beta * gamma / alpha;
      }
      alpha = den != 0.0 ? gamma / den : 0.0;
      gamma_old = gamma;
    }
    const bool first = iters == 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double pi = first ? z[i] : z[i] + beta * p[i];
      x[i] = first ? alpha * pi : x[i] + alpha * pi;
    }
  }
};

// Block preconditioner with the inner PCG on A (block lower triangular form
// of BlockSchurPreconditionerBase::vmult, linear_operators.implement.h:200-209)
struct BlockSchurPcg {
  const Stokes* S;
  InnerPcgA Ainv;
  InverseCG inverse_S;
  mutable std::vector<double> tmp;
  void vmult(double* dst, const double* src) const {
    const int64_t nu = S->nu, np = S->np;
    Ainv.vmult(dst, src);
    tmp.resize(np);
    residual(S->B, tmp.data(), dst, src + nu);
    for (int64_t i = 0; i < np; ++i) tmp[i] *= -1.0;
    std::fill(dst + nu, dst + nu + np, 0.0);
    inverse_S.vmult(dst + nu, tmp.data());
  }
};

// Flexible right-preconditioned GMRES(m) restarted from the true residual,
// stopping on ||b - K x|| <= tol (the GPU FGMRES driver, FP64)
template <class Op, class Prec>
int fgmres_true(const Op& K, const Prec& P, int64_t n, double* x, const double* b, int m, double tol,
                int max_steps, double& last) {
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), Z(m, std::vector<double>(n));
  std::vector<double> H((size_t)(m + 1) * m, 0.0), h(m + 1), gam(m + 1), cs(m), sn(m);
  int acc = 0;
  for (;;) {
    K.vmult(x, V[0].data());
    for (int64_t i = 0; i < n; ++i) V[0][i] = b[i] - V[0][i];
    const double beta = norm(V[0].data(), n);
    last = beta;
    if (beta <= tol) return acc;
    if (acc >= max_steps) return acc;
    for (int64_t i = 0; i < n; ++i) V[0][i] /= beta;
    gam.assign(m + 1, 0.0);
    gam[0] = beta;
    int dim = 0;
    for (int j = 0; j < m && acc < max_steps; ++j) {
      ++acc;
      P.vmult(Z[j].data(), V[j].data());
      K.vmult(Z[j].data(), V[j + 1].data());
      double* w = V[j + 1].data();
      std::vector<double> h1(j + 1, 0.0);
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i <= j; ++i) {
          const double c = dot(w, V[i].data(), n);
          h1[i] += c;
          const double* vi = V[i].data();
#pragma omp parallel for schedule(static)
          for (int64_t t = 0; t < n; ++t) w[t] -= c * vi[t];
        }
      for (int i = 0; i <= j; ++i) h[i] = h1[i];
      const double s = norm(w, n);
      h[j + 1] = s;
      if (s != 0)
        for (int64_t t = 0; t < n; ++t) w[t] /= s;
      givens(h, gam, cs, sn, j);
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
      dim = j + 1;
      if (std::fabs(gam[dim]) <= tol) break;
    }
    std::vector<double> y(dim);
    for (int i = dim - 1; i >= 0; --i) {
      double s = gam[i];
      for (int k = i + 1; k < dim; ++k) s -= H[(size_t)i * m + k] * y[k];
      y[i] = s / H[(size_t)i * m + i];
    }
    for (int i = 0; i < dim; ++i)
      for (int64_t t = 0; t < n; ++t) x[t] += y[i] * Z[i][t];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

void or_spmv(const or_csr* M, const double* x, double* y) {
  DMat D = from_csr(M);
  vmult(D, x, y);
}

int or_ilu_create(const or_csr* A, int k, int force_generic, or_ilu** out) {
  if (!A || !out || A->n_rows != A->n_cols || k < 0) return -1;
  auto F = std::make_unique<or_ilu>();
  F->k = k;
  F->force_generic = force_generic != 0;
  DMat D = from_csr(A);
  ilu_initialize(*F, D, k, F->force_generic, false);
  for (int64_t i = 0; i < F->N; ++i)
    if (!(F->v[F->rp[i]] != 0.0) || !std::isfinite(F->v[F->rp[i]])) { *out = F.release(); return -2; }
  *out = F.release();
  return 0;
}

int or_ilu_refactor(or_ilu* F, const or_csr* A) {
  DMat D = from_csr(A);
  ilu_initialize(*F, D, F->k, F->force_generic, true);
  return 0;
}

void or_ilu_vmult(const or_ilu* F, const double* src, double* dst) { ilu_vmult(*F, src, dst); }
void or_ilu_Tvmult(const or_ilu* F, const double* src, double* dst) { ilu_Tvmult(*F, src, dst); }

void or_ilu_get(const or_ilu* F, int64_t* n, int64_t* nnz, int64_t* rp, int32_t* ci, double* v) {
  if (n) *n = F->N;
  if (nnz) *nnz = F->rp[F->N];
  if (!rp || !ci) return;
  // sorted layout, diagonal in place
  rp[0] = 0;
  for (int64_t i = 0; i < F->N; ++i) {
    int64_t o = rp[i];
    const int64_t b = F->rp[i], oad = F->one_after_diagonal[i], e = F->rp[i + 1];
    for (int64_t p = b + 1; p < oad; ++p) { ci[o] = F->ci[p]; if (v) v[o] = F->v[p]; ++o; }
    ci[o] = (int32_t)i; if (v) v[o] = F->v[b]; ++o;
    for (int64_t p = oad; p < e; ++p) { ci[o] = F->ci[p]; if (v) v[o] = F->v[p]; ++o; }
    rp[i + 1] = o;
  }
}

void or_ilu_destroy(or_ilu* F) { delete F; }

void or_scaling_coeffs(const or_csr* Ac, const or_csr* Btc, int64_t n_p, double* s_u, double* s_p) {
  // StokesScalingOperator::initialize (stokes_common.implement.h:13-53)
  const int64_t nu = Ac->n_rows;
  std::vector<double> ghat(nu);
  for (int64_t r = 0; r < nu; ++r) {
    double mx = 0;
    for (int64_t k = Ac->rp[r]; k < Ac->rp[r + 1]; ++k) mx = std::max(std::fabs(Ac->v[k]), mx);
    s_u[r] = std::sqrt(mx);
  }
  for (int64_t r = 0; r < nu; ++r) {
    double mx = 0;
    for (int64_t k = Btc->rp[r]; k < Btc->rp[r + 1]; ++k) mx = std::max(std::fabs(Btc->v[k]), mx);
    ghat[r] = mx;
  }
  double ps = 0;
  for (int64_t r = 0; r < nu; ++r) ps += 1.0 / s_u[r] / s_u[r];
  double g2 = 0;
  for (int64_t r = 0; r < nu; ++r) g2 += ghat[r] * ghat[r];
  ps *= std::sqrt(ps * g2) / (double)n_p;
  for (int64_t i = 0; i < n_p; ++i) s_p[i] = ps;
}

void or_inner_pcg_a(const or_csr* A, int kA, int iters, const double* b, double* x) {
  const DMat M = from_csr(A);
  or_ilu F;
  ilu_initialize(F, M, kA, false, false);
  InnerPcgA{&M, &F, iters}.vmult(x, b);
}

int or_block_schur_vmult(const or_csr* A, const or_csr* B, const or_csr* Mp, int kA, int kS,
                         double inner_tol, const double* src, double* dst, long long* inner_iters) {
  Stokes S;
  S.A = from_csr(A);
  S.B = from_csr(B);
  S.Mp = from_csr(Mp);
  S.nu = S.A.m;
  S.np = S.Mp.m;
  or_ilu IA, IS;
  ilu_initialize(IA, S.A, kA, false, false);
  ilu_initialize(IS, S.Mp, kS, false, false);
  BlockSchur P{&S, &IA, InverseCG{&S.Mp, &IS, inner_tol, inner_tol}, {}};
  P.vmult(dst, src);
  if (inner_iters) *inner_iters = P.inverse_S.iters;
  return P.inverse_S.ok ? 0 : 1;
}

int or_stokes_solve(const or_csr* Ac, const or_csr* Btc, const or_csr* Bc, const or_csr* K11c,
                    const or_csr* Mpc, const or_opts* o, const double* rhs_in, double* x,
                    or_stats* stats) {
#ifdef _OPENMP
  if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
  auto t0 = Clock::now();
  Stokes S;
  S.A = from_csr(Ac);
  S.Bt = from_csr(Btc);
  S.B = from_csr(Bc);
  S.K11 = from_csr(K11c);
  S.Mp = from_csr(Mpc);
  S.nu = S.A.m;
  S.np = S.Mp.m;
  const int64_t nu = S.nu, np = S.np, n = nu + np;
  std::vector<double> rhs(rhs_in, rhs_in + n);
  std::vector<double> su(nu), sp(np);
  if (o->apply_diagonal_scaling) {
    // pre_solve_scaling (stokes_ilu.cc:418-431) with StokesScalingOperator
    // (stokes_common.implement.h:13-140): D^{-1} K D^{-1} (D x) = D^{-1} b
    or_scaling_coeffs(Ac, Btc, np, su.data(), sp.data());
    for (int64_t r = 0; r < nu; ++r)
      for (int64_t k = S.A.rp[r]; k < S.A.rp[r + 1]; ++k) S.A.v[k] /= su[r] * su[S.A.ci[k]];
    for (int64_t r = 0; r < nu; ++r)
      for (int64_t k = S.Bt.rp[r]; k < S.Bt.rp[r + 1]; ++k) S.Bt.v[k] /= su[r] * sp[S.Bt.ci[k]];
    for (int64_t r = 0; r < np; ++r)
      for (int64_t k = S.B.rp[r]; k < S.B.rp[r + 1]; ++k) S.B.v[k] /= sp[r] * su[S.B.ci[k]];
    for (int64_t r = 0; r < np; ++r)
      for (int64_t k = S.Mp.rp[r]; k < S.Mp.rp[r + 1]; ++k) S.Mp.v[k] /= sp[r] * sp[S.Mp.ci[k]];
    for (int64_t i = 0; i < nu; ++i) { x[i] *= su[i]; rhs[i] /= su[i]; }
    for (int64_t i = 0; i < np; ++i) { x[nu + i] *= sp[i]; rhs[nu + i] /= sp[i]; }
  }
  // constrained entries of the guess are zeroed by the caller (stokes_ilu.cc:523-524)
  or_ilu IA, IS;
  ilu_initialize(IA, S.A, o->ilu_k_A, false, false);
  ilu_initialize(IS, S.Mp, o->ilu_k_S, false, false);
  auto t1 = Clock::now();
  Control ctrl;
  const double bnorm = norm(rhs.data(), n);
  ctrl.tol = o->rtol * bnorm;
  ctrl.reduce = o->reduce;
  ctrl.max_steps = o->max_steps > 0 ? o->max_steps : (int)n;
  bool ok = true;
  long long inner = 0;
  if (o->method == 2) {
    // the flexible method's inner CG on M_p stops on a relative reduction only
    BlockSchurPcg P{&S, InnerPcgA{&S.A, &IA, o->inner_a_iters > 0 ? o->inner_a_iters : 30},
                    InverseCG{&S.Mp, &IS, 0.0, o->inner_tol}, {}};
    double last = 0;
    const int it = fgmres_true(S, P, n, x, rhs.data(), o->restart, o->rtol * bnorm, ctrl.max_steps, last);
    ctrl.last_step = it;
    ctrl.last_value = last;
    ok = P.inverse_S.ok && last <= o->rtol * bnorm;
    inner = P.inverse_S.iters + P.Ainv.applies;
  } else if (o->method == 0) {
    BlockSchur P{&S, &IA, InverseCG{&S.Mp, &IS, o->inner_tol, o->inner_tol}, {}};
    gmres_left(S, P, n, x, rhs.data(), o->restart, ctrl);
    ok = P.inverse_S.ok && ctrl.last_value <= std::max(ctrl.tol, ctrl.reduced_tol);
    inner = P.inverse_S.iters;
  } else {
    // solve_cg: Schur-complement reduction (stokes_ilu.cc:564-710)
    InverseCG Ainv{&S.A, &IA, 1e-15, 1e-15};
    std::vector<double> tmp(nu), srhs(np), t2(nu), t3(nu);
    Ainv.vmult(tmp.data(), rhs.data());
    vmult(S.B, tmp.data(), srhs.data());
    for (int64_t i = 0; i < np; ++i) srhs[i] -= rhs[nu + i];
    InverseCG Sinv{&S.Mp, &IS, o->inner_tol, o->inner_tol};
    ctrl.tol = o->rtol * norm(srhs.data(), np);
    ctrl.max_steps = (int)np;
    auto schur = [&](const double* a, double* b) {
      vmult(S.Bt, a, t2.data());
      std::fill(t3.begin(), t3.end(), 0.0);
      Ainv.vmult(t3.data(), t2.data());
      vmult(S.B, t3.data(), b);
    };
    bool conv = true;
    solve_cg(schur, [&](const double* a, double* b) { std::fill(b, b + np, 0.0); Sinv.vmult(b, a); },
             np, x + nu, srhs.data(), ctrl, conv);
    // velocity: A u = f - Bt p
    vmult(S.Bt, x + nu, tmp.data());
    for (int64_t i = 0; i < nu; ++i) tmp[i] = rhs[i] - tmp[i];
    Ainv.vmult(x, tmp.data());
    ok = conv && Ainv.ok && Sinv.ok;
    inner = Ainv.iters + Sinv.iters;
  }
  if (o->apply_diagonal_scaling) {
    for (int64_t i = 0; i < nu; ++i) x[i] /= su[i];
    for (int64_t i = 0; i < np; ++i) x[nu + i] /= sp[i];
  }
  auto t2 = Clock::now();
  if (stats) {
    stats->status = ok ? 0 : 1;
    stats->outer_iters = ctrl.last_step;
    stats->inner_iters = inner;
    stats->last_value = ctrl.last_value;
    stats->t_setup_s = secs(t0, t1);
    stats->t_krylov_s = secs(t1, t2);
  }
  return ok ? 0 : 1;
}

}  // extern "C"
