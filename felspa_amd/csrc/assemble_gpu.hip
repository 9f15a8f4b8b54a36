// Device assembly of the Q2-Q1 Stokes blocks on the patterns of a generated
// system (fs_assemble.h, "Device assembly").
//
// Restates on the GPU what the host generator (assemble.cpp) does per cell:
// StokesAssembler::local_assembly (src/pde/stokes_ilu.cc:739-821) for
//   A    = 2 eta eps(u):eps(v)       (element matrix sum_q eta_q G_q)
//   B    = -q div v                  (viscosity free, one table Bl)
//   M_p  = p q / eta                 (Schur approximation)
//   rhs  = rho g . v
// followed by copy_local_to_global (include/felspa/pde/src/
// stokes_common.implement.h:658-669) with the generator's condensation:
// Dirichlet columns go to the rhs, periodic slaves add into their masters,
// constrained rows keep only their diagonal.
//
// Determinism: one launch per cell colour (cells whose indices have the same
// parity in every direction share no node, also across the periodic seam),
// so every matrix / rhs entry receives at most one contribution per launch
// and the contributions arrive in the host's colour order. Within a cell the
// element entries are formed with the host's operations in the host's order
// and with floating-point contraction off, so the values are bit-identical
// to the host generator's for the same quadrature samples. The mean |A_ii|
// behind the constrained pressure rows is summed in the host's two-level
// chunk order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "assemble_internal.h"
#include "solver_internal.h"

using namespace fs;
using fsk::BLOCK;

struct fs_assembler {
  int dim = 0, N = 0, device = 0;
  int64_t n_u = 0, n_p = 0, ncell = 0, n_free_u = 0;
  int64_t nnzA = 0, nnzBt = 0, nnzB = 0, nnzM = 0;
  int nq = 0, max_row = 0;
  fs_problem prob{};
  double h = 0;
  DevBuf vpos, ppos, eq, kind, cval, dpos, Arp, Aci, Brp, Bci, Mrp, Mci, bt_src, tabs, ylay, ell, own, part, err;
  int64_t off_dN = 0, off_N = 0, off_P = 0, off_JxW = 0, off_xq = 0, off_Bl = 0;
  int n_ell = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0;
  ~fs_assembler() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

constexpr int ASM_NT = 256;
constexpr int ASM_NW = ASM_NT / 64;

struct AsmArgs {
  int N, color;
  int m[3];  // cells of this colour per direction
  int64_t nv1, np1, n_u;
  const int32_t* vpos;
  const int32_t* ppos;
  const int32_t* eq;
  const double* cval;
  const int64_t* Arp;
  const int32_t* Aci;
  double* Av;
  const int64_t* Brp;
  const int32_t* Bci;
  double* Bv;
  const int64_t* Mrp;
  const int32_t* Mci;
  double* Mv;
  double* rhs;
  double* own;
  const double* eta_q;
  const double* rho_q;
  const double* dN;
  const double* Nq;
  const double* P;
  const double* JxW;
  const double* Bl;
  double grav;  // body-force component along the last axis (-gravity)
  int gravity_on;
  int* err;
};

__device__ inline int lds_lower_bound(const int32_t* a, int n, int32_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One workgroup per cell of one colour. Waves take the cell's local rows in
// turn; a wave stages the global row's columns in LDS and its lanes take the
// row's local columns, each forming its element entry and adding it at the
// entry's position (found by binary search in the staged columns).
template <int D, int ROWCAP>
__global__ __launch_bounds__(ASM_NT) void k_asm_cells(AsmArgs a) {
#pragma clang fp contract(off)
  constexpr int NV = D == 2 ? 9 : 27, NP = D == 2 ? 4 : 8, NQ = NV, NVL = NV * D;
  __shared__ double s_dN[NQ * NV * D];
  __shared__ double s_eta[NQ], s_ie[NQ], s_jxw[NQ], s_floc[NVL];
  __shared__ int32_t s_vg[NVL], s_vd[NVL], s_pg[NP], s_pd[NP];
  __shared__ int32_t s_cols[ASM_NW][ROWCAP];
  __shared__ double s_tmp[ASM_NW][NVL];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

  int64_t b = blockIdx.x;
  int cc[3] = {0, 0, 0};
  for (int k = 0; k < D; ++k) {
    const int t = (int)(b % a.m[k]);
    b /= a.m[k];
    cc[k] = 2 * t + ((a.color >> k) & 1);
  }
  const int64_t cell = cc[0] + (int64_t)a.N * (cc[1] + (int64_t)a.N * cc[2]);

  for (int i = tid; i < NQ * NV * D; i += ASM_NT) s_dN[i] = a.dN[i];
  if (tid < NQ) {
    const double e = a.eta_q[cell * NQ + tid];
    s_eta[tid] = e;
    s_ie[tid] = 1.0 / e;
    s_jxw[tid] = a.JxW[tid];
  }
  if (tid < NV) {
    int64_t node = 0, mul = 1;
    for (int k = 0; k < D; ++k) {
      const int off = k == 0 ? tid % 3 : k == 1 ? (tid / 3) % 3 : tid / 9;
      node += (2 * (int64_t)cc[k] + off) * mul;
      mul *= a.nv1;
    }
    const int32_t vp = a.vpos[node];
    for (int c = 0; c < D; ++c) {
      const int32_t g = vp * D + c;
      s_vd[tid * D + c] = g;
      s_vg[tid * D + c] = a.eq[g];
    }
  } else if (tid >= 64 && tid < 64 + NP) {
    const int j = tid - 64;
    int64_t node = 0, mul = 1;
    for (int k = 0; k < D; ++k) {
      node += ((int64_t)cc[k] + ((j >> k) & 1)) * mul;
      mul *= a.np1;
    }
    const int32_t g = (int32_t)(a.n_u + a.ppos[node]);
    s_pd[j] = g;
    s_pg[j] = a.eq[g];
  }
  __syncthreads();
  if (tid < NVL) {
    double f = 0.0;
    if (a.gravity_on) {
      const int n = tid / D, c = tid % D;
      const double gc = c == D - 1 ? a.grav : 0.0;
      for (int q = 0; q < NQ; ++q) f += a.rho_q[cell * NQ + q] * gc * a.Nq[q * NV + n] * s_jxw[q];
    }
    s_floc[tid] = f;
  }

  // velocity rows: A entries, own diagonal, Dirichlet columns into the rhs
  for (int i0 = 0; i0 < NVL; i0 += ASM_NW) {
    const int i = i0 + w;
    const bool act = i < NVL;
    const int32_t r = act ? s_vg[i] : -1;
    int64_t rb = 0;
    int len = 0;
    if (r >= 0) {
      rb = a.Arp[r];
      len = (int)(a.Arp[r + 1] - rb);
      if (len > ROWCAP) {
        if (lane == 0) *a.err = 1;
        len = ROWCAP;
      }
      for (int k = lane; k < len; k += 64) s_cols[w][k] = a.Aci[rb + k];
    }
    __syncthreads();
    if (act) {
      const int n = i / D, c = i % D;
      for (int j = lane; j < NVL; j += 64) {
        if (r < 0 && j != i) continue;
        const int m = j / D, e = j % D;
        double val = 0.0;
        for (int q = 0; q < NQ; ++q) {
          const double* gn = &s_dN[(q * NV + n) * D];
          const double* gm = &s_dN[(q * NV + m) * D];
          double dot = 0.0;
          for (int k = 0; k < D; ++k) dot += gn[k] * gm[k];
          const double v = (c == e ? dot : 0.0) + gn[e] * gm[c];
          const double G = v * s_jxw[q];
          val += s_eta[q] * G;
        }
        if (j == i) a.own[s_vd[i]] += val;
        if (r < 0) continue;
        const int32_t cg = s_vg[j];
        if (cg >= 0) {
          const int p = lds_lower_bound(s_cols[w], len, cg);
          if (p < len && s_cols[w][p] == cg) a.Av[rb + p] += val;
          else *a.err = 2;
        } else {
          s_tmp[w][j] = val * a.cval[s_vd[j]];
        }
      }
    }
    __syncthreads();
    if (act && r >= 0 && lane == 0) {
      double x = a.rhs[r];
      x += s_floc[i];
      for (int j = 0; j < NVL; ++j)
        if (s_vg[j] < 0) x -= s_tmp[w][j];
      a.rhs[r] = x;
    }
    __syncthreads();
  }

  // pressure rows: M_p entries and own diagonal, B entries, Dirichlet columns
  for (int i0 = 0; i0 < NP; i0 += ASM_NW) {
    const int i = i0 + w;
    const bool act = i < NP;
    const int32_t r = act ? s_pg[i] : -1;
    int64_t rb = 0;
    int len = 0;
    if (r >= 0) {
      rb = a.Brp[r - a.n_u];
      len = (int)(a.Brp[r - a.n_u + 1] - rb);
      if (len > ROWCAP) {
        if (lane == 0) *a.err = 1;
        len = ROWCAP;
      }
      for (int k = lane; k < len; k += 64) s_cols[w][k] = a.Bci[rb + k];
    }
    __syncthreads();
    if (act) {
      if (lane < NP) {
        const int j = lane;
        const bool scatter = r >= 0 && s_pg[j] >= 0;
        if (j == i || scatter) {
          double mv = 0.0;
          for (int q = 0; q < NQ; ++q) mv += s_ie[q] * a.P[q * NP + i] * a.P[q * NP + j] * s_jxw[q];
          if (j == i) a.own[s_pd[i]] += mv;
          if (scatter) {
            const int32_t key = (int32_t)(s_pg[j] - a.n_u);
            const int64_t mb = a.Mrp[r - a.n_u], me = a.Mrp[r - a.n_u + 1];
            int64_t lo = mb, hi = me;
            while (lo < hi) {
              const int64_t mid = (lo + hi) >> 1;
              if (a.Mci[mid] < key) lo = mid + 1;
              else hi = mid;
            }
            if (lo < me && a.Mci[lo] == key) a.Mv[lo] += mv;
            else *a.err = 3;
          }
        }
      }
      if (r >= 0)
        for (int j = lane; j < NVL; j += 64) {
          const double val = a.Bl[i * NVL + j];
          const int32_t cg = s_vg[j];
          if (cg >= 0) {
            const int p = lds_lower_bound(s_cols[w], len, cg);
            if (p < len && s_cols[w][p] == cg) a.Bv[rb + p] += val;
            else *a.err = 4;
          } else {
            s_tmp[w][j] = val * a.cval[s_vd[j]];
          }
        }
    }
    __syncthreads();
    if (act && r >= 0 && lane == 0) {
      double x = a.rhs[r];
      for (int j = 0; j < NVL; ++j)
        if (s_vg[j] < 0) x -= s_tmp[w][j];
      a.rhs[r] = x;
    }
    __syncthreads();
  }
}

__global__ void k_asm_fill(int64_t n, double* x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// per chunk of DIAG_CHUNK velocity rows: sum of |A_ii| over the free rows
__global__ void k_asm_diag_chunks(int64_t n_u, const int32_t* eq, const int64_t* dpos, const double* Av,
                                  double* part) {
#pragma clang fp contract(off)
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t r0 = c * DIAG_CHUNK;
  if (r0 >= n_u) return;
  const int64_t r1 = r0 + DIAG_CHUNK < n_u ? r0 + DIAG_CHUNK : n_u;
  double s = 0.0;
  for (int64_t r = r0; r < r1; ++r)
    if (eq[r] == r) s += fabs(Av[dpos[r]]);
  part[c] = s;
}

// the chunk sums in order, then the mean (1 when no row is free)
__global__ void k_asm_diag_total(int64_t nchunks, double* part, int64_t cnt) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += part[c];
  part[nchunks] = cnt ? s / (double)cnt : 1.0;
}

// constrained rows keep only their diagonal: the dof's own cell value (or
// the mean |A_ii|), rhs = diagonal * value for Dirichlet dofs; constrained
// pressure rows get K11 = mean |A_ii| and M_p = own value (or 1)
__global__ void k_asm_constrained(int64_t n_u, int64_t ntot, const int32_t* eq, const int8_t* kind,
                                  const double* cval, const double* own, const double* part_dp, const int64_t* Arp,
                                  double* Av, const int64_t* Mrp, double* Mv, double* K11v, double* rhs) {
#pragma clang fp contract(off)
  const double Dp = *part_dp;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ntot; g += (int64_t)gridDim.x * blockDim.x) {
    if (eq[g] == g) continue;
    if (g < n_u) {
      double D = own[g];
      if (D == 0.0) D = Dp;
      Av[Arp[g]] = D;
      rhs[g] = kind[g] == FS_DOF_DIRICHLET ? D * cval[g] : 0.0;
    } else {
      const int64_t q = g - n_u;
      K11v[q] = Dp;
      const double Dm = own[g];
      Mv[Mrp[q]] = Dm != 0.0 ? Dm : 1.0;
      rhs[g] = 0.0;
    }
  }
}

__global__ void k_asm_transpose_vals(int64_t nnz, const int64_t* src, const double* Bv, double* Btv) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    Btv[k] = Bv[src[k]];
}

struct SampleArgs {
  int N, field;
  int64_t ncell;
  double h;
  double eta_matrix, eta_incl, rho_matrix, rho_incl, radius;
  const double* xq;
  const double* ylay;  // layered field: interface height per (cell column, Gauss index)
  const FieldEllipsoid* ell;
  int n_ell;
  double* eta_q;
  double* rho_q;
};

// the problem's viscosity / density at every quadrature point, with the
// host generator's arithmetic (Field::eval in assemble.cpp)
template <int D>
__global__ void k_asm_sample(SampleArgs a) {
#pragma clang fp contract(off)
  constexpr int NQ = D == 2 ? 9 : 27;
  const int64_t tot = a.ncell * NQ;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cell = t / NQ;
    const int q = (int)(t - cell * NQ);
    int64_t cc[3] = {0, 0, 0}, u = cell;
    for (int k = 0; k < D; ++k) {
      cc[k] = u % a.N;
      u /= a.N;
    }
    double x[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) x[k] = ((double)cc[k] + a.xq[q * D + k]) * a.h;
    double eta = a.eta_matrix, rho = a.rho_matrix;
    if (a.field == FS_FIELD_INCLUSION) {
      double dist2 = 0, norm2 = 0;
      for (int k = 0; k < D; ++k) {
        dist2 += (x[k] - 0.5) * (x[k] - 0.5);
        norm2 += x[k] * x[k];
      }
      const double dist = sqrt(dist2), nrm = sqrt(norm2);
      if (dist < a.radius) {
        eta = a.eta_incl;
        rho = a.rho_incl;
      } else if (nrm > a.radius) {
        eta = a.eta_matrix;
        rho = a.rho_matrix;
      } else {
        eta = 0.5 * (a.eta_matrix + a.eta_incl);
        rho = 0.5 * (a.rho_matrix + a.rho_incl);
      }
    } else if (a.field == FS_FIELD_LAYERS) {
      const double yif = a.ylay[cc[0] * 3 + q % 3];
      if (x[D - 1] > yif) {
        eta = a.eta_incl;
        rho = a.rho_incl;
      }
    } else if (a.field == FS_FIELD_ELLIPSES) {
      double best = -1.0, scale = 1.0;
      for (int e = 0; e < a.n_ell; ++e) {
        const FieldEllipsoid& E = a.ell[e];
        double r2 = 0;
        for (int b = 0; b < D; ++b) {
          double s = 0;
          for (int k = 0; k < D; ++k) s += E.R[b][k] * (x[k] - E.c[k]);
          r2 += (s / E.ax[b]) * (s / E.ax[b]);
        }
        if (r2 <= 1.0 && E.logc > best) {
          best = E.logc;
          scale = E.scale;
        }
      }
      eta = best >= 0 ? a.eta_matrix * scale : a.eta_matrix;
    }
    a.eta_q[t] = eta;
    a.rho_q[t] = rho;
  }
}

int grid_of(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (n + BLOCK - 1) / BLOCK)); }

constexpr int ROWCAP_2D = 128, ROWCAP_3D = 512;

void check_device(const fs_assembler* a, const char* what) {
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != a->device)
    throw FsError(FS_ERR_BAD_INPUT, std::string(what) + ": called with another device current than at create");
}

#define FS_GUARD_BEGIN try {
#define FS_GUARD_END                         \
  }                                          \
  catch (const FsError& e) {                 \
    fs::set_error("%s", e.what());           \
    return e.code;                           \
  }                                          \
  catch (const std::bad_alloc&) {            \
    fs::set_error("host allocation failed"); \
    return FS_ERR_OUT_OF_MEMORY;             \
  }                                          \
  catch (const std::exception& e) {          \
    fs::set_error("%s", e.what());           \
    return FS_ERR_BAD_INPUT;                 \
  }

}  // namespace

extern "C" {

int fs_assembler_create(const fs_system* S, fs_assembler** out) {
  if (!S || !out) {
    fs::set_error("fs_assembler_create: null argument");
    return FS_ERR_BAD_INPUT;
  }
  *out = nullptr;
  FS_GUARD_BEGIN
  ensure_device();
  std::unique_ptr<fs_assembler> a(new fs_assembler());
  const int d = S->dim;
  if (d != 2 && d != 3) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: dim must be 2 or 3");
  hip_check(hipGetDevice(&a->device), "hipGetDevice");
  a->dim = d;
  a->N = S->N;
  a->prob = S->prob;
  a->h = 1.0 / (double)S->N;
  a->n_u = S->n_u;
  a->n_p = S->n_p;
  const int64_t ntot = S->n_u + S->n_p;
  if (ntot >= (int64_t)INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "fs_assembler_create: more than 2^31 dofs");
  a->ncell = 1;
  for (int k = 0; k < d; ++k) a->ncell *= S->N;
  const int64_t nv1 = 2 * (int64_t)S->N + 1, np1 = (int64_t)S->N + 1;
  int64_t nvn = 1, npn = 1;
  for (int k = 0; k < d; ++k) {
    nvn *= nv1;
    npn *= np1;
  }
  if (nvn * d != S->n_u || npn != S->n_p) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: inconsistent system");

  // node lattice index -> dof-order position
  std::vector<int32_t> vpos(nvn, -1), ppos(npn, -1);
  for (int64_t i = 0; i < nvn; ++i) vpos[S->vnode[i * d]] = (int32_t)i;
  for (int64_t j = 0; j < npn; ++j) ppos[S->pnode[j]] = (int32_t)j;
  for (int64_t i = 0; i < nvn; ++i)
    if (vpos[i] < 0) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: velocity dof map is not a permutation");
  for (int64_t j = 0; j < npn; ++j)
    if (ppos[j] < 0) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: pressure dof map is not a permutation");
  // equation of each dof: -1 when eliminated, the master's when periodic
  std::vector<int32_t> eq(ntot);
  for (int64_t g = 0; g < ntot; ++g) {
    const int8_t k = S->kind[g];
    eq[g] = (k == FS_DOF_DIRICHLET || k == FS_DOF_PINNED) ? -1 : (int32_t)g;
  }
  for (int64_t g = 0; g < ntot; ++g)
    if (S->kind[g] == FS_DOF_PERIODIC) eq[g] = eq[S->master[g]];

  const HostCsr &A = S->A, &B = S->B, &M = S->Mp, &Bt = S->Bt;
  a->nnzA = A.nnz();
  a->nnzB = B.nnz();
  a->nnzBt = Bt.nnz();
  a->nnzM = M.nnz();
  int max_row = 0;
  for (int64_t r = 0; r < A.n_rows; ++r) max_row = std::max<int>(max_row, (int)(A.rp[r + 1] - A.rp[r]));
  for (int64_t r = 0; r < B.n_rows; ++r) max_row = std::max<int>(max_row, (int)(B.rp[r + 1] - B.rp[r]));
  a->max_row = max_row;
  if (max_row > (d == 2 ? ROWCAP_2D : ROWCAP_3D))
    throw FsError(FS_ERR_UNSUPPORTED, "fs_assembler_create: a row is longer than the assembly kernel stages");
  // diagonal position of each velocity row (free rows only are read)
  std::vector<int64_t> dpos(S->n_u);
  int64_t n_free = 0;
  for (int64_t r = 0; r < S->n_u; ++r) {
    dpos[r] = A.rp[r];
    if (eq[r] != r) continue;
    ++n_free;
    const auto b = A.ci.begin() + A.rp[r], e = A.ci.begin() + A.rp[r + 1];
    const auto it = std::lower_bound(b, e, (int32_t)r);
    if (it == e || *it != (int32_t)r) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: A has no diagonal entry");
    dpos[r] = it - A.ci.begin();
  }
  a->n_free_u = n_free;
  // B^T entry -> B entry (the counting transpose, rows by ascending pressure row)
  std::vector<int64_t> bt_src(std::max<int64_t>(1, a->nnzBt));
  {
    if (Bt.n_rows != S->n_u || a->nnzBt != a->nnzB) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: B^T mismatch");
    std::vector<int64_t> next(Bt.rp.begin(), Bt.rp.end() - 1);
    for (int64_t q = 0; q < B.n_rows; ++q)
      for (int64_t e = B.rp[q]; e < B.rp[q + 1]; ++e) {
        const int64_t k = next[B.ci[e]]++;
        if (k >= Bt.rp[B.ci[e] + 1] || Bt.ci[k] != q)
          throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_create: B^T is not the transpose of B");
        bt_src[k] = e;
      }
  }
  // reference-element tables in one buffer
  const RefTables R = ref_tables(d, a->h);
  a->nq = R.nq;
  std::vector<double> tab;
  auto put = [&](const std::vector<double>& v) {
    const int64_t o = (int64_t)tab.size();
    tab.insert(tab.end(), v.begin(), v.end());
    return o;
  };
  a->off_dN = put(R.dN);
  a->off_N = put(R.N);
  a->off_P = put(R.P);
  a->off_JxW = put(R.JxW);
  a->off_xq = put(R.xq);
  a->off_Bl = put(R.Bl);
  // field description
  std::vector<double> ylay((size_t)S->N * 3, 0.0);
  if (S->prob.field == FS_FIELD_LAYERS)
    for (int64_t c0 = 0; c0 < S->N; ++c0)
      for (int g = 0; g < 3; ++g) ylay[c0 * 3 + g] = layer_interface(S->prob, (c0 + R.xq[(size_t)g * d]) * a->h);
  std::vector<FieldEllipsoid> ell = field_ellipsoids(S->prob);
  a->n_ell = (int)ell.size();

  a->vpos.upload(vpos.data(), 4 * vpos.size());
  a->ppos.upload(ppos.data(), 4 * ppos.size());
  a->eq.upload(eq.data(), 4 * eq.size());
  a->kind.upload(S->kind.data(), ntot);
  a->cval.upload(S->cval.data(), 8 * ntot);
  a->dpos.upload(dpos.data(), 8 * std::max<int64_t>(1, S->n_u));
  a->Arp.upload(A.rp.data(), 8 * A.rp.size());
  a->Aci.upload(A.ci.data(), 4 * std::max<int64_t>(1, a->nnzA));
  a->Brp.upload(B.rp.data(), 8 * B.rp.size());
  a->Bci.upload(B.ci.data(), 4 * std::max<int64_t>(1, a->nnzB));
  a->Mrp.upload(M.rp.data(), 8 * M.rp.size());
  a->Mci.upload(M.ci.data(), 4 * std::max<int64_t>(1, a->nnzM));
  a->bt_src.upload(bt_src.data(), 8 * bt_src.size());
  a->tabs.upload(tab.data(), 8 * tab.size());
  a->ylay.upload(ylay.data(), 8 * std::max<size_t>(1, ylay.size()));
  if (!ell.empty()) a->ell.upload(ell.data(), sizeof(FieldEllipsoid) * ell.size());
  else a->ell.alloc(sizeof(FieldEllipsoid));
  a->own.alloc(8 * ntot);
  const int64_t nchunks = (S->n_u + DIAG_CHUNK - 1) / DIAG_CHUNK;
  a->part.alloc(8 * (nchunks + 1));
  a->err.alloc(sizeof(int));
  hip_check(hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking), "stream");
  hip_check(hipEventCreate(&a->e0), "event");
  hip_check(hipEventCreate(&a->e1), "event");
  *out = a.release();
  return FS_OK;
  FS_GUARD_END
}

int fs_assembler_sample_field(fs_assembler* a, double* d_eta_q, double* d_rho_q) {
  if (!a || !d_eta_q || !d_rho_q) {
    fs::set_error("fs_assembler_sample_field: null argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  check_device(a, "fs_assembler_sample_field");
  hip_check(hipStreamSynchronize(nullptr), "sync");
  SampleArgs s;
  s.N = a->N;
  s.field = a->prob.field;
  s.ncell = a->ncell;
  s.h = a->h;
  s.eta_matrix = a->prob.eta_matrix;
  s.eta_incl = a->prob.eta_incl;
  s.rho_matrix = a->prob.rho_matrix;
  s.rho_incl = a->prob.rho_incl;
  s.radius = a->prob.radius;
  s.xq = a->tabs.as<double>() + a->off_xq;
  s.ylay = a->ylay.as<double>();
  s.ell = a->ell.as<FieldEllipsoid>();
  s.n_ell = a->n_ell;
  s.eta_q = d_eta_q;
  s.rho_q = d_rho_q;
  const int g = grid_of(a->ncell * a->nq);
  if (a->dim == 2) hipLaunchKernelGGL(k_asm_sample<2>, dim3(g), dim3(BLOCK), 0, a->st, s);
  else hipLaunchKernelGGL(k_asm_sample<3>, dim3(g), dim3(BLOCK), 0, a->st, s);
  hip_check(hipGetLastError(), "sample launch");
  hip_check(hipStreamSynchronize(a->st), "sync");
  return FS_OK;
  FS_GUARD_END
}

int fs_assembler_assemble(fs_assembler* a, const double* d_eta_q, const double* d_rho_q, double* d_A, double* d_Bt,
                          double* d_B, double* d_K11, double* d_Mp, double* d_rhs) {
  if (!a || !d_eta_q || !d_rho_q || !d_A || !d_Bt || !d_B || !d_K11 || !d_Mp || !d_rhs) {
    fs::set_error("fs_assembler_assemble: null argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  check_device(a, "fs_assembler_assemble");
  hip_check(hipStreamSynchronize(nullptr), "sync");
  const int d = a->dim;
  const int64_t ntot = a->n_u + a->n_p;
  hip_check(hipEventRecord(a->e0, a->st), "event record");
  auto zero = [&](double* p, int64_t n) {
    hipLaunchKernelGGL(k_asm_fill, dim3(grid_of(n)), dim3(BLOCK), 0, a->st, n, p);
  };
  zero(d_A, a->nnzA);
  zero(d_B, a->nnzB);
  zero(d_Mp, a->nnzM);
  zero(d_K11, a->n_p);
  zero(d_rhs, ntot);
  zero(a->own.as<double>(), ntot);
  zero(a->part.as<double>(), 1);  // (the error word shares no storage with it)
  hip_check(hipMemsetAsync(a->err.p, 0, sizeof(int), a->st), "memset");

  AsmArgs g;
  g.N = a->N;
  g.nv1 = 2 * (int64_t)a->N + 1;
  g.np1 = (int64_t)a->N + 1;
  g.n_u = a->n_u;
  g.vpos = a->vpos.as<int32_t>();
  g.ppos = a->ppos.as<int32_t>();
  g.eq = a->eq.as<int32_t>();
  g.cval = a->cval.as<double>();
  g.Arp = a->Arp.as<int64_t>();
  g.Aci = a->Aci.as<int32_t>();
  g.Av = d_A;
  g.Brp = a->Brp.as<int64_t>();
  g.Bci = a->Bci.as<int32_t>();
  g.Bv = d_B;
  g.Mrp = a->Mrp.as<int64_t>();
  g.Mci = a->Mci.as<int32_t>();
  g.Mv = d_Mp;
  g.rhs = d_rhs;
  g.own = a->own.as<double>();
  g.eta_q = d_eta_q;
  g.rho_q = d_rho_q;
  const double* T = a->tabs.as<double>();
  g.dN = T + a->off_dN;
  g.Nq = T + a->off_N;
  g.P = T + a->off_P;
  g.JxW = T + a->off_JxW;
  g.Bl = T + a->off_Bl;
  g.grav = -a->prob.gravity;
  g.gravity_on = a->prob.gravity != 0.0;
  g.err = a->err.as<int>();
  for (int color = 0; color < (1 << d); ++color) {
    g.color = color;
    int64_t cnt = 1;
    for (int k = 0; k < 3; ++k) {
      g.m[k] = k < d ? (a->N - ((color >> k) & 1) + 1) / 2 : 1;
      cnt *= g.m[k];
    }
    if (cnt == 0) continue;
    if (d == 2) hipLaunchKernelGGL((k_asm_cells<2, ROWCAP_2D>), dim3(cnt), dim3(ASM_NT), 0, a->st, g);
    else hipLaunchKernelGGL((k_asm_cells<3, ROWCAP_3D>), dim3(cnt), dim3(ASM_NT), 0, a->st, g);
  }
  const int64_t nchunks = (a->n_u + DIAG_CHUNK - 1) / DIAG_CHUNK;
  hipLaunchKernelGGL(k_asm_diag_chunks, dim3((nchunks + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, a->st, a->n_u,
                     a->eq.as<int32_t>(), a->dpos.as<int64_t>(), (const double*)d_A, a->part.as<double>());
  hipLaunchKernelGGL(k_asm_diag_total, dim3(1), dim3(64), 0, a->st, nchunks, a->part.as<double>(), a->n_free_u);
  hipLaunchKernelGGL(k_asm_constrained, dim3(grid_of(ntot)), dim3(BLOCK), 0, a->st, a->n_u, ntot,
                     a->eq.as<int32_t>(), a->kind.as<int8_t>(), a->cval.as<double>(), a->own.as<double>(),
                     a->part.as<double>() + nchunks, a->Arp.as<int64_t>(), d_A, a->Mrp.as<int64_t>(), d_Mp, d_K11,
                     d_rhs);
  hipLaunchKernelGGL(k_asm_transpose_vals, dim3(grid_of(a->nnzBt)), dim3(BLOCK), 0, a->st, a->nnzBt,
                     a->bt_src.as<int64_t>(), (const double*)d_B, d_Bt);
  hip_check(hipGetLastError(), "assembly launch");
  hip_check(hipEventRecord(a->e1, a->st), "event record");
  hip_check(hipStreamSynchronize(a->st), "sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, a->e0, a->e1), "elapsed");
  a->last_ms = ms;
  int err = 0;
  hip_check(hipMemcpy(&err, a->err.p, sizeof(int), hipMemcpyDeviceToHost), "copy");
  if (err) throw FsError(FS_ERR_BAD_INPUT, "fs_assembler_assemble: an element entry has no place in the pattern (code " +
                                               std::to_string(err) + ")");
  return FS_OK;
  FS_GUARD_END
}

double fs_assembler_last_ms(const fs_assembler* a) { return a ? a->last_ms : 0.0; }

void fs_assembler_destroy(fs_assembler* a) { delete a; }

}  // extern "C"
