// fs_handle: the Stokes solve behind include/fs_stokes.h.
//
// Reference flow (src/pde/stokes_ilu.cc:442-561): optional diagonal scaling,
// block preconditioner reinitialisation (numeric ILU of A and of M_p), GMRES
// with the block-Schur preconditioner, unscaling. Two Krylov drivers:
//   FS_METHOD_GMRES_REF  left-preconditioned restarted GMRES with
//                        ReductionControl(rtol ||b||, reduce) on the
//                        preconditioned residual (deal.II semantics);
//   FS_METHOD_FGMRES     right-preconditioned flexible GMRES whose
//                        preconditioner may run in FP32 (FS_PREC_MIXED); each
//                        restart begins from the FP64 true residual, so the
//                        outer loop is iterative refinement, and the solve
//                        stops when ||b - K x|| <= rtol ||b|| in FP64.
// Orthogonalisation is classical Gram-Schmidt with one re-orthogonalisation
// (two multi-dot passes per step instead of j+1 dependent dot products).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fs_stokes.h"
#include "guess_internal.h"
#include "solver_internal.h"

using namespace fsk;
using namespace fs;

namespace {

// device scalar slots
constexpr int SC_H1 = 0;                    // [MAXM]
constexpr int SC_H2 = SC_H1 + MAXM;         // [MAXM]
constexpr int SC_HCOL = SC_H2 + MAXM;       // [MAXM+1]
constexpr int SC_CS = SC_HCOL + MAXM + 1;   // [MAXM]
constexpr int SC_SN = SC_CS + MAXM;         // [MAXM]
constexpr int SC_GAM = SC_SN + MAXM;        // [MAXM+1]
constexpr int SC_Y = SC_GAM + MAXM + 1;     // [MAXM]
constexpr int SC_NRM = SC_Y + MAXM;         // 1
constexpr int SC_EST = SC_NRM + 1;          // 1
constexpr int SC_TMP = SC_EST + 1;          // 4
constexpr int SC_PCG = SC_TMP + 4;          // 4: p.q, r.z (two slots)
constexpr int SC_DC = SC_PCG + 4;           // 4: DCGS2 1 / r, v_j.w, saved gamma_j
constexpr int SC_RAW0 = SC_DC + 4;          // [MAXM+1] DCGS2 corrected previous column
constexpr int SC_RAW1 = SC_RAW0 + MAXM + 1; // [MAXM+1] DCGS2 raw current column
constexpr int SC_H = SC_RAW1 + MAXM + 1;    // [(MAXM+1) * MAXM]
constexpr int SC_TOTAL = SC_H + (MAXM + 1) * MAXM;

struct Ctrl {  // dealii::ReductionControl::check (documented semantics)
  int max_steps = 100;
  double tol = 0, reduce = 0, reduced_tol = 0;
  int last_step = 0;
  double last_value = 0;
  enum State { ITERATE, SUCCESS, FAILURE };
  State check(int step, double v) {
    if (step == 0) reduced_tol = v * reduce;
    last_step = step;
    last_value = v;
    if (v <= reduced_tol) return SUCCESS;
    if (v <= tol) return SUCCESS;
    if (step >= max_steps || std::isnan(v)) return FAILURE;
    return ITERATE;
  }
};

template <typename T>
T* dptr(const DevBuf& b, int64_t off = 0) {
  return (T*)b.p + off;
}

// StokesScalingOperator::initialize (stokes_common.implement.h:13-53), rows
// of K in solver order: D[r] = sqrt(max_c |A_rc|) for the velocity rows;
// per-workgroup partials of sum 1/D^2 and of sum (max_c |B^T_rc|)^2
__global__ __launch_bounds__(BLOCK) void k_scale_coeffs_u(int64_t nu, const int64_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci,
                                                          const double* __restrict__ v, double* __restrict__ D,
                                                          double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  double s1 = 0, s2 = 0;
  for (int64_t r = (int64_t)blockIdx.x * BLOCK + threadIdx.x; r < nu; r += (int64_t)gridDim.x * BLOCK) {
    double ma = 0, mb = 0;
    for (int64_t p = rp[r]; p < rp[r + 1]; ++p) {
      const double a = fabs(v[p]);
      if (ci[p] < nu) ma = fmax(ma, a);
      else mb = fmax(mb, a);
    }
    D[r] = sqrt(ma);
    s1 += 1.0 / ma;
    s2 += mb * mb;
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// pressure coefficient s_p = P sqrt(P g) / n_p, P = sum 1/D_u^2, g = sum ghat^2,
// written to D[nu .. nu+np) (one workgroup; partials summed in index order)
__global__ __launch_bounds__(BLOCK) void k_scale_coeffs_p(int nb, const double* __restrict__ part, int64_t nu,
                                                          int64_t np, double* __restrict__ D) {
  __shared__ double sh[BLOCK / WAVE];
  __shared__ double sp;
  double s1 = 0, s2 = 0;
  for (int i = threadIdx.x; i < nb; i += BLOCK) {
    s1 += part[2 * i];
    s2 += part[2 * i + 1];
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) sp = s1 * sqrt(s1 * s2) / (double)np;
  __syncthreads();
  for (int64_t i = threadIdx.x; i < np; i += BLOCK) D[nu + i] = sp;
}

// y = x * D (inv = 0) or x / D (inv = 1); y may alias x
template <typename TA, typename TB>
__global__ __launch_bounds__(BLOCK) void k_mul_diag(int64_t n, const TA* x, const double* __restrict__ D, TB* y,
                                                    int inv) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    y[i] = (TB)(inv ? (double)x[i] / D[i] : (double)x[i] * D[i]);
}

// y = bs - y / D
__global__ __launch_bounds__(BLOCK) void k_resid_scaled(int64_t n, const double* __restrict__ bs,
                                                        const double* __restrict__ D, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    y[i] = bs[i] - y[i] / D[i];
}

// y = a x + b y
__global__ __launch_bounds__(BLOCK) void k_axpby(int64_t n, double a, const double* __restrict__ x, double b,
                                                 double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

struct GuessNorm {
  float mean[4], stdv[4];
  int64_t P;  // lattice points per output channel
};

// x[r] = network output at row r's lattice node and channel, denormalised
__global__ __launch_bounds__(BLOCK) void k_guess_scatter(int64_t n, const int32_t* __restrict__ idx,
                                                         const float* __restrict__ Y, GuessNorm gn,
                                                         double* __restrict__ x) {
  for (int64_t r = (int64_t)blockIdx.x * BLOCK + threadIdx.x; r < n; r += (int64_t)gridDim.x * BLOCK) {
    const int32_t id = idx[r];
    const int ch = (int)(id / gn.P);
    x[r] = (double)(Y[id] * gn.stdv[ch] + gn.mean[ch]);
  }
}

// rows holding only their diagonal (constrained dofs): x = b / K_rr
__global__ __launch_bounds__(BLOCK) void k_diag_rows(int64_t nr, const int32_t* __restrict__ rows,
                                                     const int64_t* __restrict__ rp, const double* __restrict__ kv,
                                                     const double* __restrict__ b, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nr; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t r = rows[i];
    x[r] = b[r] / kv[rp[r]];
  }
}

// Velocity hand-off (VelocityExtractor, stokes_common.implement.h:582-591):
// u[vlat[i]] = x[i] for the velocity dofs i (caller's order); vlat = comp * P
// + node on the Q2 node lattice of P = (2N+1)^d nodes, x fastest
__global__ __launch_bounds__(BLOCK) void k_velocity_lattice(int64_t nu, const int32_t* __restrict__ vlat,
                                                            const double* __restrict__ x, double* __restrict__ u) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nu; i += (int64_t)gridDim.x * BLOCK)
    u[vlat[i]] = x[i];
}

// The Q2 velocity field at points of [0,1]^D (FEValues get_function_values
// on the uniform mesh of N^D cells): the point's cell, its local coordinates
// t in [0,1]^D, and the tensor product of the 1D quadratic Lagrange basis on
// the cell's nodes 0, 1/2, 1 (lattice offsets 0, 1, 2). Points outside the
// unit cube are clamped to it. pts and vel: D values per point.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_velocity_at_points(int64_t np, const double* __restrict__ pts, int N,
                                                              const double* __restrict__ u, int64_t P,
                                                              double* __restrict__ vel) {
  const int64_t n1 = 2 * (int64_t)N + 1;
  for (int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x; q < np; q += (int64_t)gridDim.x * BLOCK) {
    double l[D][3];
    int64_t base = 0, stride = 1;
    for (int a = 0; a < D; ++a) {
      const double xa = fmin(fmax(pts[q * D + a], 0.0), 1.0) * N;
      const int c = min((int)floor(xa), N - 1);
      const double t = xa - c;
      l[a][0] = 2.0 * (t - 0.5) * (t - 1.0);
      l[a][1] = -4.0 * t * (t - 1.0);
      l[a][2] = 2.0 * t * (t - 0.5);
      base += 2 * (int64_t)c * stride;
      stride *= n1;
    }
    double acc[D];
    for (int c = 0; c < D; ++c) acc[c] = 0.0;
    int o[3] = {0, 0, 0};
    for (int k = 0; k < (D == 2 ? 9 : 27); ++k) {
      int kk = k;
      double w = 1.0;
      int64_t idx = base, st2 = 1;
      for (int a = 0; a < D; ++a) {
        o[a] = kk % 3;
        kk /= 3;
        w *= l[a][o[a]];
        idx += o[a] * st2;
        st2 *= n1;
      }
      for (int c = 0; c < D; ++c) acc[c] += w * u[c * P + idx];
    }
    for (int c = 0; c < D; ++c) vel[q * D + c] = acc[c];
  }
}

}  // namespace

// dst[map[e]] = src[e] / dst[e] = src[map[e]]
static __global__ __launch_bounds__(BLOCK) void k_values_scatter(int64_t n, const int64_t* __restrict__ map,
                                                          const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < n; e += (int64_t)gridDim.x * BLOCK)
    dst[map[e]] = src[e];
}
static __global__ __launch_bounds__(BLOCK) void k_values_gather(int64_t n, const int64_t* __restrict__ map,
                                                         const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < n; e += (int64_t)gridDim.x * BLOCK)
    dst[e] = src[map[e]];
}

struct fs_handle {
  fs_opts o{};
  int64_t nu = 0, np = 0, n = 0, ld = 0;
  hipStream_t st = nullptr;
  DevCsr K, Mp, Ablk;                                     // K, M_p and A alone, solver order
  BlkCsr Ablb;                                            // A by the ILU's lattice blocks (inner PCG SpMV)
  BlkCsr Kblb;  // K's velocity rows (A and B^T) by the same blocks, FP64 (the outer method's K z)
  // value maps, caller entry -> K / Mp position and A entry -> K position:
  // built on the host at create, then kept on the device only
  std::vector<int64_t> mapA, mapBt, mapB, mapK11, mapMp, ablk_from_k;
  int64_t nmapA = 0, nmapBt = 0, nmapB = 0, nmapK11 = 0, nmapMp = 0, nablk = 0;
  DevBuf dmapA, dmapBt, dmapB, dmapK11, dmapMp, dablk;
  IluDev FA, FS;
  DevBuf b, x, r, V, Z, ytmp, tp, cg, partial, sc, gbar, flags, perm_all, vin, pcg, pcg_part;
  int64_t ldu = 0;
  int m_alloc = 0;          // Krylov basis vectors allocated (largest restart a solve may use)
  double* h_pin = nullptr;  // pinned scalars
  bool dirty = true;
  Counters cnt;
  DevBuf scr_ws;  // Schur-complement reduction workspace (FS_METHOD_SCR)
  // diagonal scaling (StokesSolverControl::apply_diagonal_scaling): while a
  // scaled solve runs, Dsc points at D (solver order) and the Krylov methods
  // see K~ = D^-1 K D^-1 and P~^-1 = D P^-1 D on x~ = D x, b~ = D^-1 b
  const double* Dsc = nullptr;
  DevBuf dsc_buf;
  // learned guess (FS_GUESS_LEARNED)
  fs_guess_model* model = nullptr;
  int guess_nc = 0;
  DevBuf g_logeta, g_idx, g_diag_rows;
  // node lattice of the caller's velocity dofs (fs_stokes_set_dof_lattice)
  int lat_dim = 0, lat_nc = 0;
  DevBuf v_lat;
  std::vector<int32_t> diag_only;  // solver rows whose K row is the diagonal alone

  ~fs_handle() {
    if (h_pin) (void)hipHostFree(h_pin);
    if (st) (void)hipStreamDestroy(st);
  }

  double* S(int off) const { return dptr<double>(sc, off); }

  void sync() { hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"); }
  double read_scalar(int off) {
    hip_check(hipMemcpyAsync(h_pin, S(off), 8, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
    sync();
    return h_pin[0];
  }

  // ||v||_2 of an FP64 device vector into sc[off]
  void norm(const double* v, int64_t len, int off) {
    const int g = std::min<int>(RED_BLOCKS, (int)std::max<int64_t>(1, (len + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL((k_dot_partial<double>), dim3(g), dim3(BLOCK), 0, st, len, v, v, dptr<double>(partial));
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, st, g, dptr<double>(partial), S(off), 1);
    cnt.launches += 2;
    cnt.bytes += 8.0 * len;
  }

  // ||D^-1 v||_2 into sc[off] with D = Dv (the scaled system's residual
  // norm), ||v||_2 when Dv is null; D^-1 v goes through dsc_t()
  void wnorm(const double* v, int64_t len, const double* Dv, int off) {
    if (!Dv) { norm(v, len, off); return; }
    double* t = dsc_t();
    hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(len)), dim3(BLOCK), 0, st, len, v, Dv, t, 1);
    cnt.launches++;
    cnt.bytes += 24.0 * len;
    norm(t, len, off);
  }

  // y = K x (FP64) or y = b - K x
  void kmult64(const double* xv, double* y, const double* bv) {
    if (bv) spmv4_launch<double, double, double, double, double, 1>(K, n, 0, xv, y, bv, 0, st);
    else spmv4_launch<double, double, double, double, double, 0>(K, n, 0, xv, y, (const double*)nullptr, 0, st);
    cnt.launches++;
    cnt.bytes += K.nnz * 12.0 + 8.0 * (n + 1) + 8.0 * n * (bv ? 3 : 2);
  }

  // inner CG on M_p (solver order)
  template <typename T>
  void inner_cg(const T* bp, T* xp) {
    InnerCgArgs<T> a;
    a.n = np;
    a.mrp = Mp.rp.as<int64_t>();
    a.mci = Mp.ci.as<int32_t>();
    a.mv = std::is_same<T, double>::value ? (const T*)Mp.v64.p : (const T*)Mp.v32.p;
    a.F = FS.view<T>();
    a.stage_ptr = FS.stage_ptr.as<int32_t>();
    a.n_stages = FS.sched.n_stages();
    a.max_rows = FS.max_rows;
    a.cap = FS.cap;
    a.b = bp;
    a.x = xp;
    T* w = (T*)cg.p;
    const int64_t lp = (np + 63) & ~int64_t(63);
    a.g = w;
    a.h = w + lp;
    a.d = w + 2 * lp;
    a.q = w + 3 * lp;
    a.partial = dptr<double>(partial);
    const bool mixed = o.precision != FS_PREC_FP64;
    // InverseMatrixCG's ReductionControl(tol, reduce) for the reference-mode
    // GMRES and SCR; the flexible method only needs a relative reduction
    a.tol = (mixed || o.method == FS_METHOD_FGMRES) ? 0.0 : o.inner_cg_tol;
    a.reduce = mixed ? o.inner_cg_tol_mixed : o.inner_cg_tol;
    a.max_it = mixed ? (o.inner_cg_max_mixed > 0 ? o.inner_cg_max_mixed : (int)np)
                     : (o.inner_cg_max > 0 ? o.inner_cg_max : (int)np);
    a.iters_out = dptr<int>(flags, 1);
    a.fail_out = dptr<int>(flags, 2);
    a.gb = dptr<GridBarrier>(gbar);
    a.timeout = timeout_ticks();
    const int gm = Mp.G <= 8 ? 8 : Mp.G <= 16 ? 16 : 32;
    if (a.max_it == 1) {
      inner_cg_step<T>(a, gm);
      return;
    }
    const size_t lds = lds_bytes_per_wave(FS.max_rows, FS.cap, sizeof(T));
    void* fn;
    fn = pick_cg_kernel<T>(gm, FS.G);
    int per_cu = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BLOCK, lds), "occupancy");
    int nwg = (int)std::min<int64_t>((int64_t)std::max(1, per_cu) * device_cus(),
                                     std::max<int64_t>(1, (np + 1023) / 1024));
    nwg = std::max(1, std::min(nwg, 256));
    void* args[] = {&a};
    hip_check(hipLaunchCooperativeKernel(fn, dim3(nwg), dim3(BLOCK), args, (unsigned)lds, st),
              "inner CG cooperative launch");
    cnt.launches++;
    cnt.bytes += (Mp.nnz + FS.nnz) * (sizeof(T) + 4.0) + 8.0 * np * sizeof(T);
  }

  // The inner CG capped at one step, by ordinary launches: the cooperative
  // kernel's grid barriers (two per colour stage of ILU_S and one per
  // reduction) cost more than the step's work. Same arithmetic: d = ILU_S^{-1}
  // b, q = M d, alpha = b.d / d.q, x = alpha d; the step count and the
  // failure flag as the cooperative kernel sets them.
  template <typename T>
  void inner_cg_step(const InnerCgArgs<T>& a, int gm) {
    FS.apply<T, T, T>(a.b, a.h, a.d, st, cnt);
    double* part = dptr<double>(partial);
    const int rows = BLOCK / gm;
    const int nb1 = (int)std::min<int64_t>(CG_STEP_WG, std::max<int64_t>(1, (np + rows - 1) / rows));
    const int nb2 = (int)std::min<int64_t>(CG_STEP_WG, std::max<int64_t>(1, (np + BLOCK - 1) / BLOCK));
#define FS_CGS(GM)                                                                                            \
  hipLaunchKernelGGL((k_cg_step_spmv<T, GM>), dim3(nb1), dim3(BLOCK), 0, st, np, a.mrp, a.mci, a.mv, a.b, a.d, \
                     a.q, part);
    if (gm == 8) FS_CGS(8) else if (gm == 16) FS_CGS(16) else FS_CGS(32)
#undef FS_CGS
    hipLaunchKernelGGL((k_cg_step_x<T>), dim3(nb2), dim3(BLOCK), 0, st, np, nb1, a.b, a.d, a.q, a.x, part, a.tol,
                       a.reduce);
    hipLaunchKernelGGL(k_cg_step_end, dim3(1), dim3(BLOCK), 0, st, nb1, nb2, (const double*)part, a.tol, a.reduce,
                       a.iters_out, a.fail_out);
    cnt.launches += 3;
    cnt.bytes += Mp.nnz * (sizeof(T) + 4.0) + 8.0 * (np + 1) + 6.0 * np * sizeof(T);
  }

  template <typename T>
  static void* pick_cg_kernel(int gm, int gf) {
#define FS_CG(GM, GF) \
  if (gm == GM && gf == GF) return (void*)&k_inner_cg<T, GM, GF>;
    FS_CG(8, 4) FS_CG(8, 8) FS_CG(8, 16) FS_CG(8, 32) FS_CG(8, 64)
    FS_CG(16, 4) FS_CG(16, 8) FS_CG(16, 16) FS_CG(16, 32) FS_CG(16, 64)
    FS_CG(32, 4) FS_CG(32, 8) FS_CG(32, 16) FS_CG(32, 32) FS_CG(32, 64)
#undef FS_CG
    return (void*)&k_inner_cg<T, 16, 16>;
  }

  // BlockSchurPreconditionerBase::vmult (linear_operators.implement.h:200-209):
  // dst_u = ILU_A^{-1} src_u ; t = B dst_u - src_p ; dst_p = CG(M_p, ILU_S; t)
  // A^{-1} src_u: o.inner_a_iters steps of ILU(k)-preconditioned CG from x = 0,
  // in Chronopoulos-Gear form: one step is the ILU application z = M^-1 r
  // (its forward stages first apply the previous step's update to the rows
  // they read), the SpMV s = A z with z.s and r.z partials, and one
  // workgroup forming alpha and beta from both sums.
  // The same iterates as the textbook PCG (q = A p by recurrence), with one
  // reduction instead of two and no direction kernel. Scalars never leave the
  // device; sums are taken in a fixed order.
  template <typename TV, typename T>
  void inner_pcg_A(const double* src, T* x, int iters) {
    T* base = (T*)pcg.p;
    T *r = base, *z = base + ldu, *pv = base + 2 * ldu, *q = base + 3 * ldu, *y = base + 4 * ldu;
    T* sv = base + 5 * ldu;
    double* part_zs = dptr<double>(pcg_part);
    double* part_rz = part_zs + spmv_dot_grid();
    double* sc = S(SC_PCG);  // alpha, beta, gamma
    const int nd = spmv_dot_grid();
    hipLaunchKernelGGL((k_copy_convert<double, T>), dim3(512), dim3(BLOCK), 0, st, nu, src, r);
    cnt.launches++;
    for (int it = 0; it < iters; ++it) {
      // x, r (and p, q) of step it-1 inside the forward stages; z = M^-1 r
      const PcgFuse<T> pf{sc, sc + 1, z, sv, pv, q, x, r, it == 1 ? 1 : 0};
      FA.apply<TV, T, T>(r, y, z, st, cnt, it > 0 ? &pf : nullptr);
      launch_spmv_dot<TV, T>(z, sv, part_zs, r, part_rz);
      hipLaunchKernelGGL(k_cg1_scalars, dim3(1), dim3(1024), 0, st, nd, part_rz, nd, part_zs, sc, it == 0 ? 1 : 0);
      cnt.launches += 2;
      cnt.bytes += ablk_spmv_bytes(sizeof(TV), sizeof(T), true) + sizeof(T) * (it > 0 ? 10.0 : 1.0) * nu;
    }
    // the last step's x (p is not kept)
    hipLaunchKernelGGL((k_cg1_update<T>), dim3(256), dim3(BLOCK), 0, st, nu, sc, sc + 1, z, (const T*)nullptr, pv,
                       (T*)nullptr, x, (T*)nullptr, iters == 1 ? 1 : 0);
    cnt.launches++;
    cnt.bytes += sizeof(T) * 3.0 * nu;
  }

  // one ILU application as the inner PCG runs it, for kernel timing: z = M^-1 r
  // with the update step (upd: r, p, q, x, s) in its forward stages
  template <typename TV, typename T>
  void time_ilu_step(T* upd, T* y, T* z) {
    T *r = upd, *p = upd + ldu, *q = upd + 2 * ldu, *x = upd + 3 * ldu, *s = upd + 4 * ldu;
    const PcgFuse<T> pf{S(SC_PCG), S(SC_PCG) + 1, z, s, p, q, x, r, 0};
    FA.apply<TV, T, T>(r, y, z, st, cnt, &pf);
    cnt.bytes += sizeof(T) * 10.0 * nu;  // z, s, p, q, x, r read; p, q, x, r written
  }

  int spmv_dot_grid() const {
    return Ablb.nb > 0 ? Ablb.nb : (int)std::max<int64_t>(1, (nu * Ablk.G4 + BLOCK - 1) / BLOCK);
  }

  // bytes of one A-block SpMV with s_v-byte values and s_x-byte vectors:
  // the blocked layout (values + indices, column lists, row ranges, p read
  // and q written) or the CSR4 one
  double ablk_spmv_bytes(double sv, double sx, bool rdot = false) const {
    const double r = rdot ? sx * nu : 0.0;  // the second dot's other vector
    if (Ablb.nb > 0)
      return (double)Ablb.nnz * (sv + Ablb.ib) + 4.0 * (double)Ablb.n_cols + 4.0 * nu + 2.0 * nu * sx + r;
    return Ablk.nnz * (sv + 4.0) + 8.0 * (nu + 1) + sx * 2.0 * nu + r;
  }

  // q = A p and the per-workgroup partials of p.q, on the padded A block;
  // TV selects the stored values (Ablk.v4_32 or Ablk.v4_64), T the vectors
  template <typename TV, typename T>
  // q = A p with per-workgroup partials of p.q in part and, with rdot, of
  // rdot.p in part2 (spmv_dot_grid() of each)
  void launch_spmv_dot(const T* pv, T* q, double* part, const T* rdot = nullptr, double* part2 = nullptr) {
    if (Ablb.nb > 0) {  // by lattice blocks: 128 threads for up to 64 rows, 512 above
      const bool d = std::is_same<TV, double>::value;
      const TV* v = d ? (const TV*)Ablb.v64.p : (const TV*)Ablb.v32.p;
      const int ccap = (Ablb.cmax + 3) & ~3;
      const size_t lds = sizeof(T) * ((size_t)ccap + Ablb.cnmax);
#define FS_SB(IB, NT)                                                                                          \
  {                                                                                                            \
    auto fn = &k_spmv_blk_dot<TV, T, IB, NT>;                                                                  \
    if (lds > 64 * 1024)                                                                                       \
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),    \
                "hipFuncSetAttribute");                                                                        \
    hipLaunchKernelGGL(fn, dim3(Ablb.nb), dim3(NT), lds, st, Ablb.desc.as<i32x4>(), Ablb.cinfo.as<int2>(),     \
                       Ablb.rng.as<int32_t>(), Ablb.idx.p, v, Ablb.cols.as<int32_t>(), pv, q, part, ccap, rdot, \
                       part2);                                                                                 \
  }
      if (Ablb.rmax <= 64) {
        if (Ablb.ib == 1) FS_SB(1, 128) else FS_SB(2, 128)
      } else {
        if (Ablb.ib == 1) FS_SB(1, 512) else FS_SB(2, 512)
      }
#undef FS_SB
      return;
    }
    const int g = spmv_dot_grid();
    using V4 = typename Vec4<TV>::type;
    const V4* av = std::is_same<TV, double>::value ? (const V4*)Ablk.v4_64.p : (const V4*)Ablk.v4_32.p;
#define FS_SD(GG)                                                                                            \
  case GG:                                                                                                   \
    hipLaunchKernelGGL((k_spmv4_dot<TV, T, GG>), dim3(g), dim3(BLOCK), 0, st, nu, Ablk.rp4.as<int64_t>(),     \
                       Ablk.ci4.as<i32x4>(), av, pv, q, part);                                              \
    break;
    switch (Ablk.G4) {
      FS_SD(4) FS_SD(8) FS_SD(16)
      default: FS_SD(32)
    }
#undef FS_SD
    if (rdot) {
      hipLaunchKernelGGL((k_dot_partial_t<T>), dim3(g), dim3(BLOCK), 0, st, nu, rdot, pv, part2);
      cnt.launches++;
    }
  }

  // TV: stored matrix / factor values; T: vectors and arithmetic
  template <typename TV, typename T>
  void precond(const double* src, T* dst) {
    T* y = (T*)ytmp.p;
    const int ia = o.method == FS_METHOD_GMRES_REF ? 0 : o.inner_a_iters;
    if (ia > 0) inner_pcg_A<TV, T>(src, dst, ia);
    else FA.apply<TV, T, double>(src, y, dst, st, cnt);
    spmv4_launch<TV, T, T, double, T, 2>(K, np, nu, dst, (T*)tp.p, src + nu, nu, st);
    cnt.launches++;
    inner_cg<T>((const T*)tp.p, dst + nu);
  }

  // CGS2 orthogonalisation of V[j+1] against V[0..j], Givens update
  void arnoldi_tail(int j) {
    double* Vb = dptr<double>(V);
    double* w = Vb + (int64_t)(j + 1) * ld;
    const int m = j + 1;
    const int g = std::min<int>(RED_BLOCKS, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
    double* part = dptr<double>(partial);
    const dim3 gd(g, (m + MDOT_CHUNK - 1) / MDOT_CHUNK);
    hipLaunchKernelGGL((k_mdot_partial<double>), gd, dim3(BLOCK), 0, st, n, m, Vb, ld, w, part);
    hipLaunchKernelGGL(k_mdot_final, dim3(m), dim3(BLOCK), 0, st, g, m, part, S(SC_H1), 0);
    hipLaunchKernelGGL((k_gs_update<double>), dim3(g), dim3(BLOCK), 0, st, n, m, Vb, ld, S(SC_H1), w,
                       (double*)nullptr);
    hipLaunchKernelGGL((k_mdot_partial<double>), gd, dim3(BLOCK), 0, st, n, m, Vb, ld, w, part);
    hipLaunchKernelGGL(k_mdot_final, dim3(m), dim3(BLOCK), 0, st, g, m, part, S(SC_H2), 0);
    hipLaunchKernelGGL((k_gs_update<double>), dim3(g), dim3(BLOCK), 0, st, n, m, Vb, ld, S(SC_H2), w, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, st, g, part, S(SC_NRM), 1);
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(1), 0, st, j, MAXM, S(SC_H1), S(SC_H2), S(SC_NRM), S(SC_HCOL),
                       S(SC_H), S(SC_CS), S(SC_SN), S(SC_GAM), S(SC_EST));
    hipLaunchKernelGGL((k_scale_by_inv<double>), dim3(g), dim3(BLOCK), 0, st, n, w, w, S(SC_NRM), 0);
    cnt.launches += 9;
    cnt.bytes += 8.0 * n * (2.0 * (m + 1) + 2.0 * (m + 2) + 2.0);
  }

  // The flexible method's Arnoldi step with delayed reorthogonalisation
  // (DCGS2): two passes over the basis instead of classical Gram-Schmidt's
  // four (two projections, two updates). Column j holds u_j (one projection
  // pass, scaled by its norm); one pass forms V^T u_j and V^T w, the next
  // writes the final v_j and the projected w; the previous Hessenberg column
  // is corrected and its Givens rotation committed, column j's made
  // provisional with h_{j+1,j} = |w| for the residual estimate.
  void arnoldi_dcgs2(int j) {
    double* Vb = dptr<double>(V);
    double* u = Vb + (int64_t)j * ld;
    double* w = Vb + (int64_t)(j + 1) * ld;
    const int m = j + 1;
    const int g = std::min<int>(RED_BLOCKS, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
    double* part = dptr<double>(partial);
    double* part2 = part + (int64_t)RED_BLOCKS * MAXM;
    // row tiles of 8 rows per thread, 2 basis vectors per round (16 loads in
    // flight; 4 x 4 measured the same; 16-vector chunks re-reading x1 and x2
    // per chunk took 0.29 ms more per step at C5)
    hipLaunchKernelGGL((k_mdot2_rows<8, 2>), dim3(g), dim3(BLOCK), 0, st, n, m, (const double*)Vb, ld,
                       (const double*)u, (const double*)w, part, part2);
    hipLaunchKernelGGL(k_mdot2_final, dim3(2 * m), dim3(BLOCK), 0, st, g, m, (const double*)part,
                       (const double*)part2, S(SC_H1), S(SC_H2));
    hipLaunchKernelGGL(k_dcgs2_coef, dim3(1), dim3(BLOCK), 0, st, j, (const double*)S(SC_H1), (const double*)S(SC_H2),
                       (const double*)S(SC_NRM), S(SC_RAW0), S(SC_RAW1), S(SC_DC));
    hipLaunchKernelGGL(k_dcgs2_update, dim3(g), dim3(BLOCK), 0, st, n, j, Vb, ld, (const double*)S(SC_H1),
                       (const double*)S(SC_H2), (const double*)S(SC_DC), w, part);
    hipLaunchKernelGGL(k_givens_lag, dim3(1), dim3(BLOCK), 0, st, j, MAXM, g, (const double*)part,
                       (const double*)S(SC_RAW0), (const double*)S(SC_RAW1), S(SC_H), S(SC_CS), S(SC_SN),
                       S(SC_GAM), S(SC_DC) + 2, S(SC_NRM), S(SC_EST));
    hipLaunchKernelGGL((k_scale_by_inv<double>), dim3(g), dim3(BLOCK), 0, st, n, w, w, S(SC_NRM), 0);
    cnt.launches += 6;
    cnt.bytes += 8.0 * n * ((m + 2.0) + (j + 4.0) + 2.0);
  }

  void start_cycle(double* v0) {
    const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL((k_scale_by_inv<double>), dim3(g), dim3(BLOCK), 0, st, n, v0, v0, S(SC_NRM), 0);
    hipLaunchKernelGGL(k_init_gamma, dim3(1), dim3(1), 0, st, S(SC_GAM), S(SC_NRM), 0, S(SC_EST));
    cnt.launches += 2;
  }

  template <typename TZ>
  void finish_cycle(int dim, const TZ* Zb) {
    const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(1), 0, st, dim, MAXM, S(SC_H), S(SC_GAM), S(SC_Y));
    hipLaunchKernelGGL((k_update_x<TZ>), dim3(g), dim3(BLOCK), 0, st, n, dim, Zb, ld, S(SC_Y),
                       dptr<double>(x));
    cnt.launches += 2;
    cnt.bytes += (double)n * (16.0 + dim * sizeof(TZ));
  }

  // Krylov basis size: fs_opts.restart, or when 0 the method's own (deal.II
  // SolverGMRES max_n_tmp_vectors 25 -> 23 for the reference method,
  // stokes_ilu.cc:538; 60 for the flexible method)
  int restart_of(int method) const {
    return o.restart > 0 ? o.restart : (method == FS_METHOD_GMRES_REF ? 23 : 60);
  }

  // right-preconditioned flexible GMRES, FP64 true residual at each restart;
  // the preconditioner streams TV matrix / factor values and computes in T
  template <typename TV, typename T>
  void fgmres(double tol_abs, int max_steps, fs_stats& s) {
    const int m = restart_of(FS_METHOD_FGMRES);
    double* Vb = dptr<double>(V);
    T* Zb = (T*)Z.p;
    // FS_FGMRES_CGS2=1: classical Gram-Schmidt twice per step instead of the
    // delayed reorthogonalisation (for comparison)
    const char* ge = std::getenv("FS_FGMRES_CGS2");
    const bool cgs2 = ge && std::atoi(ge) == 1;
    int acc = 0, cycles = 0;
    bool conv = false;
    double est = 0;
    for (;;) {
      kop(dptr<double>(x), Vb, dptr<double>(b));  // V0 = b - K x
      norm(Vb, n, SC_NRM);
      const double beta = read_scalar(SC_NRM);
      est = beta;
      if (o.verbose) std::fprintf(stderr, "FGMRES restart %d: true residual %.6e (tol %.3e) after %d steps\n", cycles,
                                  beta, tol_abs, acc);
      if (cycles == 0) s.initial_rel_residual = beta;
      if (!(beta > tol_abs)) { conv = std::isfinite(beta); break; }
      if (acc >= max_steps) break;
      start_cycle(Vb);
      int dim = 0;
      for (int j = 0; j < m && acc < max_steps; ++j) {
        ++acc;
        T* zj = Zb + (int64_t)j * ld;
        prec<TV, T>(Vb + (int64_t)j * ld, zj);
        double* w = Vb + (int64_t)(j + 1) * ld;
        // the Arnoldi relation holds for K in FP64: only the (flexible)
        // preconditioner is approximate, so the restarts' FP64 true residual
        // and the minimised residual belong to the same operator
        kop_t<double, T>(zj, w);
        if (cgs2) arnoldi_tail(j);
        else arnoldi_dcgs2(j);
        dim = j + 1;
        est = read_scalar(SC_EST);
        if (o.verbose > 1) std::fprintf(stderr, "  FGMRES step %d: estimate %.6e\n", acc, est);
        if (!(est > tol_abs) || !std::isfinite(est)) break;
      }
      finish_cycle<T>(dim, Zb);
      ++cycles;
      if (!std::isfinite(est)) break;
    }
    s.outer_iters = acc;
    s.restarts = cycles;
    s.last_value = est;
    s.status = conv ? FS_OK : FS_NOT_CONVERGED;
  }

  // deal.II SolverGMRES, left preconditioning (stokes_ilu.cc:536-542)
  void gmres_ref(double bnorm, int max_steps, fs_stats& s) {
    const int m = restart_of(FS_METHOD_GMRES_REF);
    double* Vb = dptr<double>(V);
    double* p = dptr<double>(r);
    Ctrl c;
    c.tol = o.rtol * bnorm;
    c.reduce = o.reduce;
    c.max_steps = max_steps;
    int acc = 0, cycles = 0;
    Ctrl::State stt = Ctrl::ITERATE;
    do {
      kop(dptr<double>(x), p, dptr<double>(b));
      prec<double, double>(p, Vb);
      norm(Vb, n, SC_NRM);
      const double rho = read_scalar(SC_NRM);
      stt = c.check(acc, rho);
      if (stt != Ctrl::ITERATE) break;
      start_cycle(Vb);
      int dim = 0;
      for (int j = 0; j < m && stt == Ctrl::ITERATE; ++j) {
        ++acc;
        kop(Vb + (int64_t)j * ld, p, nullptr);
        prec<double, double>(p, Vb + (int64_t)(j + 1) * ld);
        arnoldi_tail(j);
        dim = j + 1;
        stt = c.check(acc, read_scalar(SC_EST));
      }
      finish_cycle<double>(dim, Vb);
      ++cycles;
    } while (stt == Ctrl::ITERATE);
    s.outer_iters = c.last_step;
    s.restarts = cycles;
    s.last_value = c.last_value;
    s.status = stt == Ctrl::SUCCESS ? FS_OK : FS_NOT_CONVERGED;
  }

  void numeric_factorize() {
    int* ferr = dptr<int>(flags, 0);
    hip_check(hipMemsetAsync(ferr, 0, 4, st), "memset");
    FA.factorize(K.rp.as<int64_t>(), K.ci.as<int32_t>(), K.v64.as<double>(), nu, st, ferr, cnt);
    FS.factorize(Mp.rp.as<int64_t>(), Mp.ci.as<int32_t>(), Mp.v64.as<double>(), np, st, ferr, cnt);
    int e = 0;
    hip_check(hipMemcpyAsync(&e, ferr, 4, hipMemcpyDeviceToHost, st), "D2H");
    sync();
    if (e) throw FsError(FS_ERR_ZERO_PIVOT, "ILU: zero or non-finite pivot");
    dirty = false;
  }

  // the value maps to the device; the host copies are released
  void maps_to_device() {
    auto up = [](DevBuf& d, std::vector<int64_t>& m, int64_t& cnt) {
      cnt = (int64_t)m.size();
      d.upload(m.data(), 8 * std::max<size_t>(1, m.size()));
      std::vector<int64_t>().swap(m);
    };
    up(dmapA, mapA, nmapA);
    up(dmapBt, mapBt, nmapBt);
    up(dmapB, mapB, nmapB);
    up(dmapK11, mapK11, nmapK11);
    up(dmapMp, mapMp, nmapMp);
    up(dablk, ablk_from_k, nablk);
  }

  // New values in host memory (the caller's layouts): copied to the device
  // as they are (one staging buffer, released afterwards), then scattered
  // there, so no host copy of the solver's layouts is kept.
  void set_values(const double* A, const double* Bt, const double* B, const double* K11, const double* M) {
    const int64_t tot = nmapA + nmapBt + nmapB + nmapK11 + nmapMp;
    DevBuf stage;
    stage.alloc(8 * std::max<int64_t>(1, tot));
    double* d = stage.as<double>();
    const double* src[5] = {A, Bt, B, K11, M};
    const int64_t cnt[5] = {nmapA, nmapBt, nmapB, nmapK11, nmapMp};
    double* dst[5];
    int64_t off = 0;
    for (int i = 0; i < 5; ++i) {
      dst[i] = d + off;
      if (cnt[i] > 0)
        hip_check(hipMemcpyAsync(dst[i], src[i], 8 * cnt[i], hipMemcpyHostToDevice, st), "H2D values");
      off += cnt[i];
    }
    set_values_device(dst[0], dst[1], dst[2], dst[3], dst[4]);  // synchronises the stream
  }

  // New values already in device memory (the caller's layouts): scattered
  // into K and M_p by device copies of the value maps, A alone gathered from
  // K.
  void set_values_device(const double* A, const double* Bt, const double* B, const double* K11, const double* M) {
    auto scatter = [&](const DevBuf& map, int64_t n, const double* src, DevBuf& dst) {
      if (n == 0) return;
      hipLaunchKernelGGL(k_values_scatter, dim3(grid(n)), dim3(BLOCK), 0, st, n, map.as<int64_t>(), src,
                         dst.as<double>());
    };
    if (!K.v64.p) K.v64.alloc(8 * std::max<int64_t>(1, K.nnz));
    if (!Mp.v64.p) Mp.v64.alloc(8 * std::max<int64_t>(1, Mp.nnz));
    if (!Ablk.v64.p) Ablk.v64.alloc(8 * std::max<int64_t>(1, Ablk.nnz));
    scatter(dmapA, nmapA, A, K.v64);
    scatter(dmapBt, nmapBt, Bt, K.v64);
    scatter(dmapB, nmapB, B, K.v64);
    scatter(dmapK11, nmapK11, K11, K.v64);
    scatter(dmapMp, nmapMp, M, Mp.v64);
    K.derive_from_v64(st);
    const int64_t na = nablk;
    if (na > 0)
      hipLaunchKernelGGL(k_values_gather, dim3(grid(na)), dim3(BLOCK), 0, st, na, dablk.as<int64_t>(),
                         (const double*)K.v64.as<double>(), Ablk.v64.as<double>());
    Ablk.derive_from_v64(st);
    Ablb.upload_values_device(Ablk.v64.as<double>(), st);
    Kblb.upload_values_device(K.v64.as<double>(), st);  // K's first rows, in K's entry order
    Mp.derive_from_v64(st);
    hip_check(hipGetLastError(), "device refactor");
    hip_check(hipStreamSynchronize(st), "sync");
    cnt.launches += 12;
    dirty = true;
  }

  // network guess into x (solver order); falls back to zero when its
  // residual is not below ||b||
  void learned_guess(double bnorm, fs_stats& s) {
    hipEvent_t a, e;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&e), "event");
    hip_check(hipEventRecord(a, st), "event record");
    const float* Y = fs::guess_infer(model, guess_nc, g_logeta.as<float>(), st);
    GuessNorm gn;
    for (int c = 0; c < 4; ++c) { gn.mean[c] = model->mean[c]; gn.stdv[c] = model->stdv[c]; }
    gn.P = model->P;
    const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(k_guess_scatter, dim3(g), dim3(BLOCK), 0, st, n, g_idx.as<int32_t>(), Y, gn,
                       dptr<double>(x));
    const int64_t nd = (int64_t)diag_only.size();
    if (nd > 0)
      hipLaunchKernelGGL(k_diag_rows, dim3(std::min<int64_t>(1024, (nd + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, nd,
                         g_diag_rows.as<int32_t>(), K.rp.as<int64_t>(), K.v64.as<double>(), dptr<double>(b),
                         dptr<double>(x));
    kmult64(dptr<double>(x), dptr<double>(r), dptr<double>(b));
    norm(dptr<double>(r), n, SC_TMP + 3);
    const double res = read_scalar(SC_TMP + 3);
    s.guess_rel_residual = res / bnorm;
    if (res < bnorm) {
      s.guess_used = FS_GUESS_LEARNED;
    } else {
      hip_check(hipMemsetAsync(x.p, 0, 8 * n, st), "memset");
      s.guess_used = FS_GUESS_ZERO;
    }
    hip_check(hipEventRecord(e, st), "event record");
    hip_check(hipEventSynchronize(e), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, e), "elapsed");
    s.t_guess_ms = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
    cnt.launches += 3;
  }

  // ---- diagonal scaling wrappers --------------------------------------------
  double* dsc_t() { return dptr<double>(dsc_buf, ld); }
  void* dsc_tT() { return (void*)dptr<double>(dsc_buf, 2 * ld); }
  void compute_scaling() {
    if (!dsc_buf.p) dsc_buf.alloc(8 * (size_t)(3 * ld));
    double* D = dptr<double>(dsc_buf);
    const int g = std::min<int>(RED_BLOCKS / 2, grid(nu));
    double* part = dptr<double>(partial);
    hipLaunchKernelGGL(k_scale_coeffs_u, dim3(g), dim3(BLOCK), 0, st, nu, K.rp.as<int64_t>(), K.ci.as<int32_t>(),
                       K.v64.as<double>(), D, part);
    hipLaunchKernelGGL(k_scale_coeffs_p, dim3(1), dim3(BLOCK), 0, st, g, part, nu, np, D);
    cnt.launches += 2;
    cnt.bytes += K.nnz * 12.0 * nu / n + 8.0 * n;
    // s_p is finite and positive only if every velocity coefficient is
    // (a zero row makes sum 1/D^2 infinite); the reference asserts nonzero
    double spv = 0;
    hip_check(hipMemcpyAsync(&spv, D + nu, 8, hipMemcpyDeviceToHost, st), "D2H");
    sync();
    if (!(spv > 0.0) || !std::isfinite(spv)) throw FsError(FS_ERR_BAD_INPUT, "diagonal scaling: zero velocity row");
    Dsc = D;
  }
  void scale_vec(double* v, int inv) {
    hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(n)), dim3(BLOCK), 0, st, n, v, Dsc, v, inv);
    cnt.launches++;
  }
  // y = b~ - K~ x~ (bv given) or y = K~ x~
  void kop(const double* xv, double* y, const double* bv) {
    if (!Dsc) { kmult64(xv, y, bv); return; }
    double* t = dsc_t();
    hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(n)), dim3(BLOCK), 0, st, n, xv, Dsc, t, 1);
    kmult64(t, y, nullptr);
    if (bv) hipLaunchKernelGGL(k_resid_scaled, dim3(grid(n)), dim3(BLOCK), 0, st, n, bv, Dsc, y);
    else hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(n)), dim3(BLOCK), 0, st, n, y, Dsc, y, 1);
    cnt.launches += 2;
    cnt.bytes += 40.0 * n;
  }
  // w = K~ z (TV matrix values, z in the preconditioned-vector precision T)
  template <typename TV, typename T>
  void kop_t(const T* z, double* w) {
    const T* x = z;
    if (Dsc) {
      T* t = (T*)dsc_tT();
      hipLaunchKernelGGL((k_mul_diag<T, T>), dim3(grid(n)), dim3(BLOCK), 0, st, n, z, Dsc, t, 1);
      x = t;
      cnt.launches++;
      cnt.bytes += n * (2.0 * sizeof(T) + 8.0);
    }
    if (std::is_same<TV, double>::value && Kblb.nb > 0) {
      // velocity rows by lattice blocks (2-byte column indices, the block's
      // x values gathered once into LDS), pressure rows in CSR4
      launch_kvel(x, w);
      spmv4_launch<TV, T, double, double, double, 0>(K, np, nu, x, w + nu, (const double*)nullptr, 0, st);
      cnt.launches += 2;
      cnt.bytes += (double)Kblb.nnz * (8.0 + Kblb.ib) + 4.0 * (double)Kblb.n_cols + 4.0 * nu + 8.0 * nu +
                   (K.nnz - (double)Kblb.nnz) * (sizeof(TV) + 4.0) + 8.0 * (np + 1) + n * sizeof(T) + 8.0 * np;
    } else {
      spmv4_launch<TV, T, double, double, double, 0>(K, n, 0, x, w, (const double*)nullptr, 0, st);
      cnt.launches++;
      cnt.bytes += K.nnz * (sizeof(TV) + 4.0) + 8.0 * (n + 1) + n * (sizeof(T) + 8.0);
    }
    if (Dsc) {
      hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(n)), dim3(BLOCK), 0, st, n, w, Dsc, w, 1);
      cnt.launches++;
      cnt.bytes += 24.0 * n;
    }
  }
  // w[0, nu) = (rows [0, nu) of K) x in FP64, x of type TX (all n entries)
  template <typename TX>
  void launch_kvel(const TX* x, double* w) {
    const int ccap = (Kblb.cmax + 3) & ~3;
    const size_t lds = sizeof(double) * ((size_t)ccap + Kblb.cnmax);
#define FS_KB(IB, NT)                                                                                          \
  {                                                                                                            \
    auto fn = &k_spmv_blk_dot<double, double, IB, NT, TX>;                                                     \
    if (lds > 64 * 1024)                                                                                       \
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),    \
                "hipFuncSetAttribute");                                                                        \
    hipLaunchKernelGGL(fn, dim3(Kblb.nb), dim3(NT), lds, st, Kblb.desc.as<i32x4>(), Kblb.cinfo.as<int2>(),     \
                       Kblb.rng.as<int32_t>(), Kblb.idx.p, (const double*)Kblb.v64.p, Kblb.cols.as<int32_t>(), \
                       x, w, (double*)nullptr, ccap, (const double*)nullptr, (double*)nullptr);                \
  }
    if (Kblb.rmax <= 64) {
      if (Kblb.ib == 1) FS_KB(1, 128) else FS_KB(2, 128)
    } else {
      if (Kblb.ib == 1) FS_KB(1, 512) else FS_KB(2, 512)
    }
#undef FS_KB
  }
  // dst = P~^-1 src
  template <typename TV, typename T>
  void prec(const double* src, T* dst) {
    if (!Dsc) { precond<TV, T>(src, dst); return; }
    double* t = dsc_t();
    hipLaunchKernelGGL((k_mul_diag<double, double>), dim3(grid(n)), dim3(BLOCK), 0, st, n, src, Dsc, t, 0);
    precond<TV, T>(t, dst);
    hipLaunchKernelGGL((k_mul_diag<T, T>), dim3(grid(n)), dim3(BLOCK), 0, st, n, dst, Dsc, dst, 0);
    cnt.launches += 2;
    cnt.bytes += n * (24.0 + 2.0 * sizeof(T) + 8.0);
  }

  // ---- Schur-complement reduction (FS_METHOD_SCR), FP64 -------------------
  int grid(int64_t len) const { return std::min<int>(4096, (int)std::max<int64_t>(1, (len + BLOCK - 1) / BLOCK)); }

  double dot64(const double* a, const double* b2, int64_t len) {
    const int g = std::min<int>(RED_BLOCKS, grid(len));
    hipLaunchKernelGGL((k_dot_partial<double>), dim3(g), dim3(BLOCK), 0, st, len, a, b2, dptr<double>(partial));
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(BLOCK), 0, st, g, dptr<double>(partial), S(SC_TMP + 3), 0);
    cnt.launches += 2;
    cnt.bytes += 16.0 * len;
    return read_scalar(SC_TMP + 3);
  }
  void axpby(int64_t len, double a, const double* xv, double bb, double* y) {
    hipLaunchKernelGGL(k_axpby, dim3(grid(len)), dim3(BLOCK), 0, st, len, a, xv, bb, y);
    cnt.launches++;
    cnt.bytes += 24.0 * len;
  }
  // y = B^T a (velocity rows of K applied to [0; a]); full: n-vector scratch
  void bt_mult(const double* a, double* y, double* full) {
    hip_check(hipMemsetAsync(full, 0, 8 * nu, st), "memset");
    hip_check(hipMemcpyAsync(full + nu, a, 8 * np, hipMemcpyDeviceToDevice, st), "D2D");
    spmv4_launch<double, double, double, double, double, 0>(K, nu, 0, full, y, (const double*)nullptr, 0, st);
    cnt.launches++;
    cnt.bytes += (double)K.nnz * 12.0 * nu / n;
  }
  // y = B v (pressure rows of K applied to [v; 0])
  void b_mult(const double* v, double* y, double* full) {
    hip_check(hipMemcpyAsync(full, v, 8 * nu, hipMemcpyDeviceToDevice, st), "D2D");
    hip_check(hipMemsetAsync(full + nu, 0, 8 * np, st), "memset");
    spmv4_launch<double, double, double, double, double, 0>(K, np, nu, full, y, (const double*)nullptr, 0, st);
    cnt.launches++;
    cnt.bytes += (double)K.nnz * 12.0 * np / n;
  }

  // InverseMatrixCG::vmult on A (linear_operators.implement.h:64-95): deal.II
  // SolverCG with ILU_A, ReductionControl(tol, reduce), from x (zero when
  // x_zero); written in the residual form r = b - A x, z = P r, p = z + beta p,
  // whose iterates equal deal.II's g = -r, d = -h form. With diagonal scaling
  // the CG on A~ = D^-1 A D^-1 with ILU(A~) = D^-1 ILU(A) D^-1 is run in the
  // unscaled variables (x = D^-1 x~, r = D r~): every inner product it takes
  // is the same in both, and the stopping rule measures ||r~|| = ||D^-1 r||.
  int cg_A(const double* bv, double* xv, bool x_zero, double tol, double reduce, int max_steps, bool& ok) {
    double* base = (double*)pcg.p;
    double *rr = base, *z = base + ldu, *pv = base + 2 * ldu, *q = base + 3 * ldu, *y = base + 4 * ldu;
    double* part_pq = dptr<double>(pcg_part);
    double* part_rz = part_pq + spmv_dot_grid();
    double* rz = S(SC_PCG + 1);
    const int gv = 512, gu = 256;
    if (x_zero) {
      hip_check(hipMemsetAsync(xv, 0, 8 * nu, st), "memset");
      hipLaunchKernelGGL((k_copy_convert<double, double>), dim3(gv), dim3(BLOCK), 0, st, nu, bv, rr);
    } else {
      spmv4_launch<double, double, double, double, double, 1>(Ablk, nu, 0, xv, rr, bv, 0, st);
    }
    cnt.launches++;
    Ctrl c;
    c.tol = tol;
    c.reduce = reduce;
    c.max_steps = max_steps;
    wnorm(rr, nu, Dsc, SC_TMP + 2);
    Ctrl::State stt = c.check(0, read_scalar(SC_TMP + 2));
    int it = 0, cur = 0;
    if (stt == Ctrl::ITERATE) {
      FA.apply<double, double, double>(rr, y, z, st, cnt);
      hipLaunchKernelGGL((k_dot_partial_t<double>), dim3(gv), dim3(BLOCK), 0, st, nu, rr, z, part_rz);
      hipLaunchKernelGGL((k_pcg_direction<double>), dim3(gu), dim3(BLOCK), 0, st, nu, part_rz, gv, rz, rz, z, pv, 1);
      cnt.launches += 2;
    }
    while (stt == Ctrl::ITERATE) {
      ++it;
      launch_spmv_dot<double, double>(pv, q, part_pq);
      hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, st, spmv_dot_grid(), part_pq, S(SC_PCG), 0);
      hipLaunchKernelGGL((k_pcg_update<double>), dim3(gu), dim3(BLOCK), 0, st, nu, S(SC_PCG), rz + cur, pv, q, xv, rr, 0);
      cnt.launches += 3;
      cnt.bytes += ablk_spmv_bytes(8.0, 8.0) + 8.0 * 5.0 * nu;
      wnorm(rr, nu, Dsc, SC_TMP + 2);
      const double res = read_scalar(SC_TMP + 2);
      stt = c.check(it, res);
      if (stt != Ctrl::ITERATE) break;
      FA.apply<double, double, double>(rr, y, z, st, cnt);
      hipLaunchKernelGGL((k_dot_partial_t<double>), dim3(gv), dim3(BLOCK), 0, st, nu, rr, z, part_rz);
      hipLaunchKernelGGL((k_pcg_direction<double>), dim3(gu), dim3(BLOCK), 0, st, nu, part_rz, gv, rz + cur,
                         rz + (cur ^ 1), z, pv, 0);
      cnt.launches += 2;
      cur ^= 1;
    }
    ok = stt == Ctrl::SUCCESS;
    return it;
  }

  // StokesLinearSystem::solve_cg (stokes_ilu.cc:564-710): schur_rhs = B A^{-1}
  // f - g; CG on S = B A^{-1} B^T preconditioned by InverseMatrixCG(M_p, ILU_S)
  // to rtol ||schur_rhs|| (ReductionControl with the solver's reduce); then
  // A u = f - B^T p from the current u. A^{-1} is CG with ILU_A to 1e-15.
  // With diagonal scaling (pre_solve_scaling, stokes_ilu.cc:585-586) the
  // method runs in the unscaled variables, as cg_A does: S~ = D_p^-1 S D_p^-1
  // and M~^-1 = D_p M^-1 D_p leave every inner product of the CG unchanged,
  // and its residual and right-hand-side norms are taken of D_p^-1 g.
  void scr(fs_stats& s) {
    const int64_t ldn = (n + 63) & ~int64_t(63), ldv = (nu + 63) & ~int64_t(63), ldp = (np + 63) & ~int64_t(63);
    if (!scr_ws.p) scr_ws.alloc(8 * (size_t)(ldn + 3 * ldv + 5 * ldp));
    double* full = dptr<double>(scr_ws);
    double *tmp = full + ldn, *t2 = tmp + ldv, *t3 = t2 + ldv;
    double *srhs = t3 + ldv, *g = srhs + ldp, *hh = g + ldp, *d = hh + ldp, *sv = d + ldp;
    double* xu = dptr<double>(x);
    double* xp = xu + nu;
    const double* f = dptr<double>(b);
    const double* gp = f + nu;
    int64_t a_steps = 0;
    bool ok_all = true;
    auto ainv = [&](const double* rhs_v, double* out, bool from_zero) {
      bool ok = true;
      a_steps += cg_A(rhs_v, out, from_zero, 1e-15, 1e-15, (int)nu, ok);
      ok_all = ok_all && ok;
    };
    auto schur = [&](const double* v, double* out) {
      bt_mult(v, t2, full);
      ainv(t2, t3, true);
      b_mult(t3, out, full);
    };
    // schur_rhs = B A^{-1} f - g
    ainv(f, tmp, true);
    b_mult(tmp, srhs, full);
    axpby(np, -1.0, gp, 1.0, srhs);
    const double* Dp = Dsc ? Dsc + nu : nullptr;
    auto pnorm = [&](const double* v) {
      wnorm(v, np, Dp, SC_TMP + 3);
      return read_scalar(SC_TMP + 3);
    };
    Ctrl c;
    c.tol = o.rtol * pnorm(srhs);
    c.reduce = o.reduce;
    c.max_steps = o.max_outer > 0 ? o.max_outer : (int)np;
    // deal.II SolverCG on S from x_p
    const double xpn = dot64(xp, xp, np);
    if (xpn != 0.0) {
      schur(xp, g);
      axpby(np, -1.0, srhs, 1.0, g);
    } else {
      axpby(np, -1.0, srhs, 0.0, g);
    }
    double res = pnorm(g);
    Ctrl::State stt = c.check(0, res);
    int it = 0;
    double gh = 0;
    if (stt == Ctrl::ITERATE) {
      inner_cg<double>(g, hh);
      axpby(np, -1.0, hh, 0.0, d);
      gh = dot64(g, hh, np);
    }
    while (stt == Ctrl::ITERATE) {
      ++it;
      schur(d, hh);
      const double alpha = gh / dot64(d, hh, np);
      if (!std::isfinite(alpha)) { stt = Ctrl::FAILURE; break; }
      axpby(np, alpha, d, 1.0, xp);
      axpby(np, alpha, hh, 1.0, g);
      res = pnorm(g);
      stt = c.check(it, res);
      if (stt != Ctrl::ITERATE) break;
      inner_cg<double>(g, hh);
      const double ghn = dot64(g, hh, np);
      const double beta = ghn / gh;
      gh = ghn;
      axpby(np, -1.0, hh, beta, d);
    }
    (void)sv;
    // velocity: A u = f - B^T p, from the current u
    bt_mult(xp, tmp, full);
    axpby(nu, 1.0, f, -1.0, tmp);
    ainv(tmp, xu, false);
    s.outer_iters = it;
    s.restarts = 0;
    s.last_value = c.last_value;
    s.inner_iters = a_steps;  // M_p CG steps are added from the device counter
    s.status = (stt == Ctrl::SUCCESS && ok_all) ? FS_OK : FS_NOT_CONVERGED;
  }

  // rhs, x: device FP64 in caller order
  int solve_device(const double* d_rhs, double* d_x, fs_stats& s);
};

// ---------------------------------------------------------------------------
namespace {

// Build K in solver order from the caller's blocks; maps[e] = K position of
// caller entry e of each block.
void build_K(fs_handle& h, const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11) {
  const auto& pa = h.FA.sched.perm;
  const auto& ipa = h.FA.sched.iperm;
  const auto& ps = h.FS.sched.perm;
  const auto& ips = h.FS.sched.iperm;
  const int64_t nu = h.nu, np = h.np, n = h.n;
  HostCsr Kh;
  Kh.n_rows = Kh.n_cols = n;
  Kh.rp.assign(n + 1, 0);
  auto rowlen = [](const fs_csr* m, int64_t r) { return m->row_ptr[r + 1] - m->row_ptr[r]; };
  for (int64_t r = 0; r < nu; ++r) Kh.rp[r + 1] = rowlen(A, pa[r]) + rowlen(Bt, pa[r]);
  for (int64_t q = 0; q < np; ++q) Kh.rp[nu + q + 1] = rowlen(B, ps[q]) + rowlen(K11, ps[q]);
  for (int64_t r = 0; r < n; ++r) Kh.rp[r + 1] += Kh.rp[r];
  Kh.ci.resize(Kh.rp[n]);
  h.mapA.assign(A->row_ptr[nu], -1);
  h.mapBt.assign(Bt->row_ptr[nu], -1);
  h.mapB.assign(B->row_ptr[np], -1);
  h.mapK11.assign(K11->row_ptr[np], -1);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int64_t>> tmp;  // (new col, tagged caller entry)
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < n; ++r) {
      tmp.clear();
      const bool vel = r < nu;
      const fs_csr* m1 = vel ? A : B;
      const fs_csr* m2 = vel ? Bt : K11;
      const int64_t o = vel ? pa[r] : ps[r - nu];
      for (int64_t p = m1->row_ptr[o]; p < m1->row_ptr[o + 1]; ++p) tmp.emplace_back(ipa[m1->col_idx[p]], p << 1);
      for (int64_t p = m2->row_ptr[o]; p < m2->row_ptr[o + 1]; ++p)
        tmp.emplace_back((int32_t)(nu + ips[m2->col_idx[p]]), (p << 1) | 1);
      std::sort(tmp.begin(), tmp.end());
      int64_t q = Kh.rp[r];
      for (auto& e : tmp) {
        Kh.ci[q] = e.first;
        const int64_t src = e.second >> 1;
        const bool second = e.second & 1;
        if (vel) (second ? h.mapBt : h.mapA)[src] = q;
        else (second ? h.mapK11 : h.mapB)[src] = q;
        ++q;
      }
    }
  }
  fs::phase_mark("K in solver order: rows and value maps");
  h.K.upload_pattern(Kh, true);
  fs::phase_mark("K: device pattern + padded copy");
  h.diag_only.clear();
  for (int64_t r = 0; r < n; ++r)
    if (Kh.rp[r + 1] - Kh.rp[r] == 1 && Kh.ci[Kh.rp[r]] == r) h.diag_only.push_back((int32_t)r);
  // A alone (velocity rows, velocity columns) in solver order
  HostCsr Ah;
  Ah.n_rows = Ah.n_cols = nu;
  Ah.rp.assign(nu + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nu; ++r) {
    int64_t c = 0;
    for (int64_t q = Kh.rp[r]; q < Kh.rp[r + 1]; ++q) c += Kh.ci[q] < nu;
    Ah.rp[r + 1] = c;
  }
  for (int64_t r = 0; r < nu; ++r) Ah.rp[r + 1] += Ah.rp[r];
  Ah.ci.resize(Ah.rp[nu]);
  h.ablk_from_k.resize(Ah.rp[nu]);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nu; ++r) {
    int64_t o = Ah.rp[r];
    for (int64_t q = Kh.rp[r]; q < Kh.rp[r + 1]; ++q)
      if (Kh.ci[q] < nu) { Ah.ci[o] = Kh.ci[q]; h.ablk_from_k[o] = q; ++o; }
  }
  fs::phase_mark("A alone in solver order");
  h.Ablk.upload_pattern(Ah, true);
  fs::phase_mark("A: device pattern + padded copy");
  h.Ablb.build(Ah, h.FA.sched.blk_ptr);
  fs::phase_mark("A by lattice blocks");
  {
    // K's velocity rows (all columns), for K z in FP64 by lattice blocks
    HostCsr Kv;
    Kv.n_rows = nu;
    Kv.n_cols = n;
    Kv.rp.assign(Kh.rp.begin(), Kh.rp.begin() + nu + 1);
    Kv.ci.assign(Kh.ci.begin(), Kh.ci.begin() + Kh.rp[nu]);
    h.Kblb.build(Kv, h.FA.sched.blk_ptr, false);
  }
  fs::phase_mark("K velocity rows by lattice blocks");
}

void fill_stats_err(fs_stats* s, int code) {
  if (s) {
    std::memset(s, 0, sizeof(*s));
    s->status = code;
  }
}

#define FS_GUARD_BEGIN try {
#define FS_GUARD_END                                          \
  }                                                           \
  catch (const FsError& e) {                                  \
    fs::set_error("%s", e.what());                            \
    return e.code;                                            \
  }                                                           \
  catch (const std::bad_alloc&) {                             \
    fs::set_error("host allocation failed");                  \
    return FS_ERR_OUT_OF_MEMORY;                              \
  }                                                           \
  catch (const std::exception& e) {                           \
    fs::set_error("%s", e.what());                            \
    return FS_ERR_BAD_INPUT;                                  \
  }

}  // namespace


int fs_handle::solve_device(const double* d_rhs, double* d_x, fs_stats& s) {
  std::memset(&s, 0, sizeof(s));
  cnt = Counters{};
  if (o.precision != FS_PREC_FP64 && (o.method == FS_METHOD_GMRES_REF || o.method == FS_METHOD_SCR))
    throw FsError(FS_ERR_UNSUPPORTED, "reference-mode GMRES and SCR run in FP64 only");
  hipEvent_t e0, e1, e2;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventCreate(&e2), "event");
  struct EvGuard {
    hipEvent_t a, b, c;
    ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); }
  } eg{e0, e1, e2};
  hip_check(hipEventRecord(e0, st), "event record");
  // setup: numeric factorisation (the reference refactorises in every solve,
  // BlockSchurPreconditioner::reinitialize, stokes_ilu.cc:27-37)
  numeric_factorize();
  hip_check(hipEventRecord(e1, st), "event record");
  const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
  const int32_t* pm = perm_all.as<int32_t>();
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, st, n, pm, d_rhs, dptr<double>(b), 0);
  if (o.guess == FS_GUESS_LEARNED && !model)
    throw FsError(FS_ERR_BAD_INPUT, "FS_GUESS_LEARNED needs fs_stokes_set_guess_field first");
  if (o.guess == FS_GUESS_USER)
    hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, st, n, pm, d_x, dptr<double>(x), 0);
  else
    hip_check(hipMemsetAsync(x.p, 0, 8 * n, st), "memset");
  hip_check(hipMemsetAsync(flags.p, 0, 16, st), "memset");  // flags[1] CG iterations, [2] CG failure
  cnt.launches += 2;
  norm(dptr<double>(b), n, SC_TMP);
  norm(dptr<double>(x), n, SC_TMP + 1);
  hip_check(hipMemcpyAsync(h_pin, S(SC_TMP), 16, hipMemcpyDeviceToHost, st), "D2H");
  sync();
  const double bnorm = h_pin[0], xnorm = h_pin[1];
  if (!std::isfinite(bnorm) || !std::isfinite(xnorm))
    throw FsError(FS_ERR_BAD_INPUT, "rhs or initial guess contains NaN/Inf");
  s.guess_used = o.guess == FS_GUESS_USER ? FS_GUESS_USER : FS_GUESS_ZERO;
  if (o.guess == FS_GUESS_LEARNED && bnorm > 0.0) learned_guess(bnorm, s);
  const int max_steps = o.max_outer > 0 ? o.max_outer : (int)n;
  // pre_solve_scaling (stokes_ilu.cc:418-431): x *= D, rhs /= D; the
  // tolerance is taken on the scaled rhs (stokes_ilu.cc:533)
  double bnorm_k = bnorm;
  Dsc = nullptr;
  // (SCR keeps the unscaled vectors and weights its norms instead)
  const bool scale_vectors = o.method != FS_METHOD_SCR;
  if (o.apply_diagonal_scaling && bnorm != 0.0) {
    compute_scaling();
    if (scale_vectors) {
      scale_vec(dptr<double>(b), 1);
      scale_vec(dptr<double>(x), 0);
      norm(dptr<double>(b), n, SC_TMP);
      bnorm_k = read_scalar(SC_TMP);
    }
  }
  if (bnorm == 0.0) {
    hip_check(hipMemsetAsync(x.p, 0, 8 * n, st), "memset");
    s.status = FS_OK;
  } else if (o.method == FS_METHOD_GMRES_REF) {
    gmres_ref(bnorm_k, max_steps, s);
  } else if (o.method == FS_METHOD_FGMRES) {
    if (o.precision == FS_PREC_MIXED) fgmres<float, float>(o.rtol * bnorm_k, max_steps, s);
    else if (o.precision == FS_PREC_MIXED_FP64_VECTORS) fgmres<float, double>(o.rtol * bnorm_k, max_steps, s);
    else fgmres<double, double>(o.rtol * bnorm_k, max_steps, s);
    if (s.initial_rel_residual > 0) s.initial_rel_residual /= bnorm_k;
  } else if (o.method == FS_METHOD_SCR) {
    scr(s);
  } else {
    throw FsError(FS_ERR_UNSUPPORTED, "solution method not implemented");
  }
  if (Dsc) {  // post_solve_scaling: x /= D; the rhs back to the caller's values
    if (scale_vectors) {
      scale_vec(dptr<double>(x), 1);
      scale_vec(dptr<double>(b), 0);
    }
    Dsc = nullptr;
  }
  hip_check(hipEventRecord(e2, st), "event record");
  // true residual, back to the caller's order
  kmult64(dptr<double>(x), dptr<double>(r), dptr<double>(b));
  norm(dptr<double>(r), n, SC_TMP + 2);
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, st, n, pm, dptr<double>(x), d_x, 1);
  int fl[4] = {0, 0, 0, 0};
  hip_check(hipMemcpyAsync(fl, flags.p, 16, hipMemcpyDeviceToHost, st), "D2H");
  hip_check(hipMemcpyAsync(h_pin + 2, S(SC_TMP + 2), 8, hipMemcpyDeviceToHost, st), "D2H");
  GridBarrier gbh;
  hip_check(hipMemcpyAsync(&gbh, gbar.p, sizeof(gbh), hipMemcpyDeviceToHost, st), "D2H");
  sync();
  if (gbh.err) {
    hip_check(hipMemset(gbar.p, 0, sizeof(GridBarrier)), "memset");
    throw FsError(FS_ERR_TIMEOUT, "inner CG grid barrier timed out");
  }
  float ms_setup = 0, ms_kry = 0;
  hip_check(hipEventElapsedTime(&ms_setup, e0, e1), "elapsed");
  hip_check(hipEventElapsedTime(&ms_kry, e1, e2), "elapsed");
  s.inner_iters += fl[1];
  s.true_rel_residual = h_pin[2] / (bnorm > 0 ? bnorm : 1.0);
  s.t_setup_ms = ms_setup;
  s.t_krylov_ms = ms_kry;
  s.t_total_ms = ms_setup + ms_kry;
  s.bytes_algorithmic = cnt.bytes;
  s.n_kernel_launches = cnt.launches;
  if (fl[2] && s.status == FS_OK && o.method == FS_METHOD_GMRES_REF) s.status = FS_NOT_CONVERGED;
  return s.status;
}

// ---------------------------------------------------------------------------
static void create_handle(fs_handle& h, const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11,
                          const fs_csr* Mp, const fs_opts* opts) {
  ensure_device();
  if (opts) h.o = *opts;
  else fs_opts_default(&h.o);
  auto& o = h.o;
  if (o.restart < 0 || o.restart > MAXM - 1) throw FsError(FS_ERR_BAD_INPUT, "restart must be in [0, 255]");
  if (o.precision < FS_PREC_FP64 || o.precision > FS_PREC_MIXED_FP64_VECTORS)
    throw FsError(FS_ERR_BAD_INPUT, "unknown precision");
  if (o.ilu_k_A < 0 || o.ilu_k_S < 0 || o.ilu_k_A > 20 || o.ilu_k_S > 20)
    throw FsError(FS_ERR_BAD_INPUT, "level of fill out of range");
  if (!(o.rtol > 0)) throw FsError(FS_ERR_BAD_INPUT, "rtol must be positive");
  // shapes
  HostCsr hA = csr_from_abi(A, "A", true), hBt = csr_from_abi(Bt, "Bt", true), hB = csr_from_abi(B, "B", true),
          hK11 = csr_from_abi(K11, "K11", true), hM = csr_from_abi(Mp, "Mp", true);
  h.nu = hA.n_rows;
  h.np = hM.n_rows;
  h.n = h.nu + h.np;
  if (hA.n_cols != h.nu || hBt.n_rows != h.nu || hBt.n_cols != h.np || hB.n_rows != h.np || hB.n_cols != h.nu ||
      hK11.n_rows != h.np || hK11.n_cols != h.np || hM.n_cols != h.np)
    throw FsError(FS_ERR_BAD_INPUT, "block shapes do not match");
  if (h.n >= INT32_MAX) throw FsError(FS_ERR_BAD_INPUT, "system too large");
  // phase times of the set-up (printed with o.verbose)
  fs::g_verbose = o.verbose;
  fs::phase_reset();
  auto phase = [](const char* what) { fs::phase_mark(what); };
  // orderings
  const int bmax = 256;
  HostPattern pA = pattern_of(hA), pM = pattern_of(hM);
  Schedule sA, sS;
  if (o.ordering == FS_ORDER_REFERENCE) {
    sA = schedule_reference(h.nu, 128);
    sS = schedule_reference(h.np, 128);
  } else {
    const int br = o.block_rows > 0 ? o.block_rows : 64;
    sA = schedule_block_color(pA, o.ilu_k_A, br, o.block_hint, bmax);
    sS = merge_stage_blocks(schedule_block_color(pM, o.ilu_k_S, 1, nullptr, bmax), 64);
  }
  // permuted patterns for the symbolic phase
  phase("orderings (block colouring)");
  HostCsr Anew = permute_square(hA, sA.perm, sA.iperm, nullptr);
  std::vector<int64_t> mapM;
  HostCsr Mnew = permute_square(hM, sS.perm, sS.iperm, &mapM);
  phase("permuted patterns");
  h.FA.keep_pattern = false;
  h.FA.build(pattern_of(Anew), sA, o.ilu_k_A, 8);
  phase("ILU(k) of A: symbolic + device layout");
  h.FS.dense = false;  // M_p's factor is applied inside the fused inner CG only
  h.FS.build(pattern_of(Mnew), sS, o.ilu_k_S, 8);
  phase("ILU of M_p");
  // K in solver order + value maps from the caller's layout
  build_K(h, A, Bt, B, K11);
  h.Mp.upload_pattern(Mnew);
  // mapM is indexed by the sorted-layout entries of hM; convert to caller entries
  {
    std::vector<int64_t> c2s = caller_to_sorted(Mp);
    h.mapMp.resize(c2s.size());
    for (size_t e = 0; e < c2s.size(); ++e) h.mapMp[e] = mapM[c2s[e]];
  }
  h.maps_to_device();
  // device vectors
  h.ld = (h.n + 63) & ~int64_t(63);
  // (the method can be switched later: room for either basis)
  const int m = std::max(h.restart_of(FS_METHOD_GMRES_REF), h.restart_of(FS_METHOD_FGMRES));
  h.m_alloc = m;
  h.b.alloc(8 * h.ld);
  h.x.alloc(8 * h.ld);
  h.r.alloc(8 * h.ld);
  h.V.alloc(8 * h.ld * (m + 1));
  h.Z.alloc(8 * h.ld * m);  // FP64 or FP32 preconditioned vectors (the precision can be switched)
  h.ytmp.alloc(8 * h.ld);
  h.tp.alloc(8 * h.ld);
  h.cg.alloc(8 * 4 * ((h.np + 63) & ~int64_t(63)));
  h.vin.alloc(8 * h.ld);
  h.ldu = (h.nu + 63) & ~int64_t(63);
  h.pcg.alloc(8 * 6 * h.ldu);
  // z.s (or p.q) partials, then r.z partials (as many, or 512)
  h.pcg_part.alloc(8 * (size_t)(2 * h.spmv_dot_grid() + 1024));
  h.partial.alloc(8 * RED_BLOCKS * MAXM * 2);  // two sets: the two-vector multi-dot
  h.sc.alloc(8 * SC_TOTAL);
  h.gbar.alloc(sizeof(GridBarrier));
  h.flags.alloc(64);
  hip_check(hipMemset(h.gbar.p, 0, sizeof(GridBarrier)), "memset");
  hip_check(hipMemset(h.sc.p, 0, 8 * SC_TOTAL), "memset");
  hip_check(hipMemset(h.V.p, 0, 8 * h.ld * (m + 1)), "memset");
  {
    std::vector<int32_t> pa(h.n);
    for (int64_t i = 0; i < h.nu; ++i) pa[i] = h.FA.sched.perm[i];
    for (int64_t q = 0; q < h.np; ++q) pa[h.nu + q] = (int32_t)(h.nu + h.FS.sched.perm[q]);
    h.perm_all.upload(pa.data(), 4 * h.n);
  }
  hip_check(hipStreamCreateWithFlags(&h.st, hipStreamNonBlocking), "stream");
  hip_check(hipHostMalloc((void**)&h.h_pin, 64, hipHostMallocDefault), "hipHostMalloc");
  h.set_values(A->val, Bt->val, B->val, K11->val, Mp->val);
  phase("device vectors, values");
}

extern "C" {

int fs_stokes_create(const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11, const fs_csr* Mp,
                     const fs_opts* opts, fs_handle** out) {
  if (!out) { fs::set_error("fs_stokes_create: out is null"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  FS_GUARD_BEGIN
  std::unique_ptr<fs_handle> h(new fs_handle());
  create_handle(*h, A, Bt, B, K11, Mp, opts);
  *out = h.release();
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_refactor(fs_handle* h, const double* A_val, const double* Bt_val, const double* B_val,
                       const double* K11_val, const double* Mp_val) {
  if (!h || !A_val || !Bt_val || !B_val || !K11_val || !Mp_val) {
    fs::set_error("fs_stokes_refactor: null argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  h->set_values(A_val, Bt_val, B_val, K11_val, Mp_val);
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_refactor_device(fs_handle* h, const double* d_A, const double* d_Bt, const double* d_B,
                              const double* d_K11, const double* d_Mp) {
  if (!h || !d_A || !d_Bt || !d_B || !d_K11 || !d_Mp) {
    fs::set_error("fs_stokes_refactor_device: null argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  hip_check(hipStreamSynchronize(nullptr), "sync");
  h->set_values_device(d_A, d_Bt, d_B, d_K11, d_Mp);
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_solve_device(fs_handle* h, const double* d_rhs, double* d_x, fs_stats* stats) {
  fs_stats tmp;
  if (!stats) stats = &tmp;
  if (!h || !d_rhs || !d_x) {
    fs::set_error("fs_stokes_solve_device: null argument");
    fill_stats_err(stats, FS_ERR_BAD_INPUT);
    return FS_ERR_BAD_INPUT;
  }
  try {
    return h->solve_device(d_rhs, d_x, *stats);
  } catch (const FsError& e) {
    fs::set_error("%s", e.what());
    fill_stats_err(stats, e.code);
    return e.code;
  } catch (const std::exception& e) {
    fs::set_error("%s", e.what());
    fill_stats_err(stats, FS_ERR_BAD_INPUT);
    return FS_ERR_BAD_INPUT;
  }
}

int fs_stokes_solve(fs_handle* h, const double* rhs, double* x, fs_stats* stats) {
  fs_stats tmp;
  if (!stats) stats = &tmp;
  if (!h || !rhs || !x) {
    fs::set_error("fs_stokes_solve: null argument");
    fill_stats_err(stats, FS_ERR_BAD_INPUT);
    return FS_ERR_BAD_INPUT;
  }
  try {
    const int64_t n = h->n;
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(rhs[i]) || (h->o.guess == FS_GUESS_USER && !std::isfinite(x[i])))
        throw FsError(FS_ERR_BAD_INPUT, "rhs or initial guess contains NaN/Inf");
    DevBuf drhs, dx;
    drhs.upload(rhs, 8 * n);
    dx.upload(x, 8 * n);
    int rc = h->solve_device(drhs.as<double>(), dx.as<double>(), *stats);
    hip_check(hipMemcpy(x, dx.p, 8 * n, hipMemcpyDeviceToHost), "D2H");
    return rc;
  } catch (const FsError& e) {
    fs::set_error("%s", e.what());
    fill_stats_err(stats, e.code);
    return e.code;
  } catch (const std::exception& e) {
    fs::set_error("%s", e.what());
    fill_stats_err(stats, FS_ERR_BAD_INPUT);
    return FS_ERR_BAD_INPUT;
  }
}

int fs_stokes_vmult(fs_handle* h, const double* src, double* dst, int32_t precision) {
  if (!h || !src || !dst) { fs::set_error("fs_stokes_vmult: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  const int64_t n = h->n;
  const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
  h->vin.upload(src, 8 * n);
  const int32_t* pm = h->perm_all.as<int32_t>();
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, pm, h->vin.as<double>(),
                     h->b.as<double>(), 0);
  if (precision != FS_PREC_FP64) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(BLOCK), 0, h->st, n, h->b.as<double>(),
                       (float*)h->ytmp.p);
    spmv4_launch<float, float, double, double, float, 0>(h->K, n, 0, (const float*)h->ytmp.p, h->r.as<double>(),
                                                         (const double*)nullptr, 0, h->st);
  } else {
    h->kmult64(h->b.as<double>(), h->r.as<double>(), nullptr);
  }
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, pm, h->r.as<double>(),
                     h->vin.as<double>(), 1);
  hip_check(hipMemcpyAsync(dst, h->vin.p, 8 * n, hipMemcpyDeviceToHost, h->st), "D2H");
  h->sync();
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_precond_vmult(fs_handle* h, const double* src, double* dst, int32_t precision, fs_stats* stats) {
  if (!h || !src || !dst) { fs::set_error("fs_stokes_precond_vmult: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  const int64_t n = h->n;
  const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
  if (h->dirty) h->numeric_factorize();
  h->cnt = Counters{};
  hip_check(hipMemsetAsync(h->flags.p, 0, 16, h->st), "memset");
  h->vin.upload(src, 8 * n);
  const int32_t* pm = h->perm_all.as<int32_t>();
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, pm, h->vin.as<double>(),
                     h->b.as<double>(), 0);
  if (precision == FS_PREC_MIXED) {
    if (h->o.precision != FS_PREC_MIXED) throw FsError(FS_ERR_BAD_INPUT, "FP32 vectors need a FS_PREC_MIXED handle");
    h->precond<float, float>(h->b.as<double>(), (float*)h->Z.p);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(BLOCK), 0, h->st, n, (const float*)h->Z.p,
                       h->r.as<double>());
  } else if (precision == FS_PREC_MIXED_FP64_VECTORS) {
    h->precond<float, double>(h->b.as<double>(), h->r.as<double>());
  } else {
    h->precond<double, double>(h->b.as<double>(), h->r.as<double>());
  }
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, pm, h->r.as<double>(),
                     h->vin.as<double>(), 1);
  hip_check(hipMemcpyAsync(dst, h->vin.p, 8 * n, hipMemcpyDeviceToHost, h->st), "D2H");
  int fl[4];
  hip_check(hipMemcpyAsync(fl, h->flags.p, 16, hipMemcpyDeviceToHost, h->st), "D2H");
  GridBarrier gbh;
  hip_check(hipMemcpyAsync(&gbh, h->gbar.p, sizeof(gbh), hipMemcpyDeviceToHost, h->st), "D2H");
  h->sync();
  if (gbh.err) {
    hip_check(hipMemset(h->gbar.p, 0, sizeof(GridBarrier)), "memset");
    throw FsError(FS_ERR_TIMEOUT, "inner CG grid barrier timed out");
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->inner_iters = fl[1];
    stats->status = fl[2] ? FS_NOT_CONVERGED : FS_OK;
    stats->n_kernel_launches = h->cnt.launches;
    stats->bytes_algorithmic = h->cnt.bytes;
  }
  return fl[2] ? FS_NOT_CONVERGED : FS_OK;
  FS_GUARD_END
}

int fs_stokes_get_ordering(fs_handle* h, int32_t which, int32_t* perm, int32_t* n_stages) {
  if (!h) { fs::set_error("fs_stokes_get_ordering: null handle"); return FS_ERR_BAD_INPUT; }
  const Schedule& s = which == 0 ? h->FA.sched : h->FS.sched;
  if (perm) std::copy(s.perm.begin(), s.perm.end(), perm);
  if (n_stages) *n_stages = s.n_stages();
  return FS_OK;
}

int fs_stokes_get_schedule(fs_handle* h, int32_t which, int32_t* blk_ptr, int32_t* n_blocks, int32_t* stage_ptr,
                           int32_t* n_stages) {
  if (!h || (which != 0 && which != 1)) {
    fs::set_error("fs_stokes_get_schedule: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  const Schedule& s = which == 0 ? h->FA.sched : h->FS.sched;
  if (blk_ptr) std::copy(s.blk_ptr.begin(), s.blk_ptr.end(), blk_ptr);
  if (stage_ptr) std::copy(s.stage_ptr.begin(), s.stage_ptr.end(), stage_ptr);
  if (n_blocks) *n_blocks = s.n_blocks();
  if (n_stages) *n_stages = s.n_stages();
  return FS_OK;
}

int fs_stokes_get_factor_info(fs_handle* h, int32_t which, int64_t* n, int64_t* nnz, int64_t* ikj_updates,
                              int32_t* n_stages) {
  if (!h || (which != 0 && which != 1)) {
    fs::set_error("fs_stokes_get_factor_info: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  const IluDev& F = which == 0 ? h->FA : h->FS;
  if (n) *n = F.n;
  if (nnz) *nnz = F.nnz;
  if (ikj_updates) *ikj_updates = F.ikj_updates;
  if (n_stages) *n_stages = F.sched.n_stages();
  return FS_OK;
}

int fs_stokes_set_guess_field(fs_handle* h, struct fs_guess_model* model, const float* log_eta, int32_t n_cells,
                              const int64_t* vnode, const int8_t* vcomp, const int64_t* pnode) {
  if (!h || !model || !log_eta || !vnode || !vcomp || !pnode || n_cells < 1) {
    fs::set_error("fs_stokes_set_guess_field: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  const int d = model->cfg.dim;
  int64_t P = 1, Pq = 1;
  for (int a = 0; a < d; ++a) { P *= 2 * (int64_t)n_cells + 1; Pq *= (int64_t)n_cells + 1; }
  if ((d + 1) * P >= INT32_MAX) throw FsError(FS_ERR_BAD_INPUT, "lattice too large");
  for (int64_t i = 0; i < Pq; ++i)
    if (!std::isfinite(log_eta[i])) throw FsError(FS_ERR_BAD_INPUT, "log10(eta) field has NaN/Inf");
  std::vector<int32_t> idx(h->n);
  for (int64_t r = 0; r < h->nu; ++r) {
    const int64_t c = h->FA.sched.perm[r];
    const int64_t lat = vnode[c];
    const int comp = vcomp[c];
    if (lat < 0 || lat >= P || comp < 0 || comp >= d) throw FsError(FS_ERR_BAD_INPUT, "velocity dof map out of range");
    idx[r] = (int32_t)(comp * P + lat);
  }
  for (int64_t q = 0; q < h->np; ++q) {
    const int64_t c = h->FS.sched.perm[q];
    int64_t t = pnode[c];
    if (t < 0 || t >= Pq) throw FsError(FS_ERR_BAD_INPUT, "pressure dof map out of range");
    int64_t lat = 0, stride = 1;
    for (int a = 0; a < d; ++a) {
      lat += 2 * (t % (n_cells + 1)) * stride;
      t /= (n_cells + 1);
      stride *= 2 * (int64_t)n_cells + 1;
    }
    idx[h->nu + q] = (int32_t)(d * P + lat);
  }
  h->g_idx.upload(idx.data(), 4 * idx.size());
  h->g_logeta.upload(log_eta, 4 * (size_t)Pq);
  if (!h->diag_only.empty()) h->g_diag_rows.upload(h->diag_only.data(), 4 * h->diag_only.size());
  h->model = model;
  h->guess_nc = n_cells;
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_set_dof_lattice(fs_handle* h, int32_t dim, int32_t n_cells, const int64_t* vnode,
                              const int8_t* vcomp) {
  if (!h || (dim != 2 && dim != 3) || n_cells < 1 || !vnode || !vcomp) {
    fs::set_error("fs_stokes_set_dof_lattice: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  int64_t P = 1;
  for (int a = 0; a < dim; ++a) P *= 2 * (int64_t)n_cells + 1;
  if (dim * P >= INT32_MAX) throw FsError(FS_ERR_BAD_INPUT, "lattice too large");
  std::vector<int32_t> idx(h->nu);
  std::vector<uint8_t> seen((size_t)(dim * P), 0);
  for (int64_t i = 0; i < h->nu; ++i) {
    if (vnode[i] < 0 || vnode[i] >= P || vcomp[i] < 0 || vcomp[i] >= dim)
      throw FsError(FS_ERR_BAD_INPUT, "velocity dof map out of range");
    idx[i] = (int32_t)(vcomp[i] * P + vnode[i]);
    if (seen[idx[i]]++) throw FsError(FS_ERR_BAD_INPUT, "two velocity dofs on one lattice node and component");
  }
  if (h->nu != dim * P) throw FsError(FS_ERR_BAD_INPUT, "the velocity dofs do not cover the node lattice");
  h->v_lat.upload(idx.data(), 4 * idx.size());
  h->lat_dim = dim;
  h->lat_nc = n_cells;
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_velocity_lattice(fs_handle* h, const double* d_x, double* d_u) {
  if (!h || !d_x || !d_u) { fs::set_error("fs_stokes_velocity_lattice: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  if (!h->lat_dim) throw FsError(FS_ERR_BAD_INPUT, "fs_stokes_set_dof_lattice first");
  hipLaunchKernelGGL(k_velocity_lattice, dim3(h->grid(h->nu)), dim3(BLOCK), 0, h->st, h->nu, h->v_lat.as<int32_t>(),
                     d_x, d_u);
  hip_check(hipGetLastError(), "velocity lattice");
  h->sync();
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_velocity_at_points(fs_handle* h, const double* d_u, int64_t n_points, const double* d_points,
                                 double* d_vel) {
  if (!h || !d_u || n_points < 0 || (n_points > 0 && (!d_points || !d_vel))) {
    fs::set_error("fs_stokes_velocity_at_points: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FS_GUARD_BEGIN
  if (!h->lat_dim) throw FsError(FS_ERR_BAD_INPUT, "fs_stokes_set_dof_lattice first");
  if (n_points == 0) return FS_OK;
  const int64_t P = h->nu / h->lat_dim;
  if (h->lat_dim == 2)
    hipLaunchKernelGGL(k_velocity_at_points<2>, dim3(h->grid(n_points)), dim3(BLOCK), 0, h->st, n_points, d_points,
                       h->lat_nc, d_u, P, d_vel);
  else
    hipLaunchKernelGGL(k_velocity_at_points<3>, dim3(h->grid(n_points)), dim3(BLOCK), 0, h->st, n_points, d_points,
                       h->lat_nc, d_u, P, d_vel);
  hip_check(hipGetLastError(), "velocity at points");
  h->sync();
  return FS_OK;
  FS_GUARD_END
}

int fs_stokes_set_method(fs_handle* h, int32_t method) {
  if (!h || method < FS_METHOD_GMRES_REF || method > FS_METHOD_SCR) {
    fs::set_error("fs_stokes_set_method: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  h->o.method = method;
  return FS_OK;
}

int fs_stokes_set_precision(fs_handle* h, int32_t precision) {
  if (!h || precision < FS_PREC_FP64 || precision > FS_PREC_MIXED_FP64_VECTORS) {
    fs::set_error("fs_stokes_set_precision: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  h->o.precision = precision;
  return FS_OK;
}

int fs_stokes_set_guess_mode(fs_handle* h, int32_t guess) {
  if (!h || guess < FS_GUESS_USER || guess > FS_GUESS_LEARNED) {
    fs::set_error("fs_stokes_set_guess_mode: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  h->o.guess = guess;
  return FS_OK;
}

int fs_stokes_set_krylov(fs_handle* h, int32_t restart, int32_t inner_a_iters, double inner_cg_tol,
                         int32_t inner_cg_max, double inner_cg_tol_mixed, int32_t inner_cg_max_mixed) {
  if (!h) { fs::set_error("fs_stokes_set_krylov: null handle"); return FS_ERR_BAD_INPUT; }
  fs_opts o = h->o;
  if (restart >= 0) o.restart = restart;
  if (inner_a_iters >= 0) o.inner_a_iters = inner_a_iters;
  if (inner_cg_tol >= 0) o.inner_cg_tol = inner_cg_tol;
  if (inner_cg_max >= 0) o.inner_cg_max = inner_cg_max;
  if (inner_cg_tol_mixed >= 0) o.inner_cg_tol_mixed = inner_cg_tol_mixed;
  if (inner_cg_max_mixed >= 0) o.inner_cg_max_mixed = inner_cg_max_mixed;
  const int m_ref = o.restart > 0 ? o.restart : 23, m_fg = o.restart > 0 ? o.restart : 60;
  if (std::max(m_ref, m_fg) > h->m_alloc) {
    fs::set_error("fs_stokes_set_krylov: restart exceeds the Krylov basis allocated at fs_stokes_create");
    return FS_ERR_BAD_INPUT;
  }
  if (!(o.inner_cg_tol_mixed > 0) || !(o.inner_cg_tol > 0)) {
    fs::set_error("fs_stokes_set_krylov: inner CG tolerances must be positive");
    return FS_ERR_BAD_INPUT;
  }
  h->o = o;
  return FS_OK;
}

int fs_stokes_time_kernel(fs_handle* h, int32_t kind, int32_t precision, int32_t reps, double* ms_per_launch,
                          double* bytes_per_launch) {
  if (!h || reps < 1) { fs::set_error("fs_stokes_time_kernel: bad argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  if (h->dirty) h->numeric_factorize();
  // FS_PREC_MIXED: FP32 values and vectors; FS_PREC_MIXED_FP64_VECTORS: FP32
  // values, FP64 vectors
  const bool mixed = precision == FS_PREC_MIXED_FP64_VECTORS;
  const bool f32 = precision == FS_PREC_MIXED;
  const int64_t nu = h->nu;
  // deterministic nonzero inputs in the work vectors
  const int g = std::min<int>(4096, (int)std::max<int64_t>(1, (h->n + BLOCK - 1) / BLOCK));
  std::vector<double> ones(h->n, 1.0);
  h->vin.upload(ones.data(), 8 * h->n);
  hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(BLOCK), 0, h->st, h->n, h->vin.as<double>(),
                     (float*)h->ytmp.p);
  void* base = h->pcg.p;
  // kind 2: the update step's vectors r, p, q, x, s; r = the input vector,
  // the rest and alpha = beta = 0, so every repetition sees the same data
  DevBuf upd;
  if (kind == 2) {
    const size_t es = f32 ? 4 : 8;
    upd.alloc(es * 5 * (size_t)h->ldu);
    hip_check(hipMemsetAsync(upd.p, 0, es * 5 * (size_t)h->ldu, h->st), "memset");
    hip_check(hipMemsetAsync(h->S(SC_PCG), 0, 16, h->st), "memset");
    hip_check(hipMemcpyAsync(upd.p, f32 ? h->ytmp.p : h->vin.p, es * nu, hipMemcpyDeviceToDevice, h->st), "D2D");
  }
  auto once = [&]() {
    h->cnt = Counters{};
    switch (kind) {
      case 0: {
        // as in the inner PCG: q = A p with the p.q and r.p partials
        double* pp = dptr<double>(h->pcg_part);
        double* pp2 = pp + h->spmv_dot_grid();
        if (f32) {
          const float* pv = (const float*)h->ytmp.p;
          h->launch_spmv_dot<float, float>(pv, (float*)base, pp, pv, pp2);
        } else if (mixed) {
          h->launch_spmv_dot<float, double>(h->vin.as<double>(), (double*)base, pp, h->vin.as<double>(), pp2);
        } else {
          h->launch_spmv_dot<double, double>(h->vin.as<double>(), (double*)base, pp, h->vin.as<double>(), pp2);
        }
        const double sv = (mixed || f32) ? 4.0 : 8.0, sx = f32 ? 4.0 : 8.0;
        h->cnt.bytes = h->ablk_spmv_bytes(sv, sx, true);
        break;
      }
      case 1:  // the Arnoldi product w = K z with FP64 values
        if (f32) h->kop_t<double, float>((const float*)h->ytmp.p, h->r.as<double>());
        else h->kop_t<double, double>(h->vin.as<double>(), h->r.as<double>());
        break;
      case 2:
        if (f32) h->time_ilu_step<float, float>((float*)upd.p, (float*)base, (float*)base + h->ldu);
        else if (mixed) h->time_ilu_step<float, double>((double*)upd.p, (double*)base, (double*)base + h->ldu);
        else h->time_ilu_step<double, double>((double*)upd.p, (double*)base, (double*)base + h->ldu);
        break;
      case 3:
        if (f32) h->inner_cg<float>((const float*)h->ytmp.p + nu, (float*)base);
        else h->inner_cg<double>(h->vin.as<double>() + nu, (double*)base);
        break;
      case 4:  // the numeric ILU(k) of A from K's values (FP64; each repetition the same factor)
        h->FA.factorize(h->K.rp.as<int64_t>(), h->K.ci.as<int32_t>(), h->K.v64.as<double>(), nu, h->st,
                        dptr<int>(h->flags, 0), h->cnt);
        break;
      default:
        throw FsError(FS_ERR_BAD_INPUT, "unknown kernel class");
    }
  };
  once();  // warm-up
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventRecord(e0, h->st), "event");
  for (int i = 0; i < reps; ++i) once();
  hip_check(hipEventRecord(e1, h->st), "event");
  hip_check(hipEventSynchronize(e1), "event sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_launch) *ms_per_launch = ms / reps;
  if (bytes_per_launch) *bytes_per_launch = h->cnt.bytes;
  return FS_OK;
  FS_GUARD_END
}

void fs_stokes_destroy(fs_handle* h) { delete h; }

}  // extern "C"

extern "C" {

void fs_opts_default(fs_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->method = FS_METHOD_FGMRES;
  o->precision = FS_PREC_FP64;
  o->ordering = FS_ORDER_BLOCK_COLOR;
  o->ilu_k_A = 1;
  o->ilu_k_S = 0;
  o->restart = 0;  // the method's own: 23 (reference GMRES), 60 (FGMRES)
  o->max_outer = 0;
  o->rtol = 1e-8;
  o->reduce = 1e-6;
  o->inner_cg_max = 0;
  o->inner_cg_tol = 1e-12;
  o->inner_cg_tol_mixed = 1e-2;
  o->inner_cg_max_mixed = 50;
  o->apply_diagonal_scaling = 0;
  o->guess = FS_GUESS_USER;
  o->block_rows = 0;
  o->block_hint = nullptr;
  o->verbose = 0;
  o->inner_a_iters = 60;
}

const char* fs_last_error(const fs_handle*) { return fs::get_error(); }

}  // extern "C"
