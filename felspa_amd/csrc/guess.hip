// Learned warm start for the Stokes solve: a cosine-basis spectral neural
// operator (network described in include/fs_guess.h), trained and run by the
// kernels below. Every GEMM-shaped product goes through k_gemm, a BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) strided-batched GEMM with FP32 accumulation;
// the remaining work is elementwise (bias + GELU, their gradients, the loss,
// Adam) or small deterministic reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fs_guess.h"
#include "fs_stokes.h"
#include "guess_internal.h"
#include "solver_internal.h"

using fs::FsError;
using fs::hip_check;

namespace gk {

constexpr int GT = 256;          // threads per GEMM workgroup (4 waves)
constexpr int TM = 64, TN = 64;  // workgroup tile; each wave owns 32 x 32
constexpr int TK = 32;           // one MFMA K step
constexpr int LDP = TK + 8;      // LDS row pitch in BF16 elements (80 bytes)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ inline uint16_t bf16_bits_host(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t ld_bf16(const float* p, int64_t i) { return bf16_bits(p[i]); }
__device__ __forceinline__ uint16_t ld_bf16(const uint16_t* p, int64_t i) { return p[i]; }

struct GemmArgs {
  int M, N, K, K_last, batch, tiles_m, tiles_n;
  const void* A;
  int64_t sAb, sAm, sAk;
  const void* B;
  int64_t sBb, sBk, sBn;
  float* C;
  int64_t sCb, sCm, sCn;
  float alpha, beta;
};

// C_b = alpha A_b B_b + beta C_b. One workgroup per 64 x 64 output tile of
// one batch entry; operands staged through LDS as BF16 with K contiguous
// (A as [m][k], B as [n][k]), so every MFMA operand is one 16-byte LDS read.
// MFMA 16x16x32 lane layout: operand rows/columns are lane % 16, the lane
// group lane / 16 selects 8 consecutive k (the same k for A and B), and the
// result holds rows 4 (lane / 16) + r, column lane % 16.
template <typename TA, typename TB>
__global__ __launch_bounds__(GT) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM * LDP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN * LDP];
  const int64_t tile = blockIdx.x;
  const int tn = (int)(tile % g.tiles_n);
  const int64_t t2 = tile / g.tiles_n;
  const int tm = (int)(t2 % g.tiles_m);
  const int64_t b = t2 / g.tiles_m;
  const TA* A = (const TA*)g.A + b * g.sAb;
  const TB* B = (const TB*)g.B + b * g.sBb;
  float* C = g.C + b * g.sCb;
  const int K = (b == g.batch - 1) ? g.K_last : g.K;
  const int m0 = tm * TM, n0 = tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bool a_kfast = g.sAk == 1, b_kfast = g.sBk == 1;
  for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
    for (int j = 0; j < (TM * TK) / GT; ++j) {
      const int idx = tid + j * GT;
      const int mm = a_kfast ? idx / TK : idx % TM;
      const int kk = a_kfast ? idx % TK : idx / TM;
      const int gm = m0 + mm, gkk = k0 + kk;
      uint16_t v = 0;
      if (gm < g.M && gkk < K) v = ld_bf16(A, (int64_t)gm * g.sAm + (int64_t)gkk * g.sAk);
      As[mm * LDP + kk] = v;
    }
#pragma unroll
    for (int j = 0; j < (TN * TK) / GT; ++j) {
      const int idx = tid + j * GT;
      const int nn = b_kfast ? idx / TK : idx % TN;
      const int kk = b_kfast ? idx % TK : idx / TN;
      const int gn = n0 + nn, gkk = k0 + kk;
      uint16_t v = 0;
      if (gn < g.N && gkk < K) v = ld_bf16(B, (int64_t)gkk * g.sBk + (int64_t)gn * g.sBn);
      Bs[nn * LDP + kk] = v;
    }
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      __builtin_memcpy(&af[i], &As[(wm * 32 + i * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)], 16);
      __builtin_memcpy(&bfr[i], &Bs[(wn * 32 + i * 16 + (lane & 15)) * LDP + 8 * (lane >> 4)], 16);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < g.M && col < g.N) {
          float* p = C + (int64_t)row * g.sCm + (int64_t)col * g.sCn;
          float v = g.alpha * acc[i][j][r];
          if (g.beta != 0.f) v += g.beta * *p;
          *p = v;
        }
      }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// Z[b][c][p] += bias[c]; H = act(Z) (H may alias Z when act == 0)
__global__ __launch_bounds__(256) void k_bias_act(int64_t n, int64_t P, int C, const float* __restrict__ bias,
                                                  float* Z, float* H, int act) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int c = (int)((e / P) % C);
    const float z = Z[e] + bias[c];
    Z[e] = z;
    H[e] = act ? gelu(z) : z;
  }
}

// dZ = dH * act'(Z), in place on dH
__global__ __launch_bounds__(256) void k_act_bwd(int64_t n, const float* __restrict__ Z, float* dH) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    dH[e] *= gelu_grad(Z[e]);
}

__device__ __forceinline__ double block_sum_d(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) s += sh[i];
  return s;
}

// db[c] = sum_{b,p} dZ[b][c][p]; one workgroup per channel, fixed order
__global__ __launch_bounds__(256) void k_bias_grad(int B, int C, int64_t P, const float* __restrict__ dZ,
                                                   float* __restrict__ db) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  double s = 0;
  for (int b = 0; b < B; ++b) {
    const float* p = dZ + ((int64_t)b * C + c) * P;
    for (int64_t i = threadIdx.x; i < P; i += 256) s += (double)p[i];
  }
  s = block_sum_d(s, sh);
  if (threadIdx.x == 0) db[c] = (float)s;
}

// out[i] = sum_{k < nparts} part[k * len + i]  (fixed order)
__global__ __launch_bounds__(256) void k_sum_parts(int nparts, int64_t len, const float* __restrict__ part,
                                                   float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256) {
    double s = 0;
    for (int k = 0; k < nparts; ++k) s += (double)part[(int64_t)k * len + i];
    out[i] = (float)s;
  }
}

// dY = 2 (Y - T) / n ; partial[block] = sum (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(int64_t n, const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, double* __restrict__ partial) {
  __shared__ double sh[4];
  double s = 0;
  const float sc = 2.f / (float)n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const float d = Y[e] - T[e];
    dY[e] = sc * d;
    s += (double)d * (double)d;
  }
  s = block_sum_d(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_sum_final(int nb, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double sh[4];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  s = block_sum_d(s, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// AdamW on the FP32 master weights; refreshes the BF16 copy
__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ pb,
                                              float lr, float b1, float b2, float eps, float wd, float bc1,
                                              float bc2) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] * (1.f - lr * wd) - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    pb[i] = bf16_bits(pi);
  }
}

__global__ __launch_bounds__(256) void k_to_bf16(int64_t n, const float* __restrict__ src, uint16_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = bf16_bits(src[i]);
}

// network input on the (2N+1)^d lattice from log10(eta) on the (N+1)^d
// lattice: channel 0 = (multilinear interpolation) / 3, channels 1..d the
// node coordinates in [0, 1]
__global__ __launch_bounds__(256) void k_prep_input(int dim, int nc, int batch, const float* __restrict__ q1,
                                                    float* __restrict__ X) {
  const int n1 = 2 * nc + 1, nq = nc + 1;
  int64_t P = 1, Pq = 1;
  for (int a = 0; a < dim; ++a) { P *= n1; Pq *= nq; }
  const int64_t total = (int64_t)batch * P;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t b = e / P, node = e % P;
    int idx[3] = {0, 0, 0};
    int64_t t = node;
    for (int a = 0; a < dim; ++a) { idx[a] = (int)(t % n1); t /= n1; }
    // corners of the Q1 cell (or the node itself when even)
    float s = 0.f;
    int cnt = 0;
    const int nk = 1 << dim;
    for (int k = 0; k < nk; ++k) {
      int64_t qi = 0, stride = 1;
      bool ok = true;
      for (int a = 0; a < dim; ++a) {
        const int i = idx[a];
        int q;
        if (i % 2 == 0) {
          if ((k >> a) & 1) { ok = false; break; }
          q = i / 2;
        } else {
          q = (i - 1) / 2 + ((k >> a) & 1);
        }
        qi += (int64_t)q * stride;
        stride *= nq;
      }
      if (!ok) continue;
      s += q1[b * Pq + qi];
      ++cnt;
    }
    float* xb = X + b * (int64_t)(dim + 1) * P;
    xb[node] = (s / (float)cnt) * (1.f / 3.f);
    for (int a = 0; a < dim; ++a) xb[(int64_t)(a + 1) * P + node] = (float)idx[a] / (float)(n1 - 1);
  }
}

inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// host launcher
void gemm(hipStream_t st, int M, int N, int K, int batch, const void* A, bool a_bf16, int64_t sAb, int64_t sAm,
          int64_t sAk, const void* B, bool b_bf16, int64_t sBb, int64_t sBk, int64_t sBn, float* C, int64_t sCb,
          int64_t sCm, int64_t sCn, float alpha, float beta, int K_last = -1) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.K_last = K_last >= 0 ? K_last : K; g.batch = batch;
  g.tiles_m = (M + TM - 1) / TM;
  g.tiles_n = (N + TN - 1) / TN;
  g.A = A; g.sAb = sAb; g.sAm = sAm; g.sAk = sAk;
  g.B = B; g.sBb = sBb; g.sBk = sBk; g.sBn = sBn;
  g.C = C; g.sCb = sCb; g.sCm = sCm; g.sCn = sCn;
  g.alpha = alpha; g.beta = beta;
  const int64_t tiles = (int64_t)g.tiles_m * g.tiles_n * batch;
  if (tiles > INT32_MAX) throw FsError(FS_ERR_BAD_INPUT, "GEMM grid too large");
  const dim3 grid((unsigned)tiles), blk(GT);
  if (a_bf16 && b_bf16) hipLaunchKernelGGL((k_gemm<uint16_t, uint16_t>), grid, blk, 0, st, g);
  else if (a_bf16) hipLaunchKernelGGL((k_gemm<uint16_t, float>), grid, blk, 0, st, g);
  else if (b_bf16) hipLaunchKernelGGL((k_gemm<float, uint16_t>), grid, blk, 0, st, g);
  else hipLaunchKernelGGL((k_gemm<float, float>), grid, blk, 0, st, g);
}

// out[outer][nout][inner] = T[nout][nin] . in[outer][nin][inner]  (+ beta out)
void axis_apply(hipStream_t st, const uint16_t* T, int nout, int nin, int64_t outer, int64_t inner, const float* in,
                float* out, float beta) {
  if (inner == 1)
    gemm(st, (int)outer, nout, nin, 1, in, false, 0, nin, 1, T, true, 0, 1, nin, out, 0, nout, 1, 1.f, beta);
  else
    gemm(st, nout, (int)inner, nin, (int)outer, T, true, 0, nin, 1, in, false, (int64_t)nin * inner, inner, 1, out,
         (int64_t)nout * inner, inner, 1, 1.f, beta);
}

// pointwise channel mixing: Y[b] = W X[b]  (W [O][C] BF16)
void pw_fwd(hipStream_t st, const uint16_t* W, int O, int C, int B, int64_t P, const float* X, float* Y, float beta) {
  gemm(st, O, (int)P, C, B, W, true, 0, C, 1, X, false, (int64_t)C * P, P, 1, Y, (int64_t)O * P, P, 1, 1.f, beta);
}
// dX[b] (+)= W^T dY[b]
void pw_bwd_in(hipStream_t st, const uint16_t* W, int O, int C, int B, int64_t P, const float* dY, float* dX,
               float beta) {
  gemm(st, C, (int)P, O, B, W, true, 0, 1, C, dY, false, (int64_t)O * P, P, 1, dX, (int64_t)C * P, P, 1, 1.f, beta);
}

constexpr int64_t WCHUNK = 2048;  // split-K chunk of the weight gradients

// dW = sum_b dY[b] X[b]^T  (split over lattice chunks, partials summed in order)
void pw_bwd_w(hipStream_t st, const float* dY, int O, const float* X, int C, int B, int64_t P, float* part, float* dW) {
  const int64_t nch = (P + WCHUNK - 1) / WCHUNK;
  const int64_t last = P - (nch - 1) * WCHUNK;
  for (int b = 0; b < B; ++b)
    gemm(st, O, C, (int)WCHUNK, (int)nch, dY + (int64_t)b * O * P, false, WCHUNK, P, 1, X + (int64_t)b * C * P, false,
         WCHUNK, 1, P, part + (int64_t)b * nch * O * C, (int64_t)O * C, C, 1, 1.f, 0.f, (int)last);
  hipLaunchKernelGGL(k_sum_parts, dim3(grid_for((int64_t)O * C)), dim3(256), 0, st, (int)(B * nch), (int64_t)O * C,
                     part, dW);
}

}  // namespace gk

using namespace gk;

// ---------------------------------------------------------------------------
namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int64_t ipow(int64_t a, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= a;
  return r;
}

void check_cfg(const fs_guess_config& c) {
  if (c.dim != 2 && c.dim != 3) throw FsError(FS_ERR_BAD_INPUT, "guess model: dim must be 2 or 3");
  if (c.width < 1 || c.width > 256 || c.modes < 1 || c.modes > 64 || c.layers < 1 || c.layers > 16 || c.proj < 1 ||
      c.proj > 512)
    throw FsError(FS_ERR_BAD_INPUT, "guess model: width/modes/layers/proj out of range");
}

#define FSG_GUARD_BEGIN try {
#define FSG_GUARD_END                     \
  }                                       \
  catch (const FsError& e) {              \
    fs::set_error("%s", e.what());        \
    return e.code;                        \
  }                                       \
  catch (const std::bad_alloc&) {         \
    fs::set_error("host allocation failed"); \
    return FS_ERR_OUT_OF_MEMORY;          \
  }                                       \
  catch (const std::exception& e) {       \
    fs::set_error("%s", e.what());        \
    return FS_ERR_BAD_INPUT;              \
  }

}  // namespace

void fs_guess_model::layout() {
  const int d = cfg.dim, w = cfg.width, pw = cfg.proj;
  cin = d + 1;
  cout = d + 1;
  const int64_t Mm_ = ipow(cfg.modes, d);
  int64_t o = 0;
  off_Wlift = o; o += (int64_t)w * cin;
  off_blift = o; o += w;
  lay.resize(cfg.layers);
  for (int l = 0; l < cfg.layers; ++l) {
    lay[l].R = o; o += Mm_ * w * w;
    lay[l].W = o; o += (int64_t)w * w;
    lay[l].b = o; o += w;
  }
  off_Wp1 = o; o += (int64_t)pw * w;
  off_bp1 = o; o += pw;
  off_Wp2 = o; o += (int64_t)cout * pw;
  off_bp2 = o; o += cout;
  n_params = o;
}

void fs_guess_model::ensure_workspace(int nc, int batch, bool train) {
  if (nc == ws_nc && batch <= ws_batch && (!train || ws_train)) return;
  if (nc == ws_nc) train = train || ws_train;
  const int d = cfg.dim, w = cfg.width, pw = cfg.proj, L = cfg.layers;
  N1 = 2 * (int64_t)nc + 1;
  P = ipow(N1, d);
  M = std::min<int64_t>(cfg.modes, N1);
  if (M != cfg.modes) throw FsError(FS_ERR_BAD_INPUT, "guess model: lattice smaller than the number of modes");
  Mm = ipow(M, d);
  const int64_t B = nc == ws_nc ? std::max(batch, ws_batch) : batch;
  const int64_t act = B * w * P;  // one [B][w][P] tensor
  // element counts
  std::vector<int64_t> sizes;
  auto add = [&](int64_t n) { sizes.push_back((n + 63) & ~int64_t(63)); };
  for (int l = 0; l <= L; ++l) add(act);           // H
  for (int l = 1; l <= L; ++l) add(act);           // Z
  for (int l = 1; l <= L; ++l) add(B * w * Mm);    // T2
  for (int k = 1; k < d; ++k) add(B * w * ipow(N1, d - k) * ipow(M, k));  // S
  add(B * w * Mm);                                 // O
  add(B * pw * P);                                 // P1
  add(B * pw * P);                                 // G
  add(B * cout * P);                               // Y
  add(B * cin * P);                                // Xin (inference input)
  if (train) {
    add(act); add(act);                            // dHa, dHb
    add(B * pw * P);                               // dG
    add(B * w * Mm); add(B * w * Mm);              // dO, dT2
    for (int k = 1; k < d; ++k) add(B * w * ipow(N1, d - k) * ipow(M, k));  // dS
    add(B * cout * P);                             // dY
  }
  int64_t total = 0;
  for (auto s : sizes) total += s;
  ws.alloc(4 * (size_t)total);
  float* p = (float*)ws.p;
  size_t k = 0;
  auto take = [&]() { float* r = p; p += sizes[k++]; return r; };
  H.assign(L + 1, nullptr);
  Z.assign(L + 1, nullptr);
  T2.assign(L + 1, nullptr);
  S.assign(d, nullptr);
  dS.assign(d, nullptr);
  for (int l = 0; l <= L; ++l) H[l] = take();
  for (int l = 1; l <= L; ++l) Z[l] = take();
  for (int l = 1; l <= L; ++l) T2[l] = take();
  for (int kk = 1; kk < d; ++kk) S[kk] = take();
  O = take(); P1 = take(); G = take(); Y = take();
  Xin = take();
  dHa = dHb = dG = dO = dT2 = dY = nullptr;
  if (train) {
    dHa = take(); dHb = take(); dG = take(); dO = take(); dT2 = take();
    for (int kk = 1; kk < d; ++kk) dS[kk] = take();
    dY = take();
  }
  // weight-gradient partials: max over the pointwise products of O*C
  const int64_t nch = (P + gk::WCHUNK - 1) / gk::WCHUNK;
  const int64_t oc = std::max<int64_t>({(int64_t)w * cin, (int64_t)w * w, (int64_t)pw * w, (int64_t)cout * pw});
  if (train) part.alloc(4 * (size_t)(std::max<int64_t>(B * nch * oc, 4096) + 64));
  // cosine matrices on the (2N+1) lattice, BF16
  std::vector<uint16_t> h(4 * (size_t)M * N1);
  uint16_t *af = h.data(), *aft = af + M * N1, *sb = aft + M * N1, *sbt = sb + M * N1;
  const double hstep = 1.0 / (double)(N1 - 1);
  for (int64_t m = 0; m < M; ++m)
    for (int64_t i = 0; i < N1; ++i) {
      const double x = (double)i * hstep;
      const double c = std::cos(M_PI * (double)m * x);
      const double wq = (i == 0 || i == N1 - 1) ? 0.5 * hstep : hstep;
      const float a = (float)(wq * c), s = (float)((m ? 2.0 : 1.0) * c);
      af[m * N1 + i] = bf16_bits_host(a);
      aft[i * M + m] = bf16_bits_host(a);
      sb[i * M + m] = bf16_bits_host(s);
      sbt[m * N1 + i] = bf16_bits_host(s);
    }
  phi.upload(h.data(), 2 * h.size());
  Af = (const uint16_t*)phi.p;
  AfT = Af + M * N1;
  Sb = AfT + M * N1;
  SbT = Sb + M * N1;
  ws_nc = nc;
  ws_batch = (int)B;
  ws_train = train;
}

// Y = network(X), X [B][cin][P]
void fs_guess_model::forward(const float* X, int B, hipStream_t s) {
  const int d = cfg.dim, w = cfg.width, pw = cfg.proj, L = cfg.layers;
  const uint16_t* pb = (const uint16_t*)pbf.p;
  const float* p = (const float*)p32.p;
  const int64_t act = (int64_t)B * w * P;
  // lift
  pw_fwd(s, pb + off_Wlift, w, cin, B, P, X, H[0], 0.f);
  hipLaunchKernelGGL(k_bias_act, dim3(grid_for(act)), dim3(256), 0, s, act, P, w, p + off_blift, H[0], H[0], 0);
  for (int l = 1; l <= L; ++l) {
    const auto& lo = lay[l - 1];
    pw_fwd(s, pb + lo.W, w, w, B, P, H[l - 1], Z[l], 0.f);
    // analysis along x, y (, z)
    const float* in = H[l - 1];
    for (int a = 0; a < d; ++a) {
      const int64_t outer = (int64_t)B * w * ipow(N1, d - 1 - a), inner = ipow(M, a);
      float* out = (a == d - 1) ? T2[l] : S[a + 1];
      axis_apply(s, Af, (int)M, (int)N1, outer, inner, in, out, 0.f);
      in = out;
    }
    // per-mode channel mixing O[b,o,m] = sum_c T2[b,c,m] R[m,c,o]
    gemm(s, B, w, w, (int)Mm, T2[l], false, 1, (int64_t)w * Mm, Mm, pb + lo.R, true, (int64_t)w * w, w, 1, O, 1,
         (int64_t)w * Mm, Mm, 1.f, 0.f);
    // synthesis along (z,) y, x; the last one accumulates into Z
    in = O;
    for (int a = d - 1; a >= 0; --a) {
      const int64_t outer = (int64_t)B * w * ipow(N1, d - 1 - a), inner = ipow(M, a);
      float* out = (a == 0) ? Z[l] : S[a];
      axis_apply(s, Sb, (int)N1, (int)M, outer, inner, in, out, a == 0 ? 1.f : 0.f);
      in = out;
    }
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for(act)), dim3(256), 0, s, act, P, w, p + lo.b, Z[l], H[l],
                       l < L ? 1 : 0);
  }
  const int64_t ap = (int64_t)B * pw * P, ay = (int64_t)B * cout * P;
  pw_fwd(s, pb + off_Wp1, pw, w, B, P, H[L], P1, 0.f);
  hipLaunchKernelGGL(k_bias_act, dim3(grid_for(ap)), dim3(256), 0, s, ap, P, pw, p + off_bp1, P1, G, 1);
  pw_fwd(s, pb + off_Wp2, cout, pw, B, P, G, Y, 0.f);
  hipLaunchKernelGGL(k_bias_act, dim3(grid_for(ay)), dim3(256), 0, s, ay, P, cout, p + off_bp2, Y, Y, 0);
}

// gradients of the loss whose dL/dY is in dY (forward must have run on X)
void fs_guess_model::backward(const float* X, int B, hipStream_t s) {
  const int d = cfg.dim, w = cfg.width, pw = cfg.proj, L = cfg.layers;
  const uint16_t* pb = (const uint16_t*)pbf.p;
  float* g = (float*)grad.p;
  float* pt = (float*)part.p;
  const int64_t act = (int64_t)B * w * P, ap = (int64_t)B * pw * P;
  // head
  hipLaunchKernelGGL(k_bias_grad, dim3(cout), dim3(256), 0, s, B, cout, P, dY, g + off_bp2);
  pw_bwd_w(s, dY, cout, G, pw, B, P, pt, g + off_Wp2);
  pw_bwd_in(s, pb + off_Wp2, cout, pw, B, P, dY, dG, 0.f);
  hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(ap)), dim3(256), 0, s, ap, P1, dG);
  hipLaunchKernelGGL(k_bias_grad, dim3(pw), dim3(256), 0, s, B, pw, P, dG, g + off_bp1);
  pw_bwd_w(s, dG, pw, H[L], w, B, P, pt, g + off_Wp1);
  float* dH = dHa;
  float* dHp = dHb;
  pw_bwd_in(s, pb + off_Wp1, pw, w, B, P, dG, dH, 0.f);
  for (int l = L; l >= 1; --l) {
    const auto& lo = lay[l - 1];
    if (l < L) hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(act)), dim3(256), 0, s, act, Z[l], dH);
    // dH now holds dZ_l
    hipLaunchKernelGGL(k_bias_grad, dim3(w), dim3(256), 0, s, B, w, P, dH, g + lo.b);
    pw_bwd_w(s, dH, w, H[l - 1], w, B, P, pt, g + lo.W);
    pw_bwd_in(s, pb + lo.W, w, w, B, P, dH, dHp, 0.f);
    // adjoint of the synthesis: x, y (, z)
    const float* in = dH;
    for (int a = 0; a < d; ++a) {
      const int64_t outer = (int64_t)B * w * ipow(N1, d - 1 - a), inner = ipow(M, a);
      float* out = (a == d - 1) ? dO : dS[a + 1];
      axis_apply(s, SbT, (int)M, (int)N1, outer, inner, in, out, 0.f);
      in = out;
    }
    // dR[m] = T2_m^T dO_m ; dT2_m = dO_m R_m^T
    gemm(s, w, w, B, (int)Mm, T2[l], false, 1, Mm, (int64_t)w * Mm, dO, false, 1, (int64_t)w * Mm, Mm, g + lo.R,
         (int64_t)w * w, w, 1, 1.f, 0.f);
    gemm(s, B, w, w, (int)Mm, dO, false, 1, (int64_t)w * Mm, Mm, pb + lo.R, true, (int64_t)w * w, 1, w, dT2, 1,
         (int64_t)w * Mm, Mm, 1.f, 0.f);
    // adjoint of the analysis: (z,) y, x; the last accumulates into dH_{l-1}
    in = dT2;
    for (int a = d - 1; a >= 0; --a) {
      const int64_t outer = (int64_t)B * w * ipow(N1, d - 1 - a), inner = ipow(M, a);
      float* out = (a == 0) ? dHp : dS[a];
      axis_apply(s, AfT, (int)N1, (int)M, outer, inner, in, out, a == 0 ? 1.f : 0.f);
      in = out;
    }
    std::swap(dH, dHp);
  }
  // lift (identity activation)
  hipLaunchKernelGGL(k_bias_grad, dim3(w), dim3(256), 0, s, B, w, P, dH, g + off_blift);
  pw_bwd_w(s, dH, w, X, cin, B, P, pt, g + off_Wlift);
}

namespace fs {
const float* guess_infer(fs_guess_model* m, int n_cells, const float* d_log_eta, hipStream_t s) {
  m->ensure_workspace(n_cells, 1, false);
  float* X = m->Xin;
  hipLaunchKernelGGL(k_prep_input, dim3(grid_for(m->P)), dim3(256), 0, s, m->cfg.dim, n_cells, 1, d_log_eta, X);
  m->forward(X, 1, s);
  return m->Y;
}
}  // namespace fs

// ---------------------------------------------------------------------------
extern "C" {

int fs_guess_model_create(const fs_guess_config* cfg, fs_guess_model** out) {
  if (!cfg || !out) { fs::set_error("fs_guess_model_create: null argument"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  FSG_GUARD_BEGIN
  fs::ensure_device();
  check_cfg(*cfg);
  std::unique_ptr<fs_guess_model> m(new fs_guess_model());
  m->cfg = *cfg;
  m->layout();
  const int w = cfg->width, pw = cfg->proj;
  std::vector<float> h(m->n_params, 0.f);
  uint64_t s = cfg->seed ^ 0x5DEECE66Dull;
  auto fill = [&](int64_t off, int64_t n, double a) {
    for (int64_t i = 0; i < n; ++i) h[off + i] = (float)((2.0 * uniform(s) - 1.0) * a);
  };
  fill(m->off_Wlift, (int64_t)w * m->cin, 1.0 / std::sqrt((double)m->cin));
  const int64_t Mm = ipow(cfg->modes, cfg->dim);
  for (auto& lo : m->lay) {
    fill(lo.R, Mm * w * w, 1.0 / (double)w);
    fill(lo.W, (int64_t)w * w, 1.0 / std::sqrt((double)w));
  }
  fill(m->off_Wp1, (int64_t)pw * w, 1.0 / std::sqrt((double)w));
  fill(m->off_Wp2, (int64_t)m->cout * pw, 1.0 / std::sqrt((double)pw));
  m->p32.alloc(4 * (size_t)m->n_params);
  m->pbf.alloc(2 * (size_t)m->n_params);
  m->grad.alloc(4 * (size_t)m->n_params);
  m->adam_m.alloc(4 * (size_t)m->n_params);
  m->adam_v.alloc(4 * (size_t)m->n_params);
  hip_check(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking), "stream");
  *out = m.release();
  return fs_guess_model_set_params(*out, h.data());
  FSG_GUARD_END
}

void fs_guess_model_destroy(fs_guess_model* m) { delete m; }

int fs_guess_model_get_config(const fs_guess_model* m, fs_guess_config* cfg) {
  if (!m || !cfg) { fs::set_error("fs_guess_model_get_config: null argument"); return FS_ERR_BAD_INPUT; }
  *cfg = m->cfg;
  return FS_OK;
}

int64_t fs_guess_model_num_params(const fs_guess_model* m) { return m ? m->n_params : -1; }

int fs_guess_model_set_params(fs_guess_model* m, const float* in) {
  if (!m || !in) { fs::set_error("fs_guess_model_set_params: null argument"); return FS_ERR_BAD_INPUT; }
  FSG_GUARD_BEGIN
  const int64_t n = m->n_params;
  hip_check(hipMemcpy(m->p32.p, in, 4 * n, hipMemcpyHostToDevice), "H2D");
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(256), 0, m->st, n, (const float*)m->p32.p,
                     (uint16_t*)m->pbf.p);
  hip_check(hipMemsetAsync(m->adam_m.p, 0, 4 * n, m->st), "memset");
  hip_check(hipMemsetAsync(m->adam_v.p, 0, 4 * n, m->st), "memset");
  hip_check(hipMemsetAsync(m->grad.p, 0, 4 * n, m->st), "memset");
  hip_check(hipStreamSynchronize(m->st), "sync");
  m->adam_t = 0;
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_get_params(const fs_guess_model* m, float* out) {
  if (!m || !out) { fs::set_error("fs_guess_model_get_params: null argument"); return FS_ERR_BAD_INPUT; }
  FSG_GUARD_BEGIN
  hip_check(hipStreamSynchronize(m->st), "sync");
  hip_check(hipMemcpy(out, m->p32.p, 4 * m->n_params, hipMemcpyDeviceToHost), "D2H");
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_get_grads(const fs_guess_model* m, float* out) {
  if (!m || !out) { fs::set_error("fs_guess_model_get_grads: null argument"); return FS_ERR_BAD_INPUT; }
  FSG_GUARD_BEGIN
  hip_check(hipStreamSynchronize(m->st), "sync");
  hip_check(hipMemcpy(out, m->grad.p, 4 * m->n_params, hipMemcpyDeviceToHost), "D2H");
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_set_normalization(fs_guess_model* m, const float* mean, const float* stdv) {
  if (!m || !mean || !stdv) { fs::set_error("fs_guess_model_set_normalization: null argument"); return FS_ERR_BAD_INPUT; }
  for (int c = 0; c < m->cout; ++c) {
    if (!std::isfinite(mean[c]) || !(stdv[c] > 0) || !std::isfinite(stdv[c])) {
      fs::set_error("fs_guess_model_set_normalization: bad value");
      return FS_ERR_BAD_INPUT;
    }
    m->mean[c] = mean[c];
    m->stdv[c] = stdv[c];
  }
  return FS_OK;
}

int fs_guess_model_get_normalization(const fs_guess_model* m, float* mean, float* stdv) {
  if (!m || !mean || !stdv) { fs::set_error("fs_guess_model_get_normalization: null argument"); return FS_ERR_BAD_INPUT; }
  for (int c = 0; c < m->cout; ++c) {
    mean[c] = m->mean[c];
    stdv[c] = m->stdv[c];
  }
  return FS_OK;
}

int fs_guess_prepare_input(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_log_eta, float* d_in) {
  if (!m || !d_log_eta || !d_in || n_cells < 1 || batch < 1) {
    fs::set_error("fs_guess_prepare_input: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FSG_GUARD_BEGIN
  const int64_t P = ipow(2 * (int64_t)n_cells + 1, m->cfg.dim) * batch;
  hipLaunchKernelGGL(k_prep_input, dim3(grid_for(P)), dim3(256), 0, m->st, m->cfg.dim, n_cells, batch, d_log_eta,
                     d_in);
  hip_check(hipStreamSynchronize(m->st), "sync");
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_forward(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_in, float* d_out) {
  if (!m || !d_in || !d_out || n_cells < 1 || batch < 1) {
    fs::set_error("fs_guess_model_forward: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FSG_GUARD_BEGIN
  m->ensure_workspace(n_cells, batch, false);
  m->forward(d_in, batch, m->st);
  hip_check(hipMemcpyAsync(d_out, m->Y, 4 * (size_t)batch * m->cout * m->P, hipMemcpyDeviceToDevice, m->st), "D2D");
  hip_check(hipStreamSynchronize(m->st), "sync");
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_loss_grad(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_in,
                             const float* d_target, double* loss) {
  if (!m || !d_in || !d_target || n_cells < 1 || batch < 1) {
    fs::set_error("fs_guess_model_loss_grad: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FSG_GUARD_BEGIN
  m->ensure_workspace(n_cells, batch, true);
  m->forward(d_in, batch, m->st);
  const int64_t n = (int64_t)batch * m->cout * m->P;
  const int nb = std::min(grid_for(n), 1024);
  double* partial = (double*)((float*)m->part.p);  // reuse the weight-gradient partials as scratch
  double* tot = partial + nb;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->st, n, m->Y, d_target, m->dY, partial);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, m->st, nb, partial, tot);
  double l = 0;
  hip_check(hipMemcpyAsync(&l, tot, 8, hipMemcpyDeviceToHost, m->st), "D2H");
  hip_check(hipStreamSynchronize(m->st), "sync");
  m->backward(d_in, batch, m->st);
  hip_check(hipStreamSynchronize(m->st), "sync");
  if (loss) *loss = l / (double)n;
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_adam_step(fs_guess_model* m, float lr, float beta1, float beta2, float eps, float weight_decay) {
  if (!m) { fs::set_error("fs_guess_model_adam_step: null model"); return FS_ERR_BAD_INPUT; }
  FSG_GUARD_BEGIN
  m->adam_t++;
  const float bc1 = 1.f - std::pow(beta1, (float)m->adam_t), bc2 = 1.f - std::pow(beta2, (float)m->adam_t);
  const int64_t n = m->n_params;
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, m->st, n, (float*)m->p32.p, (const float*)m->grad.p,
                     (float*)m->adam_m.p, (float*)m->adam_v.p, (uint16_t*)m->pbf.p, lr, beta1, beta2, eps,
                     weight_decay, bc1, bc2);
  hip_check(hipStreamSynchronize(m->st), "sync");
  return FS_OK;
  FSG_GUARD_END
}

int fs_gemm_bf16(int32_t M, int32_t N, int32_t K, int32_t batch, const void* A, int32_t a_bf16, int64_t sAb,
                 int64_t sAm, int64_t sAk, const void* B, int32_t b_bf16, int64_t sBb, int64_t sBk, int64_t sBn,
                 float* C, int64_t sCb, int64_t sCm, int64_t sCn, float alpha, float beta) {
  if (!A || !B || !C || M < 0 || N < 0 || K < 0 || batch < 0) {
    fs::set_error("fs_gemm_bf16: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FSG_GUARD_BEGIN
  fs::ensure_device();
  gemm(nullptr, M, N, K, batch, A, a_bf16 != 0, sAb, sAm, sAk, B, b_bf16 != 0, sBb, sBk, sBn, C, sCb, sCm, sCn, alpha,
       beta);
  hip_check(hipGetLastError(), "k_gemm launch");
  hip_check(hipDeviceSynchronize(), "sync");
  return FS_OK;
  FSG_GUARD_END
}

// weight file
static uint64_t fnv1a(const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}

struct FileHeader {
  char magic[4];
  int32_t version;
  fs_guess_config cfg;
  float mean[4], stdv[4];
  int64_t n_params;
  uint64_t hash;
};

int fs_guess_model_save(const fs_guess_model* m, const char* path) {
  if (!m || !path) { fs::set_error("fs_guess_model_save: null argument"); return FS_ERR_BAD_INPUT; }
  FSG_GUARD_BEGIN
  std::vector<float> p(m->n_params);
  hip_check(hipStreamSynchronize(m->st), "sync");
  hip_check(hipMemcpy(p.data(), m->p32.p, 4 * p.size(), hipMemcpyDeviceToHost), "D2H");
  FileHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, "FSGM", 4);
  h.version = 1;
  h.cfg = m->cfg;
  for (int c = 0; c < 4; ++c) { h.mean[c] = m->mean[c]; h.stdv[c] = m->stdv[c]; }
  h.n_params = m->n_params;
  h.hash = fnv1a(p.data(), 4 * p.size());
  FILE* f = std::fopen(path, "wb");
  if (!f) throw FsError(FS_ERR_BAD_INPUT, std::string("cannot write ") + path);
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(p.data(), 4, p.size(), f) == p.size();
  std::fclose(f);
  if (!ok) throw FsError(FS_ERR_BAD_INPUT, std::string("short write to ") + path);
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_load(const char* path, fs_guess_model** out) {
  if (!path || !out) { fs::set_error("fs_guess_model_load: null argument"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  FSG_GUARD_BEGIN
  FILE* f = std::fopen(path, "rb");
  if (!f) throw FsError(FS_ERR_BAD_INPUT, std::string("cannot open ") + path);
  FileHeader h;
  std::vector<float> p;
  bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, "FSGM", 4) == 0 && h.version == 1 &&
            h.n_params > 0 && h.n_params < (int64_t)1 << 32;
  if (ok) {
    p.resize(h.n_params);
    ok = std::fread(p.data(), 4, p.size(), f) == p.size();
  }
  std::fclose(f);
  if (!ok) throw FsError(FS_ERR_BAD_INPUT, std::string("not a guess-model file: ") + path);
  if (fnv1a(p.data(), 4 * p.size()) != h.hash) throw FsError(FS_ERR_BAD_INPUT, "guess-model file: hash mismatch");
  fs_guess_model* m = nullptr;
  int rc = fs_guess_model_create(&h.cfg, &m);
  if (rc != FS_OK) return rc;
  std::unique_ptr<fs_guess_model> g(m);
  if (g->n_params != h.n_params) throw FsError(FS_ERR_BAD_INPUT, "guess-model file: parameter count mismatch");
  rc = fs_guess_model_set_params(g.get(), p.data());
  if (rc != FS_OK) return rc;
  rc = fs_guess_model_set_normalization(g.get(), h.mean, h.stdv);
  if (rc != FS_OK) return rc;
  *out = g.release();
  return FS_OK;
  FSG_GUARD_END
}

int fs_guess_model_time_forward(fs_guess_model* m, int32_t n_cells, int32_t batch, int32_t reps, double* ms,
                                double* flops) {
  if (!m || n_cells < 1 || batch < 1 || reps < 1) {
    fs::set_error("fs_guess_model_time_forward: bad argument");
    return FS_ERR_BAD_INPUT;
  }
  FSG_GUARD_BEGIN
  m->ensure_workspace(n_cells, batch, false);
  float* X = m->Xin;
  hip_check(hipMemsetAsync(X, 0, 4 * (size_t)batch * m->cin * m->P, m->st), "memset");
  m->forward(X, batch, m->st);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventRecord(e0, m->st), "event");
  for (int i = 0; i < reps; ++i) m->forward(X, batch, m->st);
  hip_check(hipEventRecord(e1, m->st), "event");
  hip_check(hipEventSynchronize(e1), "sync");
  float t = 0;
  hip_check(hipEventElapsedTime(&t, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t / reps;
  if (flops) {
    // 2 M N K per GEMM of the forward pass
    const double B = batch, w = m->cfg.width, pw = m->cfg.proj, P = (double)m->P, N1 = (double)m->N1,
                 M = (double)m->M, Mm = (double)m->Mm;
    const int d = m->cfg.dim;
    double f = 2.0 * B * P * w * m->cin + 2.0 * B * P * pw * w + 2.0 * B * P * m->cout * pw;
    double spec = 0;
    for (int a = 0; a < d; ++a) spec += 2.0 * 2.0 * B * w * std::pow(N1, d - 1 - a) * std::pow(M, a) * N1 * M;
    spec += 2.0 * B * w * w * Mm;
    f += m->cfg.layers * (2.0 * B * P * w * w + spec);
    *flops = f;
  }
  return FS_OK;
  FSG_GUARD_END
}

}  // extern "C"
