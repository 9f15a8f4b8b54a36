// CDNA4 (gfx950) device code for the Stokes solve: CSR SpMV, deterministic
// reductions, block-scheduled ILU(k) factorisation and triangular solves, the
// fused inner CG on the pressure mass matrix, and the Krylov helpers.
// 64-wide wavefronts throughout; FP64 or FP32 values via templates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsk {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;      // threads per workgroup for streaming kernels
constexpr int RED_BLOCKS = 512; // partial-sum slots of the two-stage reductions
constexpr int MAXM = 32;        // max vectors in one multi-dot (restart + 1)

// ---------------------------------------------------------------------------
// small helpers
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

template <int G, typename T>
__device__ __forceinline__ T group_reduce(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// block-wide sum of a double into lane 0 of wave 0 (BLOCK threads)
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = group_reduce<WAVE>(v);
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / WAVE); ++i) s += sh[i];
  return s;
}

// ---------------------------------------------------------------------------
// Grid barrier for cooperative launches: arrival counter + generation word,
// with a wall-clock bounded wait so a stuck barrier ends the kernel with an
// error flag instead of hanging the device.
struct GridBarrier {
  unsigned int count;
  unsigned int gen;
  int err;
  int pad;
};

__device__ __forceinline__ bool grid_sync(GridBarrier* gb, unsigned int& gen, uint64_t timeout_ticks) {
  __syncthreads();
  bool ok = true;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned int arrived =
      __hip_atomic_fetch_add(&gb->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nb - 1) {
      __hip_atomic_store(&gb->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&gb->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint64_t t0 = wall_clock64();
      while (__hip_atomic_load(&gb->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (wall_clock64() - t0 > timeout_ticks) {
          __hip_atomic_store(&gb->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        if (__hip_atomic_load(&gb->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  ++gen;
  __syncthreads();
  if (__hip_atomic_load(&gb->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = false;
  return ok;
}

// ---------------------------------------------------------------------------
// CSR SpMV: G lanes per row, BLOCK/G rows per workgroup.
//   MODE 0: y = M x
//   MODE 1: y = b - M x
//   MODE 2: y = M x - b       (columns >= col_limit skipped)
template <typename TV, typename TX, typename TY, typename TB, typename TACC, int G, int MODE>
__global__ __launch_bounds__(BLOCK) void k_spmv(int64_t nrows, int64_t row_off,
                                                const int64_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci,
                                                const TV* __restrict__ val, const TX* __restrict__ x,
                                                TY* __restrict__ y, const TB* __restrict__ b,
                                                int64_t col_limit, int64_t x_off) {
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t row = gid / G;
  const int lane = (int)(gid % G);
  if (row >= nrows) return;
  const int64_t r = row + row_off;
  const int64_t p0 = rp[r], p1 = rp[r + 1];
  TACC acc = 0;
  int64_t p = p0 + lane;
  for (; p + G < p1; p += 2 * G) {
    const int32_t c0 = ld_nt(ci + p), c1 = ld_nt(ci + p + G);
    const TV v0 = ld_nt(val + p), v1 = ld_nt(val + p + G);
    if (MODE != 2 || c0 < col_limit) acc += (TACC)v0 * (TACC)x[c0 - x_off];
    if (MODE != 2 || c1 < col_limit) acc += (TACC)v1 * (TACC)x[c1 - x_off];
  }
  if (p < p1) {
    const int32_t c0 = ld_nt(ci + p);
    const TV v0 = ld_nt(val + p);
    if (MODE != 2 || c0 < col_limit) acc += (TACC)v0 * (TACC)x[c0 - x_off];
  }
  acc = group_reduce<G>(acc);
  if (lane == 0) {
    if (MODE == 0) y[row] = (TY)acc;
    else if (MODE == 1) y[row] = (TY)((TACC)b[row] - acc);
    else y[row] = (TY)(acc - (TACC)b[row]);
  }
}

// ---------------------------------------------------------------------------
// deterministic two-stage reductions
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_dot_partial(int64_t n, const T* __restrict__ a,
                                                       const T* __restrict__ b,
                                                       double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] = sum(partial[0..nb)); mode 1: sqrt
static __global__ void k_reduce_final(int nb, const double* __restrict__ partial, double* __restrict__ out,
                               int mode) {
  __shared__ double sh[BLOCK / WAVE];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = mode == 1 ? sqrt(s) : s;
}

// multi-dot: partial[b*MAXM + k] = sum_i V[k*ld + i] * w[i], k < m
template <typename TW>
__global__ __launch_bounds__(BLOCK) void k_mdot_partial(int64_t n, int m, const double* __restrict__ V,
                                                        int64_t ld, const TW* __restrict__ w,
                                                        double* __restrict__ partial) {
  __shared__ double sh[MAXM][BLOCK / WAVE];
  double acc[MAXM];
#pragma unroll
  for (int k = 0; k < MAXM; ++k) acc[k] = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const double wi = (double)w[i];
#pragma unroll
    for (int k = 0; k < MAXM; ++k)
      if (k < m) acc[k] += V[(int64_t)k * ld + i] * wi;
  }
  const int wv = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
#pragma unroll
  for (int k = 0; k < MAXM; ++k) {
    if (k < m) {
      double v = group_reduce<WAVE>(acc[k]);
      if (l == 0) sh[k][wv] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)m) {
    double s = 0;
    for (int i = 0; i < BLOCK / WAVE; ++i) s += sh[threadIdx.x][i];
    partial[(int64_t)blockIdx.x * MAXM + threadIdx.x] = s;
  }
}

// out[k] (+)= sum_b partial[b*MAXM + k]
static __global__ void k_mdot_final(int nb, int m, const double* __restrict__ partial, double* __restrict__ out,
                             int accumulate) {
  const int k = blockIdx.x;
  if (k >= m) return;
  __shared__ double sh[BLOCK / WAVE];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[(int64_t)i * MAXM + k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[k] = accumulate ? out[k] + s : s;
}

// w -= sum_k h[k] V_k ;   optionally partial[b] = sum_i w_i^2 (for the norm)
template <typename TW>
__global__ __launch_bounds__(BLOCK) void k_gs_update(int64_t n, int m, const double* __restrict__ V,
                                                     int64_t ld, const double* __restrict__ h,
                                                     TW* __restrict__ w, double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  double hk[MAXM];
#pragma unroll
  for (int k = 0; k < MAXM; ++k) hk[k] = k < m ? h[k] : 0.0;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double wi = (double)w[i];
#pragma unroll
    for (int k = 0; k < MAXM; ++k)
      if (k < m) wi -= hk[k] * V[(int64_t)k * ld + i];
    w[i] = (TW)wi;
    s += wi * wi;
  }
  if (partial) {
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// dst = src / scal[idx] (0 when the norm is zero: lucky breakdown)
template <typename TS>
__global__ __launch_bounds__(BLOCK) void k_scale_by_inv(int64_t n, const TS* __restrict__ src,
                                                        double* __restrict__ dst,
                                                        const double* __restrict__ scal, int idx) {
  const double s = scal[idx];
  const double f = s != 0.0 ? 1.0 / s : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    dst[i] = (double)src[i] * f;
}

// x += sum_k y[k] Z_k
template <typename TZ>
__global__ __launch_bounds__(BLOCK) void k_update_x(int64_t n, int m, const TZ* __restrict__ Z, int64_t ld,
                                                    const double* __restrict__ y, double* __restrict__ x) {
  double yk[MAXM];
#pragma unroll
  for (int k = 0; k < MAXM; ++k) yk[k] = k < m ? y[k] : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double s = x[i];
#pragma unroll
    for (int k = 0; k < MAXM; ++k)
      if (k < m) s += yk[k] * (double)Z[(int64_t)k * ld + i];
    x[i] = s;
  }
}

template <typename TA, typename TB>
__global__ __launch_bounds__(BLOCK) void k_convert(int64_t n, const TA* __restrict__ a, TB* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    b[i] = (TB)a[i];
}

// y[i] = x[perm[i]]  (gather)   or  y[perm[i]] = x[i] (scatter)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_permute(int64_t n, const int32_t* __restrict__ perm,
                                                   const T* __restrict__ x, T* __restrict__ y, int scatter) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    if (scatter) y[perm[i]] = x[i];
    else y[i] = x[perm[i]];
  }
}

// dst[map[e]] = src[e]
static __global__ __launch_bounds__(BLOCK) void k_scatter_vals(int64_t n, const int64_t* __restrict__ map,
                                                        const double* __restrict__ src,
                                                        double* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    dst[map[i]] = src[i];
}

// ---------------------------------------------------------------------------
// ILU factor description (rows in factor order, sorted columns, diagonal in
// place). Per row, offsets relative to rp[r]:
//   [0, o_lint)        L entries in earlier blocks (final before the block runs)
//   [o_lint, o_diag)   L entries inside the row's block
//   o_diag             diagonal
//   (o_diag, o_uext)   U entries inside the block
//   [o_uext, len)      U entries in later blocks
// lpre / upre: start of the row's in-block L / U entries within the block's
// compacted list (prefix over the block's rows).
template <typename T>
struct IluView {
  const int64_t* rp;
  const int32_t* ci;
  T* v;
  const int32_t* o_lint;
  const int32_t* o_diag;
  const int32_t* o_uext;
  const int32_t* lpre;
  const int32_t* upre;
  const int32_t* blk_ptr;
};

// LDS layout of one wave's block workspace
template <typename T>
struct BlockLds {
  T* s;        // [max_rows]   running values
  T* d;        // [max_rows]   diagonal (backward)
  int32_t* st; // [max_rows]   start in compacted list
  int32_t* ct; // [max_rows]   count
  T* ev;       // [cap]        in-block entry values
  uint16_t* ec;// [cap]        in-block entry local columns
};

template <typename T>
__device__ __forceinline__ BlockLds<T> carve_lds(char* base, int max_rows, int cap) {
  BlockLds<T> L;
  char* p = base;
  L.s = (T*)p; p += sizeof(T) * max_rows;
  L.d = (T*)p; p += sizeof(T) * max_rows;
  L.ev = (T*)p; p += sizeof(T) * cap;
  L.st = (int32_t*)p; p += 4 * max_rows;
  L.ct = (int32_t*)p; p += 4 * max_rows;
  L.ec = (uint16_t*)p;
  return L;
}

__host__ __device__ inline size_t lds_bytes_per_wave(int max_rows, int cap, int tsize) {
  size_t b = (size_t)tsize * (2 * max_rows + cap) + 8 * (size_t)max_rows + 2 * (size_t)cap;
  return (b + 15) & ~(size_t)15;
}

// Forward solve of one block: y = L^{-1} b restricted to the block's rows,
// L unit lower (UNIT=true) or with its diagonal (transposed factors).
// One wave; G lanes per row for the part that reads earlier blocks.
template <typename T, typename TB, typename TY, int G, bool UNIT>
__device__ void ilu_fwd_block(const IluView<T>& F, int blk, const TB* __restrict__ b, TY* __restrict__ y,
                              BlockLds<T> L, int cap) {
  const int lane = threadIdx.x % WAVE;
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  const int nr = r1 - r0;
  const int grp = lane / G, gl = lane % G;
  constexpr int NG = WAVE / G;
  // phase A: contributions of earlier blocks + stage in-block entries
  for (int i = grp; i < nr; i += NG) {
    const int r = r0 + i;
    const int64_t base = F.rp[r];
    const int olint = F.o_lint[r], odiag = F.o_diag[r];
    T acc = 0;
    for (int p = gl; p < olint; p += G) acc += ld_nt(F.v + base + p) * (T)y[F.ci[base + p]];
    acc = group_reduce<G>(acc);
    const int st = F.lpre[r], cnt = odiag - olint;
    if (gl == 0) {
      L.s[i] = (T)b[r] - acc;
      L.st[i] = st;
      L.ct[i] = cnt;
      if (!UNIT) L.d[i] = F.v[base + odiag];
    }
    if (st + cnt <= cap)
      for (int t = gl; t < cnt; t += G) {
        L.ev[st + t] = F.v[base + olint + t];
        L.ec[st + t] = (uint16_t)(F.ci[base + olint + t] - r0);
      }
  }
  wave_sync();
  // phase B: rows in order, dependencies inside the block from LDS
  for (int i = 0; i < nr; ++i) {
    const int st = L.st[i], cnt = L.ct[i];
    T acc = 0;
    if (st + cnt <= cap) {
      for (int t = lane; t < cnt; t += WAVE) acc += L.ev[st + t] * L.s[L.ec[st + t]];
    } else {  // entries did not fit in LDS: read them from memory
      const int r = r0 + i;
      const int64_t base = F.rp[r] + F.o_lint[r];
      for (int t = lane; t < cnt; t += WAVE) acc += F.v[base + t] * L.s[F.ci[base + t] - r0];
    }
    if (cnt > 0) acc = group_reduce<WAVE>(acc);
    if (lane == 0) {
      T yi = L.s[i] - acc;
      if (!UNIT) yi /= L.d[i];
      L.s[i] = yi;
    }
    wave_sync();
  }
  for (int i = lane; i < nr; i += WAVE) y[r0 + i] = (TY)L.s[i];
}

// Backward solve of one block: x = U^{-1} y, U with diagonal (UNIT=false) or
// unit upper (transposed factors).
template <typename T, typename TB, typename TY, int G, bool UNIT>
__device__ void ilu_bwd_block(const IluView<T>& F, int blk, const TB* __restrict__ yin, TY* __restrict__ x,
                              BlockLds<T> L, int cap) {
  const int lane = threadIdx.x % WAVE;
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  const int nr = r1 - r0;
  const int grp = lane / G, gl = lane % G;
  constexpr int NG = WAVE / G;
  for (int i = grp; i < nr; i += NG) {
    const int r = r0 + i;
    const int64_t base = F.rp[r];
    const int len = (int)(F.rp[r + 1] - base);
    const int odiag = F.o_diag[r], ouext = F.o_uext[r];
    T acc = 0;
    for (int p = ouext + gl; p < len; p += G) acc += ld_nt(F.v + base + p) * (T)x[F.ci[base + p]];
    acc = group_reduce<G>(acc);
    const int st = F.upre[r], cnt = ouext - odiag - 1;
    if (gl == 0) {
      L.s[i] = (T)yin[r] - acc;
      L.st[i] = st;
      L.ct[i] = cnt;
      L.d[i] = UNIT ? (T)1 : F.v[base + odiag];
    }
    if (st + cnt <= cap)
      for (int t = gl; t < cnt; t += G) {
        L.ev[st + t] = F.v[base + odiag + 1 + t];
        L.ec[st + t] = (uint16_t)(F.ci[base + odiag + 1 + t] - r0);
      }
  }
  wave_sync();
  for (int i = nr - 1; i >= 0; --i) {
    const int st = L.st[i], cnt = L.ct[i];
    T acc = 0;
    if (st + cnt <= cap) {
      for (int t = lane; t < cnt; t += WAVE) acc += L.ev[st + t] * L.s[L.ec[st + t]];
    } else {
      const int r = r0 + i;
      const int64_t base = F.rp[r] + F.o_diag[r] + 1;
      for (int t = lane; t < cnt; t += WAVE) acc += F.v[base + t] * L.s[F.ci[base + t] - r0];
    }
    if (cnt > 0) acc = group_reduce<WAVE>(acc);
    if (lane == 0) L.s[i] = (L.s[i] - acc) / L.d[i];
    wave_sync();
  }
  for (int i = lane; i < nr; i += WAVE) x[r0 + i] = (TY)L.s[i];
}

// one stage: each wave takes blocks blk0 + wave_id (+ stride)
template <typename T, typename TB, typename TY, int G, bool UNIT, bool FWD>
__global__ __launch_bounds__(BLOCK) void k_ilu_stage(IluView<T> F, int blk0, int blk1, const TB* __restrict__ b,
                                                     TY* __restrict__ y, int max_rows, int cap) {
  extern __shared__ char smem[];
  const int wv = threadIdx.x / WAVE;
  const size_t per_wave = lds_bytes_per_wave(max_rows, cap, sizeof(T));
  BlockLds<T> L = carve_lds<T>(smem + per_wave * wv, max_rows, cap);
  const int nw = gridDim.x * (BLOCK / WAVE);
  for (int blk = blk0 + blockIdx.x * (BLOCK / WAVE) + wv; blk < blk1; blk += nw) {
    if (FWD) ilu_fwd_block<T, TB, TY, G, UNIT>(F, blk, b, y, L, cap);
    else ilu_bwd_block<T, TB, TY, G, UNIT>(F, blk, b, y, L, cap);
  }
}

// ---------------------------------------------------------------------------
// Numeric ILU (include/felspa/linear_algebra/src/ilu.implement.h:438-541):
// each wave factorises the rows of one block in order. The row is held in LDS
// (sorted columns + working values); the source matrix row (columns <
// col_limit) is scattered in, then for each L column k in ascending order
// m = w_k / u_kk and w_j -= m u_kj for the columns j > k of row k that row i
// stores. FP64 throughout; err[0] is set on a zero or non-finite pivot.
static __global__ __launch_bounds__(WAVE) void k_ilu_numeric_stage(
  IluView<double> F, int blk0, int blk1, const int64_t* __restrict__ srp, const int32_t* __restrict__ sci,
  const double* __restrict__ sv, int64_t col_limit, int max_len, int* __restrict__ err) {
  extern __shared__ char smem[];
  int32_t* lc = (int32_t*)smem;
  double* w = (double*)(smem + ((4 * (size_t)max_len + 15) & ~(size_t)15));
  const int lane = threadIdx.x;
  for (int blk = blk0 + blockIdx.x; blk < blk1; blk += gridDim.x) {
    const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
    for (int r = r0; r < r1; ++r) {
      const int64_t base = F.rp[r];
      const int len = (int)(F.rp[r + 1] - base);
      for (int t = lane; t < len; t += WAVE) {
        lc[t] = F.ci[base + t];
        w[t] = 0.0;
      }
      wave_sync();
      for (int64_t p = srp[r] + lane; p < srp[r + 1]; p += WAVE) {
        const int32_t c = sci[p];
        if (c >= col_limit) continue;
        int lo = 0, hi = len;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (lc[mid] < c) lo = mid + 1; else hi = mid;
        }
        if (lo < len && lc[lo] == c) w[lo] = sv[p];
      }
      wave_sync();
      const int nl = F.o_diag[r];
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      for (int t = 0; t < nl; ++t) {
        const int32_t kk = lc[t];
        const int64_t kb = F.rp[kk];
        const int kd = F.o_diag[kk];
        const double ukk = F.v[kb + kd];
        const double mult = w[t] / ukk;
        wave_sync();
        if (lane == 0) w[t] = mult;
        const int64_t q1 = F.rp[kk + 1];
        for (int64_t q = kb + kd + 1 + lane; q < q1; q += WAVE) {
          const int32_t j = F.ci[q];
          int lo = t + 1, hi = len;
          while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (lc[mid] < j) lo = mid + 1; else hi = mid;
          }
          if (lo < len && lc[lo] == j) w[lo] -= mult * F.v[q];
        }
        wave_sync();
      }
      for (int t = lane; t < len; t += WAVE) F.v[base + t] = w[t];
      if (lane == 0) {
        const double d = w[F.o_diag[r]];
        if (!(d != 0.0) || !isfinite(d)) atomicExch(err, 1);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// Inner CG on the pressure mass matrix (InverseMatrixCG::vmult,
// include/felspa/linear_algebra/src/linear_operators.implement.h:64-95, with
// deal.II SolverCG steps and ReductionControl(tol, reduce)), preconditioned by
// ILU_S, as one cooperative kernel. Every workgroup computes the same reduced
// scalars from the same partial sums in the same order.
template <typename T>
struct InnerCgArgs {
  int64_t n;
  const int64_t* mrp;
  const int32_t* mci;
  const T* mv;
  IluView<T> F;
  const int32_t* stage_ptr;  // device copy
  int n_stages;
  int max_rows, cap;
  const T* b;   // rhs
  T* x;         // solution (starts at zero)
  T* g;
  T* h;
  T* d;
  T* q;
  double* partial;  // [2][gridDim.x]
  double tol, reduce;
  int max_it;
  int* iters_out;   // total CG iterations (accumulated)
  int* fail_out;    // set when CG did not converge
  GridBarrier* gb;
  uint64_t timeout;
};

template <typename T, int GM, int GF>
__global__ __launch_bounds__(BLOCK) void k_inner_cg(InnerCgArgs<T> a) {
  extern __shared__ char smem[];
  __shared__ double shr[BLOCK / WAVE];
  __shared__ double tot;
  const int wv = threadIdx.x / WAVE;
  const size_t per_wave = lds_bytes_per_wave(a.max_rows, a.cap, sizeof(T));
  BlockLds<T> L = carve_lds<T>(smem + per_wave * wv, a.max_rows, a.cap);
  unsigned int gen = __hip_atomic_load(&a.gb->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int nwg = gridDim.x;
  const int64_t tid = (int64_t)blockIdx.x * BLOCK + threadIdx.x, nthr = (int64_t)nwg * BLOCK;
  const int gw = blockIdx.x * (BLOCK / WAVE) + wv, nw = nwg * (BLOCK / WAVE);
  int red = 0;
  // reduce a per-thread value over the grid; all workgroups get the sum
  auto allreduce = [&](double v, bool& ok) -> double {
    double* part = a.partial + (red & 1) * nwg;
    ++red;
    double s = block_sum(v, shr);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    ok = grid_sync(a.gb, gen, a.timeout);
    if (threadIdx.x == 0) {
      double t = 0;
      for (int i = 0; i < nwg; ++i) t += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      tot = t;
    }
    __syncthreads();
    return tot;
  };
  auto precond = [&](const T* src, T* dst, bool& ok) {
    for (int s = 0; s < a.n_stages && ok; ++s) {
      for (int blk = a.stage_ptr[s] + gw; blk < a.stage_ptr[s + 1]; blk += nw)
        ilu_fwd_block<T, T, T, GF, true>(a.F, blk, src, dst, L, a.cap);
      ok = grid_sync(a.gb, gen, a.timeout);
    }
    for (int s = a.n_stages - 1; s >= 0 && ok; --s) {
      for (int blk = a.stage_ptr[s] + gw; blk < a.stage_ptr[s + 1]; blk += nw)
        ilu_bwd_block<T, T, T, GF, false>(a.F, blk, dst, dst, L, a.cap);
      ok = grid_sync(a.gb, gen, a.timeout);
    }
  };
  bool ok = true;
  const int64_t n = a.n;
  // x = 0, g = -b
  double s = 0;
  for (int64_t i = tid; i < n; i += nthr) {
    a.x[i] = 0;
    const T gi = -a.b[i];
    a.g[i] = gi;
    s += (double)gi * (double)gi;
  }
  double res = sqrt(allreduce(s, ok));
  if (!ok) return;
  const double reduced_tol = res * a.reduce;
  int it = 0;
  bool done = (res <= reduced_tol) || (res <= a.tol);
  bool failed = false;
  double gh = 0;
  if (!done) {
    precond(a.g, a.h, ok);
    if (!ok) return;
    s = 0;
    for (int64_t i = tid; i < n; i += nthr) {
      const T hi = a.h[i];
      a.d[i] = -hi;
      s += (double)a.g[i] * (double)hi;
    }
    gh = allreduce(s, ok);
    if (!ok) return;
  }
  while (!done) {
    ++it;
    // q = M d, dq = d.q
    s = 0;
    for (int64_t row = tid / GM; row < n; row += nthr / GM) {
      const int gl = (int)(tid % GM);
      T acc = 0;
      for (int64_t p = a.mrp[row] + gl; p < a.mrp[row + 1]; p += GM) acc += a.mv[p] * a.d[a.mci[p]];
      acc = group_reduce<GM>(acc);
      if (gl == 0) {
        a.q[row] = acc;
        s += (double)acc * (double)a.d[row];
      }
    }
    const double dq = allreduce(s, ok);
    if (!ok) return;
    const double alpha = gh / dq;
    s = 0;
    for (int64_t i = tid; i < n; i += nthr) {
      a.x[i] += (T)alpha * a.d[i];
      const T gi = a.g[i] + (T)alpha * a.q[i];
      a.g[i] = gi;
      s += (double)gi * (double)gi;
    }
    res = sqrt(allreduce(s, ok));
    if (!ok) return;
    if (res <= reduced_tol || res <= a.tol) break;
    if (it >= a.max_it || isnan(res)) { failed = true; break; }
    precond(a.g, a.h, ok);
    if (!ok) return;
    s = 0;
    for (int64_t i = tid; i < n; i += nthr) s += (double)a.g[i] * (double)a.h[i];
    const double ghn = allreduce(s, ok);
    if (!ok) return;
    const double beta = ghn / gh;
    gh = ghn;
    for (int64_t i = tid; i < n; i += nthr) a.d[i] = (T)beta * a.d[i] - a.h[i];
    ok = grid_sync(a.gb, gen, a.timeout);
    if (!ok) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.iters_out[0] += it;
    if (failed) a.fail_out[0] = 1;
  }
}

// ---------------------------------------------------------------------------
// Inner PCG on the velocity block A (the A^{-1} of the block preconditioner),
// one launch per step; the scalars alpha/beta are formed inside the consuming
// kernel from the producer's per-workgroup partial sums, summed in a fixed
// order by every workgroup, so no separate reduction launch is needed.

// sum of nb partials (every workgroup computes the same value)
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nb, double* sh) {
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  s = block_sum(s, sh);
  __shared__ double tot;
  if (threadIdx.x == 0) tot = s;
  __syncthreads();
  return tot;
}

// q = A p ; partial[b] = sum over the workgroup's rows of p_i q_i
// (grid-stride over rows so the partial count stays bounded)
template <typename T, int G>
__global__ __launch_bounds__(BLOCK) void k_spmv_dot(int64_t nrows, const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const T* __restrict__ val,
                                                    const T* __restrict__ p, T* __restrict__ q,
                                                    double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  const int lane = threadIdx.x % G;
  const int64_t rows_per_pass = (int64_t)gridDim.x * (BLOCK / G);
  double contrib = 0;
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < nrows; row += rows_per_pass) {
    const int64_t p0 = rp[row], p1 = rp[row + 1];
    T acc = 0;
    int64_t k = p0 + lane;
    for (; k + G < p1; k += 2 * G) {
      const int32_t c0 = ld_nt(ci + k), c1 = ld_nt(ci + k + G);
      const T v0 = ld_nt(val + k), v1 = ld_nt(val + k + G);
      acc += v0 * p[c0];
      acc += v1 * p[c1];
    }
    if (k < p1) acc += ld_nt(val + k) * p[ld_nt(ci + k)];
    acc = group_reduce<G>(acc);
    if (lane == 0) {
      q[row] = acc;
      contrib += (double)acc * (double)p[row];
    }
  }
  contrib = block_sum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = contrib;
}

// alpha = rz / sum(pq partials); x += alpha p; r -= alpha q
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_pcg_update(int64_t n, const double* __restrict__ pq_part, int nb,
                                                      const double* __restrict__ rz, const T* __restrict__ p,
                                                      const T* __restrict__ q, T* __restrict__ x,
                                                      T* __restrict__ r, int first) {
  __shared__ double sh[BLOCK / WAVE];
  const double pq = sum_partials(pq_part, nb, sh);
  const double a = pq != 0.0 ? rz[0] / pq : 0.0;
  const T alpha = (T)a;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const T pi = p[i];
    x[i] = first ? alpha * pi : x[i] + alpha * pi;
    r[i] -= alpha * q[i];
  }
}

// partial[b] = sum r_i z_i
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_dot_partial_t(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                         double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// rz_new = sum(rz partials); beta = rz_new / rz (0 on the first step);
// p = z + beta p; rz <- rz_new (workgroup 0 writes after every workgroup read)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_pcg_direction(int64_t n, const double* __restrict__ rz_part, int nb,
                                                         const double* __restrict__ rz_old,
                                                         double* __restrict__ rz_new, const T* __restrict__ z,
                                                         T* __restrict__ p, int first) {
  __shared__ double sh[BLOCK / WAVE];
  const double rzn = sum_partials(rz_part, nb, sh);
  const double b = (first || rz_old[0] == 0.0) ? 0.0 : rzn / rz_old[0];
  const T beta = (T)b;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    p[i] = first ? z[i] : z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) rz_new[0] = rzn;
}

// r = (T) src
template <typename TS, typename T>
__global__ __launch_bounds__(BLOCK) void k_copy_convert(int64_t n, const TS* __restrict__ src, T* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    dst[i] = (T)src[i];
}

// ---------------------------------------------------------------------------
// Givens / least-squares update of one GMRES column (deal.II SolverGMRES
// documented algorithm): the column is h1 + h2 (the two classical
// Gram-Schmidt passes) over the new norm; previous rotations are applied, a new one is
// formed, gamma updated, |gamma[j+1]| written to est[0].
static __global__ void k_givens(int j, int ldh, const double* __restrict__ h1, const double* __restrict__ h2,
                         const double* __restrict__ nrm, double* __restrict__ hcol, double* __restrict__ H,
                         double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ gam,
                         double* __restrict__ est) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i <= j; ++i) hcol[i] = h1[i] + h2[i];
  hcol[j + 1] = nrm[0];
  for (int i = 0; i < j; ++i) {
    const double hi = hcol[i];
    hcol[i] = cs[i] * hi + sn[i] * hcol[i + 1];
    hcol[i + 1] = -sn[i] * hi + cs[i] * hcol[i + 1];
  }
  const double r = 1.0 / sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
  sn[j] = hcol[j + 1] * r;
  cs[j] = hcol[j] * r;
  hcol[j] = cs[j] * hcol[j] + sn[j] * hcol[j + 1];
  gam[j + 1] = -sn[j] * gam[j];
  gam[j] *= cs[j];
  for (int i = 0; i <= j; ++i) H[(int64_t)i * ldh + j] = hcol[i];
  est[0] = fabs(gam[j + 1]);
}

// y = H^{-1} gamma (upper triangular dim x dim)
static __global__ void k_backsolve(int dim, int ldh, const double* __restrict__ H, const double* __restrict__ gam,
                            double* __restrict__ y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = dim - 1; i >= 0; --i) {
    double s = gam[i];
    for (int k = i + 1; k < dim; ++k) s -= H[(int64_t)i * ldh + k] * y[k];
    y[i] = s / H[(int64_t)i * ldh + i];
  }
}

// gamma[0] = scal[idx]; est[0] = gamma[0]
static __global__ void k_init_gamma(double* __restrict__ gam, const double* __restrict__ scal, int idx,
                             double* __restrict__ est) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    gam[0] = scal[idx];
    est[0] = scal[idx];
  }
}

}  // namespace fsk
