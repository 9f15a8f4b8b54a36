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
constexpr int MAXM = 256;       // max vectors in one multi-dot (restart + 1)
constexpr int MDOT_CHUNK = 32;  // vectors per workgroup row of a multi-dot

// ---------------------------------------------------------------------------
// small helpers
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Orders LDS accesses of one wave for the compiler only: the LDS executes a
// wave's operations in issue order, so no wait is needed between a lane's
// write and another lane's later read.
__device__ __forceinline__ void wave_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
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

// block-wide sum of a double into thread 0 (blockDim.x <= BLOCK threads)
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

// block_sum of two values at once (each summed in block_sum's order)
__device__ __forceinline__ void block_sum2(double& a, double& b, double2* sh) {
  a = group_reduce<WAVE>(a);
  b = group_reduce<WAVE>(b);
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) sh[w] = make_double2(a, b);
  __syncthreads();
  double sa = 0, sb = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / WAVE); ++i) {
      sa += sh[i].x;
      sb += sh[i].y;
    }
  a = sa;
  b = sb;
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
// CSR row dot product with G lanes per row. Each lane keeps 4 column
// indices, 4 values and 4 gathered x values in flight: loads go to clamped
// in-range positions and out-of-range terms are masked afterwards, so the
// compiler emits no per-element branch (a branch per gather serialises the
// gathers on their memory latency).
template <typename TV, typename TX, typename TACC, int G, bool LIMIT>
__device__ __forceinline__ TACC row_dot(const int32_t* __restrict__ ci, const TV* __restrict__ val,
                                        const TX* __restrict__ x, int64_t p0, int64_t p1, int lane,
                                        int64_t col_limit, int64_t x_off) {
  TACC acc = 0;
  const int64_t last = p1 - 1;
  for (int64_t k = p0 + lane; k < p1; k += 4 * G) {
    int32_t c[4];
    TV v[4];
    bool in[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t kk = k + u * G;
      in[u] = kk < p1;
      const int64_t kc = in[u] ? kk : last;
      c[u] = ld_nt(ci + kc);
      v[u] = ld_nt(val + kc);
    }
    TX xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = !LIMIT || c[u] < col_limit;
      in[u] = in[u] && ok;
      xv[u] = x[(ok ? (int64_t)c[u] : x_off) - x_off];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += in[u] ? (TACC)v[u] * (TACC)xv[u] : (TACC)0;
  }
  return group_reduce<G>(acc);
}

// Padded CSR ("CSR4"): every row starts at a multiple of 4 entries and is
// padded to one (padding: value 0, column = a valid column of the row), so a
// lane streams 4 consecutive entries with one 16-byte column load and one
// 16/32-byte value load, and the G lanes of a row read one contiguous span.
template <typename TV>
struct Vec4;
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <>
struct Vec4<float> {
  using type = f32x4;
};
template <>
struct Vec4<double> {
  using type = f64x4;
};

template <typename TV, typename TX, typename TACC, int G, bool LIMIT>
__device__ __forceinline__ TACC row_dot4(const i32x4* __restrict__ ci4, const typename Vec4<TV>::type* __restrict__ val4,
                                         const TX* __restrict__ x, int64_t q0, int64_t q1, int lane,
                                         int64_t col_limit) {
  TACC acc = 0;
  for (int64_t q = q0 + lane; q < q1; q += G) {
    const i32x4 c = __builtin_nontemporal_load(ci4 + q);
    const typename Vec4<TV>::type v = __builtin_nontemporal_load(val4 + q);
    if (LIMIT) {  // clamped gathers, masked values: no branch per gather
      const bool m0 = c.x < col_limit, m1 = c.y < col_limit, m2 = c.z < col_limit, m3 = c.w < col_limit;
      const TX x0 = x[m0 ? c.x : 0], x1 = x[m1 ? c.y : 0], x2 = x[m2 ? c.z : 0], x3 = x[m3 ? c.w : 0];
      acc += (TACC)(m0 ? v.x : (TV)0) * (TACC)x0 + (TACC)(m1 ? v.y : (TV)0) * (TACC)x1 +
             (TACC)(m2 ? v.z : (TV)0) * (TACC)x2 + (TACC)(m3 ? v.w : (TV)0) * (TACC)x3;
    } else {
      const TX x0 = x[c.x], x1 = x[c.y], x2 = x[c.z], x3 = x[c.w];
      acc += (TACC)v.x * (TACC)x0 + (TACC)v.y * (TACC)x1 + (TACC)v.z * (TACC)x2 + (TACC)v.w * (TACC)x3;
    }
  }
  return group_reduce<G>(acc);
}

// Same product with two groups of 4 per lane in flight: both groups' column
// and value loads are issued before any gather (the second group is clamped
// to the first when the row has no more, and masked), so a lane pays one
// load latency and one gather latency per two groups.
template <typename TV, typename TX, typename TACC, int G>
__device__ __forceinline__ TACC row_dot4x2(const i32x4* __restrict__ ci4,
                                           const typename Vec4<TV>::type* __restrict__ val4,
                                           const TX* __restrict__ x, int64_t q0, int64_t q1, int lane) {
  TACC acc = 0, acc2 = 0;
  for (int64_t q = q0 + lane; q < q1; q += 2 * G) {
    const bool two = q + G < q1;
    const int64_t qb = two ? q + G : q;
    const i32x4 c = __builtin_nontemporal_load(ci4 + q);
    const typename Vec4<TV>::type v = __builtin_nontemporal_load(val4 + q);
    const i32x4 c2 = __builtin_nontemporal_load(ci4 + qb);
    const typename Vec4<TV>::type v2 = __builtin_nontemporal_load(val4 + qb);
    const TX x0 = x[c.x], x1 = x[c.y], x2 = x[c.z], x3 = x[c.w];
    const TX y0 = x[c2.x], y1 = x[c2.y], y2 = x[c2.z], y3 = x[c2.w];
    acc += (TACC)v.x * (TACC)x0 + (TACC)v.y * (TACC)x1 + (TACC)v.z * (TACC)x2 + (TACC)v.w * (TACC)x3;
    const TACC s2 = (TACC)v2.x * (TACC)y0 + (TACC)v2.y * (TACC)y1 + (TACC)v2.z * (TACC)y2 + (TACC)v2.w * (TACC)y3;
    acc2 += two ? s2 : (TACC)0;
  }
  return group_reduce<G>(acc + acc2);
}

// y = M x / b - M x / M x - b on a CSR4 matrix (rp4 counts groups of 4)
template <typename TV, typename TX, typename TY, typename TB, typename TACC, int G, int MODE>
__global__ __launch_bounds__(BLOCK) void k_spmv4(int64_t nrows, int64_t row_off, const int64_t* __restrict__ rp4,
                                                 const i32x4* __restrict__ ci4,
                                                 const typename Vec4<TV>::type* __restrict__ val4,
                                                 const TX* __restrict__ x, TY* __restrict__ y, const TB* __restrict__ b,
                                                 int64_t col_limit) {
  const int64_t gid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t row = gid / G;
  const int lane = (int)(gid % G);
  if (row >= nrows) return;
  const int64_t r = row + row_off;
  const TACC acc = MODE == 2 ? row_dot4<TV, TX, TACC, G, true>(ci4, val4, x, rp4[r], rp4[r + 1], lane, col_limit)
                             : row_dot4x2<TV, TX, TACC, G>(ci4, val4, x, rp4[r], rp4[r + 1], lane);
  if (lane == 0) {
    if (MODE == 0) y[row] = (TY)acc;
    else if (MODE == 1) y[row] = (TY)((TACC)b[row] - acc);
    else y[row] = (TY)(acc - (TACC)b[row]);
  }
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i mod 8), and each XCD has its own L2. Renumbering workgroup i as
// (i mod 8) * (n / 8) + i / 8 gives every XCD one contiguous range of the
// work (rows, blocks), so the values it gathers from its neighbourhood stay
// in its own L2; the tail beyond 8 (n / 8) keeps its numbers. (Used by the
// ILU stages, 5 % faster on the 2D 256x256 Q2 velocity system; the CSR4
// A-block SpMV ran slower with it.)
__device__ __forceinline__ int xcd_contiguous(int i, int n) {
  constexpr int NX = 8;
  const int per = n / NX;
  return i < per * NX ? (i % NX) * per + i / NX : i;
}

// q = A p on a CSR4 matrix with TV values and T vectors / arithmetic;
// partial[b] = sum over the workgroup's rows of p_i q_i
template <typename TV, typename T, int G>
__global__ __launch_bounds__(BLOCK) void k_spmv4_dot(int64_t nrows, const int64_t* __restrict__ rp4,
                                                     const i32x4* __restrict__ ci4,
                                                     const typename Vec4<TV>::type* __restrict__ val4,
                                                     const T* __restrict__ p, T* __restrict__ q,
                                                     double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  const int lane = threadIdx.x % G;
  const int wg = blockIdx.x;
  const int64_t row = (int64_t)wg * (BLOCK / G) + threadIdx.x / G;
  double contrib = 0;
  if (row < nrows) {
    const T acc = row_dot4x2<TV, T, T, G>(ci4, val4, p, rp4[row], rp4[row + 1], lane);
    if (lane == 0) {
      q[row] = acc;
      contrib = (double)acc * (double)p[row];
    }
  }
  contrib = block_sum(contrib, sh);
  if (threadIdx.x == 0) partial[wg] = contrib;
}

// q = A p and per-block partials of p.q on the A block stored by lattice
// blocks (the ILU's blocks): each block's entries reference the block's own
// column list (its distinct columns, ascending) through 1- or 2-byte indices,
// rows padded to chunks of 4. One workgroup per block: the chunks (16-byte
// value vectors + index words) and the column list load in one round, the
// block's p values are gathered once into LDS, chunk sums go to LDS, G
// threads per row add the row's chunks (thread g: chunks g, g + G, ...) and
// the G partial sums are added in order.
template <int IB>
struct Idx4;
// TX: the type of p when it differs from T (the product K z of the outer
// method: FP32 z, FP64 sums and q); partial may then be null (no dot).
template <typename TV, typename T, int IB, int NT, typename TX = T>
__global__ __launch_bounds__(NT) void k_spmv_blk_dot(const i32x4* __restrict__ desc, const int2* __restrict__ cinfo,
                                                     const int32_t* __restrict__ rng, const void* __restrict__ xidx,
                                                     const TV* __restrict__ xv, const int32_t* __restrict__ cols,
                                                     const TX* __restrict__ p, T* __restrict__ q,
                                                     double* __restrict__ partial, int ccap,
                                                     const T* __restrict__ rdot, double* __restrict__ partial2) {
  extern __shared__ __attribute__((aligned(16))) char dsm[];
  __shared__ double sh[NT / WAVE];
  typedef TV V4 __attribute__((ext_vector_type(4)));
  typedef typename Idx4<IB>::type IX;
  constexpr int R = 4;
  T* xs = (T*)dsm;
  T* part = xs + ccap;
  const int tid = threadIdx.x;
  const int blk = blockIdx.x;
  const i32x4 d = desc[blk];
  const int2 ci = cinfo[blk];
  const int r0 = d.x, nr = d.y, nc = d.w;
  const IX* isrc = (const IX*)xidx + d.z;
  const V4* vsrc = (const V4*)xv + d.z;
  IX cx[R];
  V4 vv[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int c = u * NT + tid;
    const int cc = c < nc ? c : 0;
    cx[u] = __builtin_nontemporal_load(isrc + cc);
    vv[u] = __builtin_nontemporal_load(vsrc + cc);
  }
  // this thread's row range, p and rdot, and the first round of the column
  // list, all in the first burst: unconditional loads of clamped indices
  // (nr >= 1; the column list holds at least the block's own rows)
  const int tr = tid < nr ? tid : nr - 1;
  const T pr = (T)p[r0 + tr];
  const T rr = rdot ? rdot[r0 + tr] : pr;
  // G threads per row in the row sums
  const int G = 4 * nr <= NT ? 4 : 2 * nr <= NT ? 2 : 1;
  const int grow = tid / G, gq = tid - grow * G;
  const uint32_t rg = (uint32_t)rng[r0 + (grow < nr ? grow : nr - 1)];
  const int cs = grow < nr ? (int)(rg & 0xffffu) : 0;
  const int cn = grow < nr ? (int)(rg >> 16) : 0;
  const i32x4* c4 = (const i32x4*)(cols + ci.x);
  const i32x4 cf = __builtin_nontemporal_load(c4 + (4 * tid < ci.y ? tid : 0));
  {
    const T a0 = (T)p[cf.x], a1 = (T)p[cf.y], a2 = (T)p[cf.z], a3 = (T)p[cf.w];
    if (4 * tid < ci.y) {
      xs[4 * tid] = a0;
      xs[4 * tid + 1] = a1;
      xs[4 * tid + 2] = a2;
      xs[4 * tid + 3] = a3;
    }
  }
  for (int k = NT + tid; 4 * k < ci.y; k += NT) {
    const i32x4 c = __builtin_nontemporal_load(c4 + k);
    const T a0 = (T)p[c.x], a1 = (T)p[c.y], a2 = (T)p[c.z], a3 = (T)p[c.w];
    xs[4 * k] = a0;
    xs[4 * k + 1] = a1;
    xs[4 * k + 2] = a2;
    xs[4 * k + 3] = a3;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int c = u * NT + tid;
    if (c < nc) {
      T s = (T)vv[u].x * xs[Idx4<IB>::get(cx[u], 0)];
      s += (T)vv[u].y * xs[Idx4<IB>::get(cx[u], 1)];
      s += (T)vv[u].z * xs[Idx4<IB>::get(cx[u], 2)];
      s += (T)vv[u].w * xs[Idx4<IB>::get(cx[u], 3)];
      part[c] = s;
    }
  }
  for (int c0 = R * NT + tid; c0 < nc; c0 += R * NT) {  // blocks beyond one round
    IX xr[R];
    V4 vr[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int c = c0 + u * NT < nc ? c0 + u * NT : c0;
      xr[u] = __builtin_nontemporal_load(isrc + c);
      vr[u] = __builtin_nontemporal_load(vsrc + c);
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int c = c0 + u * NT < nc ? c0 + u * NT : c0;
      T s = (T)vr[u].x * xs[Idx4<IB>::get(xr[u], 0)];
      s += (T)vr[u].y * xs[Idx4<IB>::get(xr[u], 1)];
      s += (T)vr[u].z * xs[Idx4<IB>::get(xr[u], 2)];
      s += (T)vr[u].w * xs[Idx4<IB>::get(xr[u], 3)];
      part[c] = s;
    }
  }
  __syncthreads();
  __shared__ T red[NT];
  if (grow < nr) {  // thread gq of a row: its chunks gq, gq + G, ...
    T s0 = 0, s1 = 0;
    int k = gq;
    for (; k + G < cn; k += 2 * G) {
      s0 += part[cs + k];
      s1 += part[cs + k + G];
    }
    if (k < cn) s0 += part[cs + k];
    red[gq * nr + grow] = s0 + s1;
  }
  __syncthreads();
  double contrib = 0, c2 = 0;
  if (tid < nr) {
    T acc = red[tid];
    for (int g = 1; g < G; ++g) acc += red[g * nr + tid];
    q[r0 + tid] = acc;
    contrib = (double)acc * (double)pr;
    c2 = rdot ? (double)rr * (double)pr : 0.0;
  }
  if (rdot) {  // also partial2[blk] = sum rdot_i p_i over the block's rows
    __shared__ double2 sh2[NT / WAVE];
    block_sum2(contrib, c2, sh2);
    if (tid == 0) {
      partial[blk] = contrib;
      partial2[blk] = c2;
    }
  } else if (partial) {
    contrib = block_sum(contrib, sh);
    if (tid == 0) partial[blk] = contrib;
  }
}

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
  const TACC acc = row_dot<TV, TX, TACC, G, MODE == 2>(ci, val, x, rp[r], rp[r + 1], lane, col_limit, x_off);
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
  __shared__ double sh[1024 / WAVE];  // up to 1024 threads
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = mode == 1 ? sqrt(s) : s;
}

// multi-dot: partial[b*MAXM + k] = sum_i V[k*ld + i] * w[i], k < m. The
// vectors are taken MDOT_CHUNK at a time, one chunk per blockIdx.y, so each
// thread keeps MDOT_CHUNK sums in registers whatever the basis size.
template <typename TW>
__global__ __launch_bounds__(BLOCK) void k_mdot_partial(int64_t n, int m, const double* __restrict__ V,
                                                        int64_t ld, const TW* __restrict__ w,
                                                        double* __restrict__ partial) {
  __shared__ double sh[MDOT_CHUNK][BLOCK / WAVE];
  const int k0 = blockIdx.y * MDOT_CHUNK;
  const int mc = min(MDOT_CHUNK, m - k0);
  const double* Vc = V + (int64_t)k0 * ld;
  double acc[MDOT_CHUNK];
#pragma unroll
  for (int k = 0; k < MDOT_CHUNK; ++k) acc[k] = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const double wi = (double)w[i];
#pragma unroll
    for (int k = 0; k < MDOT_CHUNK; ++k)
      if (k < mc) acc[k] += Vc[(int64_t)k * ld + i] * wi;
  }
  const int wv = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
#pragma unroll
  for (int k = 0; k < MDOT_CHUNK; ++k) {
    if (k < mc) {
      double v = group_reduce<WAVE>(acc[k]);
      if (l == 0) sh[k][wv] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)mc) {
    double s = 0;
    for (int i = 0; i < BLOCK / WAVE; ++i) s += sh[threadIdx.x][i];
    partial[(int64_t)blockIdx.x * MAXM + k0 + threadIdx.x] = s;
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

// w -= sum_k h[k] V_k ;   optionally partial[b] = sum_i w_i^2 (for the norm).
// The coefficients sit in LDS; the sum runs over k in order for every row.
template <typename TW>
__global__ __launch_bounds__(BLOCK) void k_gs_update(int64_t n, int m, const double* __restrict__ V,
                                                     int64_t ld, const double* __restrict__ h,
                                                     TW* __restrict__ w, double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  __shared__ double hk[MAXM];
  for (int k = threadIdx.x; k < m; k += BLOCK) hk[k] = h[k];
  __syncthreads();
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double wi = (double)w[i];
    int k = 0;
    for (; k + 8 <= m; k += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = V[(int64_t)(k + u) * ld + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) wi -= hk[k + u] * v[u];
    }
    for (; k < m; ++k) wi -= hk[k] * V[(int64_t)k * ld + i];
    w[i] = (TW)wi;
    s += wi * wi;
  }
  if (partial) {
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// Arnoldi step with delayed reorthogonalisation (DCGS2: Bielich, Langou,
// Thomas, Swirydowicz, Yamazaki, Boman, "Low-synch Gram-Schmidt with delayed
// reorthogonalization for Krylov solvers", Parallel Computing 2022). At step
// j the basis column j holds u_j, the previous step's vector after one
// projection pass, scaled by that pass's norm nu_{j-1}; columns 0..j-1 are
// final. One pass over V gives both a = V_{<j}^T u_j, alpha = u_j.u_j (the
// reorthogonalisation of u_j) and t = V_{<j}^T w, tau = u_j.w (the projection
// of the new vector w); then v_j = (u_j - V_{<j} a) / r with r^2 = alpha -
// a.a, v_j.w = (tau - a.t) / r, and one more pass writes v_j and the
// projected w. The previous Hessenberg column gets nu_{j-1} a added and
// h_{j,j-1} = nu_{j-1} r.

// partial[b*MAXM + k] = V_k . x1, partial2[b*MAXM + k] = V_k . x2 (k < m),
// by row tiles: a workgroup takes tiles of BLOCK R rows (grid-stride),
// holds its rows of x1 and x2 in registers and walks the
// m basis vectors U at a time (U R loads in flight per thread), so V, x1
// and x2 are each read once; per vector every wave sums its lanes and keeps
// the sum in LDS, and the waves' sums are added in wave order at the end.
template <int R, int U>
__global__ __launch_bounds__(BLOCK) void k_mdot2_rows(int64_t n, int m, const double* __restrict__ V, int64_t ld,
                                                      const double* __restrict__ x1, const double* __restrict__ x2,
                                                      double* __restrict__ partial, double* __restrict__ partial2) {
  constexpr int NW = BLOCK / WAVE;
  __shared__ double s1[NW][MAXM], s2[NW][MAXM];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  for (int k = lane; k < m; k += WAVE) {  // each wave its own row (in order within the wave)
    s1[wv][k] = 0;
    s2[wv][k] = 0;
  }
  for (int64_t i0 = (int64_t)blockIdx.x * BLOCK * R; i0 < n; i0 += (int64_t)gridDim.x * BLOCK * R) {
    double a1[R], a2[R];
    int64_t ri[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = i0 + (int64_t)r * BLOCK + threadIdx.x;
      ri[r] = i < n ? i : 0;  // past the end: row 0 times zero weights
      a1[r] = i < n ? x1[i] : 0.0;
      a2[r] = i < n ? x2[i] : 0.0;
    }
    for (int k0 = 0; k0 < m; k0 += U) {
      double v[U][R];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t col = (int64_t)(k0 + u < m ? k0 + u : m - 1) * ld;
#pragma unroll
        for (int r = 0; r < R; ++r) v[u][r] = V[col + ri[r]];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double p1 = 0, p2 = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          p1 += v[u][r] * a1[r];
          p2 += v[u][r] * a2[r];
        }
        p1 = group_reduce<WAVE>(p1);
        p2 = group_reduce<WAVE>(p2);
        if (lane == 0 && k0 + u < m) {
          s1[wv][k0 + u] += p1;
          s2[wv][k0 + u] += p2;
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < m; k += BLOCK) {
    double t1 = 0, t2 = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      t1 += s1[w][k];
      t2 += s2[w][k];
    }
    partial[(int64_t)blockIdx.x * MAXM + k] = t1;
    partial2[(int64_t)blockIdx.x * MAXM + k] = t2;
  }
}

// out[k] = sum_b partial[b*MAXM + k] (blocks k < m), out2[k] from partial2
// (blocks m + k)
static __global__ void k_mdot2_final(int nb, int m, const double* __restrict__ partial,
                                     const double* __restrict__ partial2, double* __restrict__ out,
                                     double* __restrict__ out2) {
  const bool second = (int)blockIdx.x >= m;
  const int k = second ? blockIdx.x - m : blockIdx.x;
  const double* p = second ? partial2 : partial;
  __shared__ double sh[BLOCK / WAVE];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[(int64_t)i * MAXM + k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) (second ? out2 : out)[k] = s;
}

// From a, alpha (h1[0..j]) and t, tau (h2[0..j]) the step's 1 / r and
// v_j.w (dc[0], dc[1]); the previous column corrected into raw0 (rows 0..j,
// from the raw column j - 1 in raw1 and nu_{j-1} = nrm[0]) and the new raw
// column j (rows 0..j) into raw1. u_0 = v_0 is final: r = 1. The workgroup
// stages the inputs in LDS; thread 0 sums in index order.
static __global__ __launch_bounds__(BLOCK) void k_dcgs2_coef(int j, const double* __restrict__ h1,
                                                             const double* __restrict__ h2,
                                                             const double* __restrict__ nrm,
                                                             double* __restrict__ raw0, double* __restrict__ raw1,
                                                             double* __restrict__ dc) {
  __shared__ double la[MAXM], lt[MAXM], lr[MAXM];
  for (int i = threadIdx.x; i <= j; i += blockDim.x) {
    la[i] = h1[i];
    lt[i] = h2[i];
    lr[i] = raw1[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (j == 0) {
    dc[0] = 1.0;
    dc[1] = lt[0];
    raw1[0] = lt[0];
    return;
  }
  double aa = 0, at = 0;
  for (int i = 0; i < j; ++i) {
    aa += la[i] * la[i];
    at += la[i] * lt[i];
  }
  // r^2 <= 0: u_j lies in the span of the basis (a breakdown the projection
  // pass did not see); v_j = 0 then
  const double r2 = la[j] - aa;
  const double r = r2 > 0.0 ? sqrt(r2) : 0.0;
  const double rinv = r > 0.0 ? 1.0 / r : 0.0;
  const double hjj = (lt[j] - at) * rinv;
  const double nu = nrm[0];
  for (int i = 0; i < j; ++i) raw0[i] = lr[i] + nu * la[i];
  raw0[j] = nu * r;
  for (int i = 0; i < j; ++i) raw1[i] = lt[i];
  raw1[j] = hjj;
  dc[0] = rinv;
  dc[1] = hjj;
}

// v_j = (u_j - sum_{k<j} a_k V_k) / r into column j (j > 0) and
// w -= sum_{k<j} t_k V_k + (v_j.w) v_j; partial[b] = sum_i w_i^2
static __global__ __launch_bounds__(BLOCK) void k_dcgs2_update(int64_t n, int j, double* __restrict__ V, int64_t ld,
                                                               const double* __restrict__ a,
                                                               const double* __restrict__ t,
                                                               const double* __restrict__ dc,
                                                               double* __restrict__ w,
                                                               double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  __shared__ double ak[MAXM], tk[MAXM];
  for (int k = threadIdx.x; k < j; k += BLOCK) {
    ak[k] = a[k];
    tk[k] = t[k];
  }
  __syncthreads();
  const double rinv = dc[0], hjj = dc[1];
  double* u = V + (int64_t)j * ld;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double sa = 0, st = 0;
    int k = 0;
    for (; k + 8 <= j; k += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = V[(int64_t)(k + q) * ld + i];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        sa += ak[k + q] * v[q];
        st += tk[k + q] * v[q];
      }
    }
    for (; k < j; ++k) {
      const double v = V[(int64_t)k * ld + i];
      sa += ak[k] * v;
      st += tk[k] * v;
    }
    double vj = u[i];
    if (j > 0) {
      vj = (vj - sa) * rinv;
      u[i] = vj;
    }
    const double wi = w[i] - st - hjj * vj;
    w[i] = wi;
    s += wi * wi;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// One workgroup: nu_j = |w| from the update's partials (nrm[0]); then
// thread 0 commits column j - 1 (its corrected raw column raw0, rows 0..j:
// rotations 0..j-2, its own rotation, gamma from gsave = gamma_{j-1} before
// that rotation) and makes column j provisional (raw1 rows 0..j and h_{j+1,j}
// = nu_j, the norm before u_{j+1}'s reorthogonalisation), saving gamma_j
// first; est = |gamma_{j+1}|. A cycle's last column stays provisional. The
// rotations and raw columns are staged in LDS by the workgroup; the
// rotation sweep carries the running element in a register (the same
// arithmetic as k_givens).
static __global__ __launch_bounds__(BLOCK) void k_givens_lag(int j, int ldh, int nb, const double* __restrict__ partial,
                                                             const double* __restrict__ raw0,
                                                             const double* __restrict__ raw1,
                                                             double* __restrict__ H, double* __restrict__ cs,
                                                             double* __restrict__ sn, double* __restrict__ gam,
                                                             double* __restrict__ gsave, double* __restrict__ nrm,
                                                             double* __restrict__ est) {
  __shared__ double sh[BLOCK / WAVE];
  __shared__ double lc[MAXM], ls[MAXM], c0[MAXM + 1], c1[MAXM + 2];
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    lc[i] = cs[i];
    ls[i] = sn[i];
  }
  for (int i = threadIdx.x; i <= j; i += blockDim.x) {
    c0[i] = raw0[i];
    c1[i] = raw1[i];
  }
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  s = block_sum(s, sh);  // (its barriers also publish the staged values)
  if (threadIdx.x != 0) return;
  const double nu = sqrt(s);
  nrm[0] = nu;
  auto rotate = [&](double* col, int c) {  // column c in col[0..c+1]
    double carry = col[0];
    for (int i = 0; i < c; ++i) {
      const double nx = col[i + 1];
      const double ci = lc[i], si = ls[i];
      col[i] = ci * carry + si * nx;
      carry = -si * carry + ci * nx;
    }
    const double b = col[c + 1];
    const double r = 1.0 / sqrt(carry * carry + b * b);
    const double sc = b * r, cc = carry * r;
    ls[c] = sc;
    lc[c] = cc;
    sn[c] = sc;
    cs[c] = cc;
    col[c] = cc * carry + sc * b;
    for (int i = 0; i <= c; ++i) H[(int64_t)i * ldh + c] = col[i];
  };
  if (j > 0) {
    rotate(c0, j - 1);
    const double g = gsave[0];
    gam[j] = -ls[j - 1] * g;
    gam[j - 1] = lc[j - 1] * g;
  }
  c1[j + 1] = nu;
  rotate(c1, j);
  const double gj = gam[j];
  gsave[0] = gj;
  gam[j + 1] = -ls[j] * gj;
  gam[j] = gj * lc[j];
  est[0] = fabs(gam[j + 1]);
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

// x += sum_k y[k] Z_k (coefficients in LDS, the sum in k order per row)
template <typename TZ>
__global__ __launch_bounds__(BLOCK) void k_update_x(int64_t n, int m, const TZ* __restrict__ Z, int64_t ld,
                                                    const double* __restrict__ y, double* __restrict__ x) {
  __shared__ double yk[MAXM];
  for (int k = threadIdx.x; k < m; k += BLOCK) yk[k] = y[k];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double s = x[i];
    int k = 0;
    for (; k + 8 <= m; k += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (double)Z[(int64_t)(k + u) * ld + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += yk[k + u] * v[u];
    }
    for (; k < m; ++k) s += yk[k] * (double)Z[(int64_t)k * ld + i];
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

// d64[i] = src[pos[i]], d32[i] = (float) src[pos[i]]
static __global__ __launch_bounds__(BLOCK) void k_gather_vals(int64_t n, const int64_t* __restrict__ pos,
                                                              const double* __restrict__ src,
                                                              double* __restrict__ d64, float* __restrict__ d32) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const double v = src[pos[i]];
    d64[i] = v;
    d32[i] = (float)v;
  }
}

// External entries of one triangle of an ILU factor, block after block in
// chunks of 4 (the dense stage kernels' layout): block b = desc[b] {first
// row, rows, first chunk, chunks}; row r's entries start at chunk first +
// (rng[r] & 0xffff) and fill rng[r] >> 16 chunks. Its entries are the factor
// row's positions [0, o_lint) (lower) or [o_uext, len) (upper), in order,
// then zeros up to the chunk count. FP64 and FP32 copies.
static __global__ __launch_bounds__(BLOCK) void k_gather_ext(const i32x4* __restrict__ desc,
                                                             const int32_t* __restrict__ rng,
                                                             const int64_t* __restrict__ rp,
                                                             const int32_t* __restrict__ o_lint,
                                                             const int32_t* __restrict__ o_uext, int lower,
                                                             const double* __restrict__ v, double* __restrict__ x64,
                                                             float* __restrict__ x32, int nb) {
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const i32x4 d = desc[b];
    for (int r = d.x; r < d.x + d.y; ++r) {
      const uint32_t g = (uint32_t)rng[r];
      const int64_t dst = 4 * ((int64_t)d.z + (g & 0xffffu));
      const int ne = 4 * (int)(g >> 16);
      const int64_t src = rp[r] + (lower ? 0 : o_uext[r]);
      const int cnt = lower ? o_lint[r] : (int)(rp[r + 1] - rp[r]) - o_uext[r];
      for (int e = threadIdx.x; e < ne; e += BLOCK) {
        const double val = e < cnt ? v[src + e] : 0.0;
        x64[dst + e] = val;
        x32[dst + e] = (float)val;
      }
    }
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
// The in-block entries are also kept transposed (column order) so the
// sequential part of a block solve pushes each finished value to the rows
// that depend on it: lcp[c]..lcp[c+1] are the in-block L entries of column c
// (their local rows in lrow, values in lv), ucp / urow / uv likewise for U.
template <typename T>
struct IluView {
  const int64_t* rp;
  const int32_t* ci;
  T* v;
  const int32_t* o_lint;
  const int32_t* o_diag;
  const int32_t* o_uext;
  const int32_t* blk_ptr;
  const int64_t* lcp;
  const uint16_t* lrow;
  const T* lv;
  const int64_t* ucp;
  const uint16_t* urow;
  const T* uv;
};

// LDS layout of one wave's block workspace
template <typename T>
struct BlockLds {
  T* s;         // [max_rows]   running values
  T* d;         // [max_rows]   diagonal
  T* ev;        // [cap]        in-block entry values (column order)
  int32_t* st;  // [max_rows+1] start of each column's entries
  uint16_t* ec; // [cap]        in-block entry local rows
};

template <typename T>
__device__ __forceinline__ BlockLds<T> carve_lds(char* base, int max_rows, int cap) {
  BlockLds<T> L;
  char* p = base;
  L.s = (T*)p; p += sizeof(T) * max_rows;
  L.d = (T*)p; p += sizeof(T) * max_rows;
  L.ev = (T*)p; p += sizeof(T) * cap;
  L.st = (int32_t*)p; p += 4 * (max_rows + 1);
  L.ec = (uint16_t*)p;
  return L;
}

__host__ __device__ inline size_t lds_bytes_per_wave(int max_rows, int cap, int tsize) {
  size_t b = (size_t)tsize * (2 * max_rows + cap) + 4 * (size_t)(max_rows + 1) + 2 * (size_t)cap;
  return (b + 15) & ~(size_t)15;
}

// Sequential part of a block solve. Rows are finished in order (forward:
// 0..nr-1, backward: nr-1..0); a finished value v_i = s_i (/ d_i) is pushed
// to its dependent rows, s_r -= e_ri v_i, one entry per lane. The next row's
// column list is fetched while the current row is pushed, so the chain per
// row is one LDS read, one FMA and one LDS write. EV/EC are LDS pointers, or
// global memory for blocks whose entries did not fit.
template <typename T, bool UNIT, bool FWD, typename PV, typename PC>
__device__ __forceinline__ void push_phase(BlockLds<T>& L, PV ev, PC ec, int nr, int lane) {
  if (nr <= 0) return;
  const bool divide = FWD ? !UNIT : true;
  int i = FWD ? 0 : nr - 1;
  const int step = FWD ? 1 : -1;
  int k0 = L.st[i], k1 = L.st[i + 1];
  int k = k0 + lane;
  bool act = k < k1;
  int kc = act ? k : 0;
  int row = (int)ec[kc];
  T val = (T)ev[kc];
  for (int it = 0; it < nr; ++it, i += step) {
    const int cur_k0 = k0, cur_k1 = k1;
    const bool cur_act = act;
    const int cur_row = row;
    const T cur_val = val;
    // prefetch the next row's column list
    if (it + 1 < nr) {
      const int ni = i + step;
      k0 = L.st[ni];
      k1 = L.st[ni + 1];
      k = k0 + lane;
      act = k < k1;
      kc = act ? k : 0;
      row = (int)ec[kc];
      val = (T)ev[kc];
    }
    T vi = L.s[i];
    if (divide) vi /= L.d[i];
    const T sr = cur_act ? L.s[cur_row] : (T)0;
    if (divide && lane == 0) L.s[i] = vi;
    if (cur_act) L.s[cur_row] = sr - cur_val * vi;
    // rows with more than 64 dependents (3D blocks)
    for (int kk = cur_k0 + WAVE + lane; kk < cur_k1; kk += WAVE) {
      const int rr = ec[kk];
      L.s[rr] -= (T)ev[kk] * vi;
    }
    wave_order();
  }
}

// Forward solve of one block: y = L^{-1} b on the block's rows; L unit lower
// (UNIT=true) or with its diagonal (transposed factors). NT threads of the
// caller cooperate (a whole workgroup, or one wave): the part that reads
// earlier blocks runs over all rows at once, G threads per row; the in-block
// part then runs on the first wave (push_phase), the others wait.
template <typename T, typename TB, typename TY, int G, bool UNIT, int NT>
__device__ void ilu_fwd_block(const IluView<T>& F, int blk, const TB* __restrict__ b, TY* __restrict__ y,
                              BlockLds<T> L, int cap, int tid) {
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  const int nr = r1 - r0;
  const int grp = tid / G, gl = tid % G;
  constexpr int NG = NT / G;
  const int64_t e0 = F.lcp[r0], ne = F.lcp[r1] - e0;
  const bool in_lds = ne <= cap;
  for (int i = grp; i < nr; i += NG) {
    const int r = r0 + i;
    const int64_t base = F.rp[r];
    const int olint = F.o_lint[r];
    const T acc = row_dot<T, TY, T, G, false>(F.ci, F.v, y, base, base + olint, gl, 0, 0);
    if (gl == 0) {
      L.s[i] = (T)b[r] - acc;
      if (!UNIT) L.d[i] = F.v[base + F.o_diag[r]];
    }
  }
  for (int i = tid; i <= nr; i += NT) L.st[i] = (int32_t)(F.lcp[r0 + i] - e0);
  if (in_lds)
    for (int64_t k = tid; k < ne; k += NT) {
      L.ev[k] = F.lv[e0 + k];
      L.ec[k] = F.lrow[e0 + k];
    }
  if (NT > WAVE) __syncthreads(); else wave_sync();
  if (ne == 0) {  // no in-block entries (point-coloured blocks): rows independent
    if (!UNIT)
      for (int i = tid; i < nr; i += NT) L.s[i] /= L.d[i];
  } else if (tid < WAVE) {
    if (in_lds) push_phase<T, UNIT, true>(L, L.ev, L.ec, nr, tid);
    else push_phase<T, UNIT, true>(L, F.lv + e0, F.lrow + e0, nr, tid);
  }
  if (NT > WAVE) __syncthreads(); else wave_sync();
  for (int i = tid; i < nr; i += NT) y[r0 + i] = (TY)L.s[i];
}

// Backward solve of one block: x = U^{-1} y, U with its diagonal (UNIT=false)
// or unit upper (transposed factors); same organisation as the forward solve.
template <typename T, typename TB, typename TY, int G, bool UNIT, int NT>
__device__ void ilu_bwd_block(const IluView<T>& F, int blk, const TB* __restrict__ yin, TY* __restrict__ x,
                              BlockLds<T> L, int cap, int tid) {
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  const int nr = r1 - r0;
  const int grp = tid / G, gl = tid % G;
  constexpr int NG = NT / G;
  const int64_t e0 = F.ucp[r0], ne = F.ucp[r1] - e0;
  const bool in_lds = ne <= cap;
  for (int i = grp; i < nr; i += NG) {
    const int r = r0 + i;
    const int64_t base = F.rp[r];
    const int len = (int)(F.rp[r + 1] - base);
    const int ouext = F.o_uext[r];
    const T acc = row_dot<T, TY, T, G, false>(F.ci, F.v, x, base + ouext, base + len, gl, 0, 0);
    if (gl == 0) {
      L.s[i] = (T)yin[r] - acc;
      L.d[i] = UNIT ? (T)1 : F.v[base + F.o_diag[r]];
    }
  }
  for (int i = tid; i <= nr; i += NT) L.st[i] = (int32_t)(F.ucp[r0 + i] - e0);
  if (in_lds)
    for (int64_t k = tid; k < ne; k += NT) {
      L.ev[k] = F.uv[e0 + k];
      L.ec[k] = F.urow[e0 + k];
    }
  if (NT > WAVE) __syncthreads(); else wave_sync();
  if (ne == 0) {  // no in-block entries (point-coloured blocks): rows independent
    for (int i = tid; i < nr; i += NT) L.s[i] /= L.d[i];
  } else if (tid < WAVE) {
    if (in_lds) push_phase<T, UNIT, false>(L, L.ev, L.ec, nr, tid);
    else push_phase<T, UNIT, false>(L, F.uv + e0, F.urow + e0, nr, tid);
  }
  if (NT > WAVE) __syncthreads(); else wave_sync();
  for (int i = tid; i < nr; i += NT) x[r0 + i] = (TY)L.s[i];
}

// one stage: each workgroup solves blocks blk0 + blockIdx.x (+ stride)
template <typename T, typename TB, typename TY, int G, bool UNIT, bool FWD>
__global__ __launch_bounds__(BLOCK) void k_ilu_stage(IluView<T> F, int blk0, int blk1, const TB* __restrict__ b,
                                                     TY* __restrict__ y, int max_rows, int cap) {
  extern __shared__ char smem[];
  BlockLds<T> L = carve_lds<T>(smem, max_rows, cap);
  for (int blk = blk0 + blockIdx.x; blk < blk1; blk += gridDim.x) {
    if (FWD) ilu_fwd_block<T, TB, TY, G, UNIT, BLOCK>(F, blk, b, y, L, cap, threadIdx.x);
    else ilu_bwd_block<T, TB, TY, G, UNIT, BLOCK>(F, blk, b, y, L, cap, threadIdx.x);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Triangular solves with explicit inverses of the in-block factors.
// For block b (rows r0 .. r0+nr-1) the stage computes
//   forward   y_b = L_bb^{-1} (b_b - L_b,ext y_ext)
//   backward  x_b = U_bb^{-1} (y_b - U_b,ext x_ext)
// which is the substitution of ilu.implement.h:568-594 regrouped by blocks:
// the rows of one block depend on each other only through L_bb / U_bb, whose
// inverses (dense triangles, packed by columns) are formed once per numeric
// factorisation. The sequential chain through a block becomes a dense
// mat-vec whose column reads are coalesced across the lanes (rows).
// Packing (nr rows): L^{-1}(i,j), i >= j, at j*nr - j(j-1)/2 + (i-j);
//                    U^{-1}(i,j), i <= j, at j(j+1)/2 + i.
__host__ __device__ inline int64_t packed_tri_size(int64_t nr) { return nr * (nr + 1) / 2; }

// Per-block operands of one triangle (L or U) for the dense stage kernel:
// the block's external entries (columns in other blocks) stored block after
// block, row by row (xptr[r] .. xptr[r+1] for row r, so a block's entries
// are one contiguous span), and the packed inverse of its in-block triangle.
template <typename T>
struct DenseTri {
  const i32x4* desc;      // per block: first row, rows, first chunk, chunks
  const int32_t* rng;     // per row: first chunk in the block | chunks << 16
  const void* xidx;       // per external entry: its column's index in the block's halo (IB bytes)
  const T* xv;            // external values, chunks of 4
  const int32_t* hcol;    // halo columns, every block's from a multiple of 4
  const int2* halo;       // per block: {first halo column, halo columns}
  const int64_t* inv_off; // per block: start of its packed triangle(s)
  const T* inv;           // packed inverse of the in-block factor
  const int2* split;      // per block: {S, first partial}; S > 1: its chunks
                          // were summed by S workgroups of k_ilu_ext_partials
  void* part;             // row partial sums of split blocks, S x rows each
};

// The update step of the inner PCG on A in Chronopoulos-Gear form (one
// reduction per step; s = A z, q = A p by recurrence):
//   p = z + beta p, q = s + beta q   (p = z, q = s on the first step)
//   x += alpha p (x = alpha p on the first step), r -= alpha q
// carried out by the forward stage kernels of the next ILU application on the
// rows they read r from: every row belongs to exactly one stage, so the
// vectors are updated once.
template <typename T>
struct PcgFuse {
  const double* alpha;  // device scalars of this step
  const double* beta;
  const T* z;  // z and s = A z of the previous step
  const T* s;
  T* p;
  T* q;
  T* x;
  T* r;  // null: no update
  int first;
};

// 4 halo indices of one chunk: IB = 1 byte each (one word) or 2 (two words)
template <int IB>
struct Idx4;
template <>
struct Idx4<1> {
  typedef uint32_t type;
  static __device__ __forceinline__ int get(type v, int e) { return (int)((v >> (8 * e)) & 0xffu); }
};
template <>
struct Idx4<2> {
  typedef uint32_t type __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ int get(type v, int e) {
    return (int)(((e < 2 ? v.x : v.y) >> (16 * (e & 1))) & 0xffffu);
  }
};

// Blocks with more external chunks than a piece are summed in pieces by
// k_ilu_ext_partials, one workgroup per piece: pieces of 1536 chunks, or of
// 3072 for factors with long rows (more bytes in flight per workgroup and
// fewer, larger halos)
constexpr int DENSE_PIECE_SMALL = 1536;
constexpr int DENSE_PIECE_LARGE = 3072;
constexpr int DENSE_SPLIT_NT = 512;

// Chunks of external entries per thread loaded with the inverse in the
// dense stage kernel; part_cap >= dense_rc * threads or >= every block's chunks
template <typename TV, typename TY>
__host__ __device__ constexpr int dense_rc() {
  return sizeof(TV) + sizeof(TY) <= 8 ? 5 : 3;
}

// The block's halo values (T) into LDS hv: thread k gathers halo columns
// 4k..4k+3, 4(k + NT).., U groups of 4 in flight, from the finished values yin
template <typename T, typename TY, int NT, int U>
__device__ __forceinline__ void gather_halo(const int32_t* __restrict__ hcol, int nh, const TY* yin, T* hv,
                                            int tid, int k0 = 0) {
  const i32x4* h4 = (const i32x4*)hcol;
  const int nu = (nh + 3) / 4;
  for (int k = k0 + tid; k < nu; k += U * NT) {
    i32x4 c[U];
    TY v[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(h4 + (k + u * NT < nu ? k + u * NT : k));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u][0] = yin[c[u].x];
      v[u][1] = yin[c[u].y];
      v[u][2] = yin[c[u].z];
      v[u][3] = yin[c[u].w];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int ku = k + u * NT < nu ? k + u * NT : k;
      hv[4 * ku] = (T)v[u][0];
      hv[4 * ku + 1] = (T)v[u][1];
      hv[4 * ku + 2] = (T)v[u][2];
      hv[4 * ku + 3] = (T)v[u][3];
    }
  }
}

// one chunk's sum v . hv[idx] in entry order
template <typename T, typename V4, int IB>
__device__ __forceinline__ T chunk_dot(const V4& v, typename Idx4<IB>::type ix, const T* hv) {
  T s = (T)v.x * hv[Idx4<IB>::get(ix, 0)];
  s += (T)v.y * hv[Idx4<IB>::get(ix, 1)];
  s += (T)v.z * hv[Idx4<IB>::get(ix, 2)];
  s += (T)v.w * hv[Idx4<IB>::get(ix, 3)];
  return s;
}

// External sums of one piece (chunks [ca, ce) of block blk, ce - ca <=
// PIECE) of a block with many external entries: the piece's halo (its
// entries' distinct columns; their indices point into it) is gathered into
// LDS, every chunk's sum goes to LDS, then G threads per row sum the row's
// chunks inside the piece (thread g takes chunks g, g + G, ... of the row)
// and the G partial sums are added in order into part[first + piece rows +
// row]. One workgroup per piece, all chunk loads in one round. LDS: PIECE
// chunk sums, the halo, NT partial row sums.
template <typename TV, typename T, typename TY, int IB, int PIECE>
__global__ __launch_bounds__(DENSE_SPLIT_NT) void k_ilu_ext_partials(DenseTri<TV> F, const i32x4* __restrict__ pieces,
                                                                     int pc0, int hcap, const TY* yin) {
  typedef TV V4 __attribute__((ext_vector_type(4)));
  typedef typename Idx4<IB>::type IX;
  constexpr int NT = DENSE_SPLIT_NT;
  constexpr int R = (PIECE + NT - 1) / NT;
  extern __shared__ __attribute__((aligned(16))) char dsm[];
  T* part = (T*)dsm;
  T* hv = part + PIECE;
  T* red = hv + hcap;
  const int tid = threadIdx.x;
  const i32x4 pc = pieces[2 * (pc0 + blockIdx.x)];      // {block, first chunk, end chunk, piece}
  const i32x4 ph = pieces[2 * (pc0 + blockIdx.x) + 1];  // {the piece's halo, halo columns, 0, 0}
  const int blk = pc.x, ca = pc.y, ce = pc.z, piece = pc.w;
  const i32x4 d = F.desc[blk];
  const int2 hl = make_int2(ph.x, ph.y);
  const int r0 = d.x, nr = d.y;
  const IX* isrc = (const IX*)F.xidx + d.z;
  const V4* vsrc = (const V4*)F.xv + d.z;
  // the first round of halo columns goes out before the chunks: the halo
  // gather then waits only for them (loads complete in issue order) and
  // overlaps the chunk loads (a piece's halo is never empty)
  constexpr int UH = 2;
  const int nhu = (hl.y + 3) / 4;
  const i32x4* h4 = (const i32x4*)(F.hcol + hl.x);
  i32x4 hc[UH];
#pragma unroll
  for (int u = 0; u < UH; ++u) {
    const int k = u * NT + tid;
    hc[u] = __builtin_nontemporal_load(h4 + (k < nhu ? k : 0));
  }
  IX ix[R];
  V4 vv[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int c = ca + u * NT + tid;
    const int cc = c < ce ? c : ca;
    ix[u] = __builtin_nontemporal_load(isrc + cc);
    vv[u] = __builtin_nontemporal_load(vsrc + cc);
  }
  // G threads per row (nr <= 256 rows per block)
  const int G = 4 * nr <= NT ? 4 : 2 * nr <= NT ? 2 : 1;
  const int row = tid / G, g = tid - row * G;
  TY hy[UH][4];
#pragma unroll
  for (int u = 0; u < UH; ++u) {  // (column hc[u] of entry 0 stands in past the end)
    hy[u][0] = yin[hc[u].x];
    hy[u][1] = yin[hc[u].y];
    hy[u][2] = yin[hc[u].z];
    hy[u][3] = yin[hc[u].w];
  }
  int cs = 0, cn = 0;
  if (row < nr) {
    const uint32_t rg = (uint32_t)F.rng[r0 + row];
    cs = (int)(rg & 0xffffu);
    cn = (int)(rg >> 16);
  }
#pragma unroll
  for (int u = 0; u < UH; ++u) {
    const int k = u * NT + tid;
    if (k < nhu) {
      hv[4 * k] = (T)hy[u][0];
      hv[4 * k + 1] = (T)hy[u][1];
      hv[4 * k + 2] = (T)hy[u][2];
      hv[4 * k + 3] = (T)hy[u][3];
    }
  }
  gather_halo<T, TY, NT, UH>(F.hcol + hl.x, hl.y, yin, hv, tid, UH * NT);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int c = ca + u * NT + tid;
    if (c < ce) part[c - ca] = chunk_dot<T, V4, IB>(vv[u], ix[u], hv);
  }
  __syncthreads();
  if (row < nr) {
    const int lo = cs > ca ? cs : ca, hi = cs + cn < ce ? cs + cn : ce;
    T s0 = 0, s1 = 0;
    int k = lo + g;
    for (; k + G < hi; k += 2 * G) {
      s0 += part[k - ca];
      s1 += part[k + G - ca];
    }
    if (k < hi) s0 += part[k - ca];
    red[g * nr + row] = s0 + s1;
  }
  __syncthreads();
  if (tid < nr) {
    const int2 sp = F.split[blk];
    T s = red[tid];
    for (int q = 1; q < G; ++q) s += red[q * nr + tid];
    ((T*)F.part)[sp.y + (size_t)piece * nr + tid] = s;
  }
}

// Lane gl's share of row i of a packed inverse times t, columns gl, gl + G,
// ...: lower (columns j <= i, column j starting at j n - j (j - 1) / 2) and
// upper (columns j >= i, column j starting at j (j + 1) / 2); four
// independent partial sums so that the LDS reads overlap
template <typename TV, typename T, int G>
__device__ __forceinline__ T lower_row_dot(const TV* M, const T* t, int n, int i, int gl) {
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = gl;
  for (; j + 3 * G <= i; j += 4 * G) {
    const int j1 = j + G, j2 = j + 2 * G, j3 = j + 3 * G;
    a0 += (T)M[j * n - j * (j - 1) / 2 + i - j] * t[j];
    a1 += (T)M[j1 * n - j1 * (j1 - 1) / 2 + i - j1] * t[j1];
    a2 += (T)M[j2 * n - j2 * (j2 - 1) / 2 + i - j2] * t[j2];
    a3 += (T)M[j3 * n - j3 * (j3 - 1) / 2 + i - j3] * t[j3];
  }
  for (; j <= i; j += G) a0 += (T)M[j * n - j * (j - 1) / 2 + i - j] * t[j];
  return (a0 + a1) + (a2 + a3);
}
template <typename TV, typename T, int G>
__device__ __forceinline__ T upper_row_dot(const TV* M, const T* t, int n, int i, int gl) {
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = i + gl;
  for (; j + 3 * G < n; j += 4 * G) {
    const int j1 = j + G, j2 = j + 2 * G, j3 = j + 3 * G;
    a0 += (T)M[j * (j + 1) / 2 + i] * t[j];
    a1 += (T)M[j1 * (j1 + 1) / 2 + i] * t[j1];
    a2 += (T)M[j2 * (j2 + 1) / 2 + i] * t[j2];
    a3 += (T)M[j3 * (j3 + 1) / 2 + i] * t[j3];
  }
  for (; j < n; j += G) a0 += (T)M[j * (j + 1) / 2 + i] * t[j];
  return (a0 + a1) + (a2 + a3);
}

// Dense stage kernels: MODE 0 forward stage, 1 backward stage, 2 the last
// stage forward then backward in one launch (its rows have no external U
// entries: every later column of a last-stage row lies in its own block).
enum { DS_FWD = 0, DS_BWD = 1, DS_BOTH = 2 };

// LDS of one dense stage workgroup: packed inverse(s) (tv bytes), the
// block's halo values, one partial sum per external chunk and the
// right-hand side(s) (ta bytes)
__host__ __device__ inline size_t dense_lds_bytes(int inv_cap, int hcap, int part_cap, int tv, int ta, int mode) {
  const int ninv = mode == DS_BOTH ? 2 : 1;
  return (size_t)tv * ninv * inv_cap + (size_t)ta * (hcap + part_cap + 256 * ninv);
}

// One stage, one block per workgroup of NT threads. The block's external
// entries are chunks of 4 (every row's list padded with value-0 entries),
// each entry a value and the index of its column in the block's halo (the
// block's distinct external columns); thread k reads chunks k, k + NT, ... as
// 16-byte value vectors and 4- or 8-byte index words, consecutive threads at
// consecutive addresses; the packed inverse(s) (starting at multiples of 4
// elements) and the halo columns load in the same round:
//  1. issue the first RI inverse units, RC chunks and the halo columns of
//     every thread; then gather the halo's finished values yin[col] into LDS;
//  2. chunk sums v . halo[idx] into LDS; per row: t_i = b_i - (sum of the
//     row's chunk sums in order), in passes of part_cap chunks;
//  3. G threads per row: y_i = sum_j inv(i, j) t_j from LDS (MODE 2 then
//     y_i = sum_j U_inv(i, j) t'_j).
// The order of every sum is fixed, so results do not depend on scheduling.
// Blocks split into pieces (S > 1) take their external sums from the pieces.
template <typename TV, typename T, typename TB, typename TY, int NT, int G, int MODE, int IB>
__global__ __launch_bounds__(NT) void k_ilu_stage_dense(DenseTri<TV> F, const TV* __restrict__ uinv, int blk0,
                                                        int inv_cap, int hcap, int part_cap, const TB* b,
                                                        const TY* yin, TY* y, PcgFuse<T> pf) {
  extern __shared__ __attribute__((aligned(16))) char dsm[];
  typedef TV V4 __attribute__((ext_vector_type(4)));
  typedef typename Idx4<IB>::type IX;
  constexpr int VN = 16 / sizeof(TV);
  constexpr bool FWD = MODE != DS_BWD;
  constexpr int NINV = MODE == DS_BOTH ? 2 : 1;
  constexpr int RI = sizeof(TV) == 4 ? 3 : 5;  // inverse units per thread, first round
  constexpr int RC = dense_rc<TV, TY>();         // chunks per thread, first round
  TV* Ms = (TV*)dsm;
  TV* Us = Ms + inv_cap;
  T* hv = (T*)(Ms + NINV * inv_cap);
  T* part = hv + hcap;
  T* t = part + part_cap;
  T* t2 = t + 256;
  const int tid = threadIdx.x;
  const int blk = blk0 + xcd_contiguous(blockIdx.x, gridDim.x);
  const i32x4 d = F.desc[blk];
  const int2 sp = F.split[blk];
  const int2 hl = F.halo[blk];
  // a split block's external sums come from the pieces
  const bool whole = sp.x <= 1;
  const int r0 = d.x, nr = d.y, nc = whole ? d.w : 0, nh = whole ? hl.y : 0;
  const int64_t mo = F.inv_off[blk];
  const int nv = (nr * (nr + 1) / 2 + VN - 1) / VN;
  const i32x4* isrc = (const i32x4*)(F.inv + mo);
  const i32x4* usrc = MODE == DS_BOTH ? (const i32x4*)(uinv + mo) : nullptr;
  const IX* xsrc = (const IX*)F.xidx + d.z;
  const V4* vsrc = (const V4*)F.xv + d.z;
  // unconditional loads of a clamped index (nv >= 1): a load under a branch
  // into an element of iv makes the compiler wait for it before the next one
  i32x4 iv[RI], uv[RI];
#pragma unroll
  for (int u = 0; u < RI; ++u) {
    const int k = u * NT + tid < nv ? u * NT + tid : 0;
    iv[u] = __builtin_nontemporal_load(isrc + k);
    if (MODE == DS_BOTH) uv[u] = __builtin_nontemporal_load(usrc + k);
  }
  IX cx[RC];
  V4 vv[RC];
#pragma unroll
  for (int u = 0; u < RC; ++u) {
    const int c = u * NT + tid;
    if (c < nc) {
      cx[u] = __builtin_nontemporal_load(xsrc + c);
      vv[u] = __builtin_nontemporal_load(vsrc + c);
    }
  }
  // the first round of the halo's column indices in the same burst (from a
  // valid address when the block has no halo: the values are then unused)
  constexpr int UH = NT >= 512 ? 2 : 1;
  const int nhu = (nh + 3) / 4;
  const i32x4* h4 = nhu > 0 ? (const i32x4*)(F.hcol + hl.x) : F.desc;
  i32x4 hc[UH];
#pragma unroll
  for (int u = 0; u < UH; ++u) {
    const int k = u * NT + tid;
    hc[u] = __builtin_nontemporal_load(h4 + (k < nhu ? k : 0));
  }
  // one row per thread (nr <= NT)
  int cs = 0, cn = 0;
  T bi = 0;
  if (tid < nr) {
    const uint32_t rg = (uint32_t)F.rng[r0 + tid];
    cs = (int)(rg & 0xffffu);
    cn = (int)(rg >> 16);
    if (FWD && pf.r) {  // the PCG update of this row (k_cg1_update's arithmetic)
      const T alpha = (T)pf.alpha[0], beta = (T)pf.beta[0];
      const int i = r0 + tid;
      // every load before the first store (the pointers may alias, so the
      // compiler keeps program order), unconditionally (values the first
      // step does not use are discarded by the selects)
      const T zi = pf.z[i], si = pf.s[i], po = pf.p[i], qo = pf.q[i], xo = pf.x[i], ro = pf.r[i];
      const T pi = pf.first ? zi : zi + beta * po;
      const T qi = pf.first ? si : si + beta * qo;
      pf.p[i] = pi;
      pf.q[i] = qi;
      pf.x[i] = pf.first ? alpha * pi : xo + alpha * pi;
      const T ri = ro - alpha * qi;
      pf.r[i] = ri;
      bi = ri;
    } else {
      bi = (T)b[r0 + tid];
    }
    if (!whole) {  // partials in piece order, 8 loads in flight
      const T* P = (const T*)F.part + sp.y + tid;
      int q = 0;
      for (; q + 8 <= sp.x; q += 8) {
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = P[(size_t)(q + u) * nr];
#pragma unroll
        for (int u = 0; u < 8; ++u) bi -= a[u];
      }
      for (; q < sp.x; ++q) bi -= P[(size_t)q * nr];
    }
  }
  // the first round's halo values (column 0 stands in past the end), in
  // flight while the inverse goes to LDS
  TY hy[UH][4];
#pragma unroll
  for (int u = 0; u < UH; ++u) {
    const bool in = u * NT + tid < nhu;
    hy[u][0] = yin[in ? hc[u].x : 0];
    hy[u][1] = yin[in ? hc[u].y : 0];
    hy[u][2] = yin[in ? hc[u].z : 0];
    hy[u][3] = yin[in ? hc[u].w : 0];
  }
#pragma unroll
  for (int u = 0; u < RI; ++u) {
    const int k = u * NT + tid;
    if (k < nv) {
      ((i32x4*)Ms)[k] = iv[u];
      if (MODE == DS_BOTH) ((i32x4*)Us)[k] = uv[u];
    }
  }
#pragma unroll
  for (int u = 0; u < UH; ++u) {
    const int k = u * NT + tid;
    if (k < nhu) {
      hv[4 * k] = (T)hy[u][0];
      hv[4 * k + 1] = (T)hy[u][1];
      hv[4 * k + 2] = (T)hy[u][2];
      hv[4 * k + 3] = (T)hy[u][3];
    }
  }
  gather_halo<T, TY, NT, UH>(F.hcol + hl.x, nh, yin, hv, tid, UH * NT);
  // the rest of the inverse: two units per thread in flight
  for (int k = RI * NT + tid; k < nv; k += 2 * NT) {
    const bool two = k + NT < nv;
    const i32x4 a0 = __builtin_nontemporal_load(isrc + k);
    const i32x4 a1 = two ? __builtin_nontemporal_load(isrc + k + NT) : a0;
    ((i32x4*)Ms)[k] = a0;
    if (two) ((i32x4*)Ms)[k + NT] = a1;
    if (MODE == DS_BOTH) {
      const i32x4 u0 = __builtin_nontemporal_load(usrc + k);
      const i32x4 u1 = two ? __builtin_nontemporal_load(usrc + k + NT) : u0;
      ((i32x4*)Us)[k] = u0;
      if (two) ((i32x4*)Us)[k + NT] = u1;
    }
  }
  __syncthreads();
  // chunk sums in passes of part_cap (>= RC NT) partial sums; row sums in
  // chunk order, two partial sums by the parity of the chunk within its row
  T s0 = 0, s1 = 0;
  for (int cb = 0; cb < nc; cb += part_cap) {
    const int ce = nc < cb + part_cap ? nc : cb + part_cap;
    if (cb == 0) {
#pragma unroll
      for (int u = 0; u < RC; ++u) {
        const int c = u * NT + tid;
        if (c < nc) part[c] = chunk_dot<T, V4, IB>(vv[u], cx[u], hv);
      }
    }
    // chunks [cb, ce) not yet summed into part: RR per thread in flight
    // (2 for FP32, which keeps the 128-thread workgroups of a 2D stage
    // resident in one round)
    constexpr int RR = sizeof(TV) == 4 && sizeof(TY) == 4 ? 2 : 4;
    for (int c = (cb == 0 ? RC * NT : cb) + tid; c < ce; c += RR * NT) {
      IX xr[RR];
      V4 vr[RR];
#pragma unroll
      for (int u = 0; u < RR; ++u) {
        const int cu = c + u * NT < ce ? c + u * NT : c;
        xr[u] = __builtin_nontemporal_load(xsrc + cu);
        vr[u] = __builtin_nontemporal_load(vsrc + cu);
      }
#pragma unroll
      for (int u = 0; u < RR; ++u) {
        const int cu = c + u * NT < ce ? c + u * NT : c;
        part[cu - cb] = chunk_dot<T, V4, IB>(vr[u], xr[u], hv);
      }
    }
    __syncthreads();
    const int lo = cs > cb ? cs : cb, hi = cs + cn < ce ? cs + cn : ce;
    for (int k = lo; k < hi; ++k) {
      if ((k - cs) & 1) s1 += part[k - cb];
      else s0 += part[k - cb];
    }
    if (ce < nc) __syncthreads();  // part is overwritten by the next pass
  }
  if (tid < nr) t[tid] = bi - (s0 + s1);
  __syncthreads();
  const int grp = tid / G, gl = tid % G;
  for (int i = grp; i < nr; i += NT / G) {
    T acc = FWD ? lower_row_dot<TV, T, G>(Ms, t, nr, i, gl) : upper_row_dot<TV, T, G>(Ms, t, nr, i, gl);
    acc = group_reduce<G>(acc);
    if (gl == 0) {
      if (MODE == DS_BOTH) t2[i] = acc;
      else y[r0 + i] = (TY)acc;
    }
  }
  if (MODE == DS_BOTH) {
    __syncthreads();
    for (int i = grp; i < nr; i += NT / G) {
      T acc = upper_row_dot<TV, T, G>(Us, t2, nr, i, gl);
      acc = group_reduce<G>(acc);
      if (gl == 0) y[r0 + i] = (TY)acc;
    }
  }
}

// Inverses of the in-block factors of every block, FP64: thread j of the
// block's workgroup forms column j of L_bb^{-1} (unit lower) by forward
// substitution and column j of U_bb^{-1} by backward substitution, reading
// the in-block entries of the factor rows and its own finished column.
static __global__ __launch_bounds__(BLOCK) void k_ilu_block_inverse(IluView<double> F, int nblk,
                                                                     const int64_t* __restrict__ inv_off,
                                                                     double* __restrict__ linv,
                                                                     double* __restrict__ uinv) {
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int r0 = F.blk_ptr[blk], nr = F.blk_ptr[blk + 1] - r0;
    double* Lb = linv + inv_off[blk];
    double* Ub = uinv + inv_off[blk];
    for (int j = threadIdx.x; j < nr; j += blockDim.x) {
      double* xl = Lb + (int64_t)j * nr - (int64_t)j * (j - 1) / 2;  // rows j..nr-1
      xl[0] = 1.0;
      for (int i = j + 1; i < nr; ++i) {
        const int r = r0 + i;
        const int64_t base = F.rp[r];
        double s = 0;
        for (int p = F.o_lint[r]; p < F.o_diag[r]; ++p) {
          const int k = F.ci[base + p] - r0;
          if (k >= j) s += F.v[base + p] * xl[k - j];
        }
        xl[i - j] = -s;
      }
      double* xu = Ub + (int64_t)j * (j + 1) / 2;  // rows 0..j
      {
        const int r = r0 + j;
        xu[j] = 1.0 / F.v[F.rp[r] + F.o_diag[r]];
      }
      for (int i = j - 1; i >= 0; --i) {
        const int r = r0 + i;
        const int64_t base = F.rp[r];
        double s = 0;
        for (int p = F.o_diag[r] + 1; p < F.o_uext[r]; ++p) {
          const int k = F.ci[base + p] - r0;
          if (k <= j) s += F.v[base + p] * xu[k];
        }
        xu[i] = -s / F.v[base + F.o_diag[r]];
      }
    }
  }
}

// The same inverses with the block in LDS: the in-block L and U entries
// scattered into a dense nr x nr tile D (row-major, even leading dimension
// ld), then thread j forms column j of L_bb^{-1} in the tile X (forward
// substitution; the sum over k ascending, as above), the workgroup writes
// X's columns to the packed output, and likewise U_bb^{-1} (backward
// substitution). Lane j reads element (j + t) (ld + 1) + c of D and X:
// consecutive lanes an odd number of elements apart, no bank conflicts.
// LDS: 2 nr_max ld doubles.
template <int NT>
__global__ __launch_bounds__(NT) void k_ilu_block_inverse_lds(IluView<double> F, int nblk,
                                                              const int64_t* __restrict__ inv_off,
                                                              double* __restrict__ linv, double* __restrict__ uinv,
                                                              int ld) {
  extern __shared__ __attribute__((aligned(16))) char ism[];
  double* D = (double*)ism;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  constexpr int NW = NT / WAVE;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int r0 = F.blk_ptr[blk], nr = F.blk_ptr[blk + 1] - r0;
    double* X = D + (size_t)nr * ld;
    for (int e = tid; e < nr * ld; e += NT) D[e] = 0.0;
    __syncthreads();
    for (int i = wv; i < nr; i += NW) {
      const int r = r0 + i;
      const int64_t base = F.rp[r];
      for (int p = F.o_lint[r] + lane; p < F.o_uext[r]; p += WAVE) D[i * ld + (F.ci[base + p] - r0)] = F.v[base + p];
    }
    __syncthreads();
    for (int j = tid; j < nr; j += NT) {
      X[j * ld + j] = 1.0;
      for (int i = j + 1; i < nr; ++i) {
        double s = 0;
        for (int k = j; k < i; ++k) s += D[i * ld + k] * X[k * ld + j];
        X[i * ld + j] = -s;
      }
    }
    __syncthreads();
    double* Lb = linv + inv_off[blk];
    for (int j = wv; j < nr; j += NW) {
      double* col = Lb + (int64_t)j * nr - (int64_t)j * (j - 1) / 2;  // rows j..nr-1
      for (int i = j + lane; i < nr; i += WAVE) col[i - j] = X[i * ld + j];
    }
    __syncthreads();
    for (int j = tid; j < nr; j += NT) {
      X[j * ld + j] = 1.0 / D[j * ld + j];
      for (int i = j - 1; i >= 0; --i) {
        double s = 0;
        for (int k = i + 1; k <= j; ++k) s += D[i * ld + k] * X[k * ld + j];
        X[i * ld + j] = -s / D[i * ld + i];
      }
    }
    __syncthreads();
    double* Ub = uinv + inv_off[blk];
    for (int j = wv; j < nr; j += NW) {
      double* col = Ub + (int64_t)j * (j + 1) / 2;  // rows 0..j
      for (int i = lane; i <= j; i += WAVE) col[i] = X[i * ld + j];
    }
    __syncthreads();  // X and D are rewritten for the next block
  }
}

// ---------------------------------------------------------------------------
// Numeric ILU (include/felspa/linear_algebra/src/ilu.implement.h:438-541):
// each workgroup of NT threads factorises the rows of one block in order.
// The row is held in LDS: its columns, its working values and a hash map
// column -> position. The source matrix row (columns < col_limit) is
// scattered in; then for each L column k in ascending order m = w_k / u_kk
// and w_j -= m u_kj for the columns j > k of row k that row i stores. Rows
// k < i are final. For every L column the start and length of row k's U part
// and the pivot u_kk sit in an LDS ring of two windows of `win` columns,
// refilled one window ahead; row k's U part loads three L columns ahead of
// its use (the barrier between L columns waits for LDS only). FP64
// throughout; err[0] is set on a zero or non-finite pivot.
//
// Hash slots are 32 bits: a 16-bit tag of the column's hash above the
// position in the row (rows of at most 65534 entries); a slot whose tag
// matches is confirmed against the column list. Empty slots are all ones.
constexpr uint32_t NUM_EMPTY = 0xFFFFFFFFu;
constexpr int NUM_MAX_LEN = 65534;

// LDS plan for rows of up to max_len entries of which up to max_nl lie left
// of the diagonal: the L-column window (a power of two, 64..512), the hash
// size (a power of two, load factor <= 1/4 when it fits the limit, never
// above 1/2) and the bytes; 0 bytes when no plan fits.
__host__ __device__ inline size_t ilu_numeric_plan(int max_len, int max_nl, size_t limit, int& hs, int& win) {
  win = 64;
  while (win < max_nl && win < 512) win *= 2;
  const size_t fixed = (size_t)(max_len + 1) * 8 + (size_t)2 * win * (8 + 8 + 4) + (size_t)max_len * 4;
  hs = 64;
  while (hs < 4 * max_len) hs *= 2;
  while (fixed + (size_t)4 * hs > limit && hs / 2 >= 2 * max_len) hs /= 2;
  const size_t bytes = fixed + (size_t)4 * hs;
  return (bytes <= limit && max_len <= NUM_MAX_LEN) ? bytes : 0;
}
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// multiplicative (Fibonacci) hashing: the top lg bits of c * 2^32 / phi pick
// the slot, the low 16 bits are the tag
__device__ __forceinline__ uint32_t num_hash(int32_t c) { return (uint32_t)c * 2654435769u; }
// linear probing from slot h, first slot already read into e
__device__ __forceinline__ int num_probe(const uint32_t* tab, const int32_t* lc, uint32_t mask, int32_t c,
                                         uint32_t tag, uint32_t h, uint32_t e) {
  while (e != NUM_EMPTY) {
    if ((e >> 16) == tag && lc[e & 0xFFFFu] == c) return (int)(e & 0xFFFFu);
    h = (h + 1) & mask;
    e = tab[h];
  }
  return -1;
}
__device__ __forceinline__ int num_find(const uint32_t* tab, const int32_t* lc, uint32_t mask, int lg, int32_t c) {
  const uint32_t hh = num_hash(c);
  const uint32_t h = hh >> (32 - lg);
  return num_probe(tab, lc, mask, c, hh & 0xFFFFu, h, tab[h]);
}
template <int NT, int M>
__global__ __launch_bounds__(NT) void k_ilu_numeric_stage(IluView<double> F, int blk0, int blk1,
                                                          const int64_t* __restrict__ srp,
                                                          const int32_t* __restrict__ sci,
                                                          const double* __restrict__ sv, int64_t col_limit,
                                                          int max_len, int hs, int win, int from_factor,
                                                          int* __restrict__ err) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const uint32_t mask = (uint32_t)hs - 1;
  const int lg = __builtin_ctz((uint32_t)hs);
  const int ring = 2 * win, rmask = ring - 1;
  double* w = (double*)smem;                  // working row; w[max_len]: dump slot
  int64_t* mq = (int64_t*)(w + max_len + 1);  // ring: first U entry of row k
  double* mu = (double*)(mq + ring);          // ring: pivot u_kk
  int32_t* mn = (int32_t*)(mu + ring);        // ring: U entries of row k
  int32_t* lc = mn + ring;                    // columns of the row
  uint32_t* tab = (uint32_t*)(lc + max_len);  // hash map
  const int tid = threadIdx.x;
  const int blk = blk0 + blockIdx.x;
  if (blk >= blk1) return;
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  for (int r = r0; r < r1; ++r) {
    const int64_t base = F.rp[r];
    const int len = (int)(F.rp[r + 1] - base);
    const int nl = F.o_diag[r];
    // from_factor: the external L columns were applied by the stage updates
    // and the row's values are in the factor; only the in-block L columns
    // (positions [o_lint, o_diag)) remain
    const int t0 = from_factor ? F.o_lint[r] : 0;
    for (int h = tid; h < hs; h += NT) tab[h] = NUM_EMPTY;
    lds_barrier();
    for (int t = tid; t < len; t += NT) {
      const int32_t c = F.ci[base + t];
      lc[t] = c;
      w[t] = from_factor && t >= t0 ? F.v[base + t] : 0.0;
      const uint32_t hh = num_hash(c);
      const uint32_t ent = ((hh & 0xFFFFu) << 16) | (uint32_t)t;
      for (uint32_t h = hh >> (32 - lg);; h = (h + 1) & mask)
        if (atomicCAS(&tab[h], NUM_EMPTY, ent) == NUM_EMPTY) break;
    }
    lds_barrier();
    // the U part and pivot of every L column of window wi
    auto load_window = [&](int wi) {
      for (int j = tid; j < win; j += NT) {
        const int t = t0 + wi * win + j;
        if (t < nl) {
          const int32_t k = lc[t];
          const int64_t kb = F.rp[k];
          const int kd = F.o_diag[k];
          const int s = (t - t0) & rmask;
          mq[s] = kb + kd + 1;
          mn[s] = (int)(F.rp[k + 1] - kb - kd - 1);
          mu[s] = F.v[kb + kd];
        }
      }
    };
    load_window(0);
    load_window(1);
    if (!from_factor)
      for (int64_t p = srp[r] + tid; p < srp[r + 1]; p += NT) {
        const int32_t c = sci[p];
        if (c >= col_limit) continue;
        const int pos = num_find(tab, lc, mask, lg, c);
        if (pos >= 0) w[pos] = sv[p];
      }
    lds_barrier();
    int loaded = 1;  // last window in the ring
    // three register sets in rotation: the U part of L column t + 3 loads
    // while columns t + 1 and t + 2 are applied
    int32_t ja[M], jb[M], jc[M];
    double va[M], vb[M], vc[M];
    // (every load is issued, from entry 0 when out of range, so the number
    // of loads in flight is the same on every path and the waits before a
    // set's use leave the two younger sets in flight)
    auto fetch = [&](int t, int32_t* j, double* v) {
      const bool ok = t < nl;
      const int s = ((ok ? t : t0) - t0) & rmask;
      const int64_t q = mq[s];
      const int n = ok ? mn[s] : 0;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int idx = tid + m * NT;
        const int64_t a = idx < n ? q + idx : 0;
        j[m] = F.ci[a];
        v[m] = F.v[a];
      }
    };
    // w_j -= mult u_kj for M entries of row k's U part, first probes of all
    // M lookups together, then the rare collisions; entries out of range or
    // not stored update the dump slot, so the reads and writes of w issue
    // back to back
    auto update = [&](double mult, int n, int off, const int32_t* j, const double* v) {
      uint32_t h[M], e[M];
#pragma unroll
      for (int m = 0; m < M; ++m) {
        h[m] = num_hash(j[m]) >> (32 - lg);
        e[m] = tab[h[m]];
      }
      int pos[M];
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int p = num_probe(tab, lc, mask, j[m], num_hash(j[m]) & 0xFFFFu, h[m], e[m]);
        pos[m] = off + tid + m * NT < n && p >= 0 ? p : max_len;
      }
      double o[M];
#pragma unroll
      for (int m = 0; m < M; ++m) o[m] = w[pos[m]];
#pragma unroll
      for (int m = 0; m < M; ++m) w[pos[m]] = o[m] - mult * v[m];
    };
    auto apply = [&](int t, const int32_t* j, const double* v) {
      // w[t] is final: later L columns only update positions > t
      const int s = (t - t0) & rmask;
      const double mult = w[t] / mu[s];
      const int n = mn[s];
      if (tid == 0) F.v[base + t] = mult;  // l_it is final (no row reads row r's L part)
      update(mult, n, 0, j, v);
      if (n > M * NT) {  // U parts longer than M NT: further rounds of M entries
        const int64_t q = mq[s];
        for (int off = M * NT; off < n; off += M * NT) {
          int32_t jt[M];
          double vt[M];
#pragma unroll
          for (int m = 0; m < M; ++m) {
            const int idx = off + tid + m * NT;
            const int64_t a = idx < n ? q + idx : 0;
            jt[m] = F.ci[a];
            vt[m] = F.v[a];
          }
          update(mult, n, off, jt, vt);
        }
      }
    };
    // the window after the current one, once every column still to be
    // applied or fetched lies in the current window or later (win >= 64, so
    // the fetches up to t + 5 never reach the window being loaded)
    auto refill = [&](int t) {
      if (t - t0 >= loaded * win && (loaded + 1) * win < nl - t0) {
        ++loaded;
        load_window(loaded);
      }
    };
    fetch(t0, ja, va);
    fetch(t0 + 1, jb, vb);
    fetch(t0 + 2, jc, vc);
    int t = t0;
    for (; t + 2 < nl; t += 3) {  // (no exit inside the body: every path
                                  // into the loop has the same loads in flight)
      refill(t);
      apply(t, ja, va);
      fetch(t + 3, ja, va);
      lds_barrier();
      apply(t + 1, jb, vb);
      fetch(t + 4, jb, vb);
      lds_barrier();
      apply(t + 2, jc, vc);
      fetch(t + 5, jc, vc);
      lds_barrier();
    }
    if (t < nl) {
      apply(t, ja, va);
      lds_barrier();
    }
    if (t + 1 < nl) {
      apply(t + 1, jb, vb);
      lds_barrier();
    }
    for (int t2 = nl + tid; t2 < len; t2 += NT) F.v[base + t2] = w[t2];
    if (tid == 0) {
      const double d = w[nl];
      if (!(d != 0.0) || !isfinite(d)) atomicExch(err, 1);
    }
    __syncthreads();  // the row's values are read by the block's later rows
  }
}

// Left-looking ILU by stages. Every external L column of a row (a column in
// another block) lies in an earlier stage, and in IKJ order a row applies its
// L columns in ascending order: all external ones, stage by stage, then the
// in-block ones. So once every earlier stage is final, the rows of a stage
// can apply all their external L columns at once, one workgroup per row
// group, their values kept in the factor; the per-block kernel above
// (from_factor) then applies the in-block columns and forms the pivots. Each
// entry receives the same updates in the same order as in the row-by-row
// IKJ loop.
//
// Source values first: the factor zeroed, then every entry of the source
// row (columns < col_limit) stored at its position (binary search in the
// factor row). One wave per row.
static __global__ __launch_bounds__(256) void k_ilu_source_scatter(IluView<double> F, int64_t n,
                                                            const int64_t* __restrict__ srp,
                                                            const int32_t* __restrict__ sci,
                                                            const double* __restrict__ sv, int64_t col_limit) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;
  const int64_t base = F.rp[i];
  const int len = (int)(F.rp[i + 1] - base);
  for (int64_t p = srp[i] + lane; p < srp[i + 1]; p += 64) {
    const int32_t c = sci[p];
    if (c >= col_limit) continue;
    int lo = 0, hi = len;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (F.ci[base + mid] < c) lo = mid + 1;
      else hi = mid;
    }
    if (lo < len && F.ci[base + lo] == c) F.v[base + lo] = sv[p];
  }
}

// first position p in [0, n) with a[p] >= key (a ascending), by the 64 lanes
// of a wave: probes spread over the range narrow it to 64 entries, then one
// exact round; every lane returns the same value
__device__ __forceinline__ int wave_lower_bound(const int32_t* __restrict__ a, int n, int32_t key, int lane) {
  int lo = 0, hi = n;
  while (hi - lo > 64) {
    const int step = (hi - lo + 63) >> 6;
    const int p = lo + lane * step;
    const bool lt = p < hi && a[p] < key;
    const int cnt = __popcll(__ballot(lt));
    if (cnt == 0) {
      hi = lo;
      break;
    }
    const int nlo = lo + (cnt - 1) * step + 1;
    hi = min(hi, lo + cnt * step);
    lo = nlo;
  }
  const int p = lo + lane;
  const bool lt = p < hi && a[p] < key;
  return lo + __popcll(__ballot(lt));
}

// Node-keyed hash map of the stage updates. A 3D velocity row names the
// columns of a lattice node (its three components 3q, 3q+1, 3q+2) together,
// so the map is keyed by the node q = column / 3, with a third of the keys of
// a column-keyed map and shorter probe chains at the same size (a step waits
// for the longest chain among its threads). Entry: a 15-bit tag above the
// node's present components (3 bits) above the position of its first
// present column (13 bits: rows of at most 8192 entries); the present
// columns of a node are consecutive in the sorted row. Empty slots are all
// ones (a valid entry has its top bit clear).
__device__ __forceinline__ uint32_t node_of(int32_t c) { return (uint32_t)c / 3u; }
__device__ __forceinline__ int node_probe(const uint32_t* tab, const int32_t* lc, uint32_t mask, uint32_t q,
                                          uint32_t comp, uint32_t tag, uint32_t h, uint32_t e) {
  while (e != NUM_EMPTY) {
    if ((e >> 16) == tag) {
      const int base = (int)(e & 0x1FFFu);
      if (node_of(lc[base]) == q) {
        const uint32_t pm = (e >> 13) & 7u;
        return ((pm >> comp) & 1u) ? base + __popc(pm & ((1u << comp) - 1u)) : -1;
      }
    }
    h = (h + 1) & mask;
    e = tab[h];
  }
  return -1;
}

// LDS of one stage-update workgroup: the G rows' working values from the
// first updated column on (and a dump slot each), a batch of 64 L columns'
// U spans, pivots, first two U values, node-triple flags, the steps with
// their streams' starts and lengths, the columns and the hash map
constexpr int UPD_BATCH = 64;
__host__ __device__ inline size_t ilu_update_lds_bytes(int mcap, int hs, int G) {
  return (size_t)G * (mcap + 1) * 8 + (size_t)UPD_BATCH * (8 + 8 + 8 + 8 + 24 + 4 + 4 + 4 + 4) + 16 +
         (size_t)mcap * 4 + (size_t)hs * 4;
}

// One stage update: the row groups starting at rows[0 .. nrows) (G
// consecutive rows of one block with the same columns, e.g. the velocity
// components of one node; ascending, each row of at most mcap entries) apply
// their L columns in [col_a, col_b), rows of finished stages. One workgroup
// of NT threads per group, workgroups XCD-contiguous so that neighbouring
// rows, which name the same finished rows, share an L2. The rows from their first such column on go to
// LDS (values per row, the shared columns and hash map); then the L columns
// in order, in batches of 64 whose U spans, pivots and first U values are
// read first. A step is one L column k, or three, k, k+1, k+2, when those
// rows start a node triple (tri[k]: three consecutive rows of one block with
// the same columns, so row k's U part starts with k+1, k+2 and row k+1's
// with k+2, and past k+2 all three U parts name the same columns): the
// step's multipliers m = w_t (1 / u_kk) (for a triple: with the in-node updates
// of w_{t+1}, w_{t+2} first) go to the factor, then one stream of the U
// part(s) (chunks of NT M entries, loaded three chunks ahead, each column
// looked up once for the G rows and the step's 1 or 3 source rows) updates
// w_j -= m_k u_kj (-= m_k+1 u_k+1,j -= m_k+2 u_k+2,j) where the rows store j.
// Every w_j receives the same fused multiply-adds in the same order as in
// the column-by-column IKJ loop; a multiplier is w_t times the pivot's
// reciprocal (taken once per column and batch), within an ulp of w_t / u_kk.
// Then the rest goes back to the factor.
// The row group i .. i + G - 1 (same columns) from position pa of its rows
// (m positions) in LDS, with the node-keyed hash of its columns; then its L
// columns at positions [0, nc) applied in ascending order in node-triple
// steps (sources: finished rows). The values stay in LDS (w, ld = mcap + 1)
// for the caller; the multipliers of positions [0, nc) are already stored.
template <int NT, int M, int G>
__device__ __forceinline__ void ilu_group_columns(const IluView<double>& F, const uint8_t* __restrict__ tri,
                                                  int32_t i, int pa, int nc, int m, int mcap, int hs, char* smem) {
  double* w = (double*)smem;                     // [G][mcap + 1]; w[g][mcap]: dump slot
  int64_t* mq = (int64_t*)(w + G * (mcap + 1));  // [UPD_BATCH] first U entry of row k
  double* mu = (double*)(mq + UPD_BATCH);         // [UPD_BATCH] 1 / u_kk
  double* mf = mu + UPD_BATCH;                    // [UPD_BATCH] first U value of row k
  double* ms = mf + UPD_BATCH;                    // [UPD_BATCH] second U value of row k
  int64_t* sq = (int64_t*)(ms + UPD_BATCH);       // [3][UPD_BATCH] per step: its three U streams
  int32_t* mn = (int32_t*)(sq + 3 * UPD_BATCH);   // [UPD_BATCH] U entries of row k
  int32_t* tf = mn + UPD_BATCH;                   // [UPD_BATCH] slot starts a node triple
  int32_t* stp = tf + UPD_BATCH;                  // [UPD_BATCH] steps: slot | width << 8
  int32_t* sn = stp + UPD_BATCH;                  // [UPD_BATCH] per step: its stream's length
  int32_t* nsteps = sn + UPD_BATCH;               // [4] the batch's step count
  int32_t* lc = nsteps + 4;                       // [mcap] columns
  uint32_t* tab = (uint32_t*)(lc + mcap);         // [hs] hash map
  const uint32_t mask = (uint32_t)hs - 1;
  const int lg = __builtin_ctz((uint32_t)hs);
  const int tid = threadIdx.x, lane = tid & 63;
  const int ld = mcap + 1;
  const int64_t base = F.rp[i];
  int64_t bg[G];  // the group's rows: row i + g starts at bg[g] + pa
#pragma unroll
  for (int g = 0; g < G; ++g) bg[g] = F.rp[i + g] + pa;
  for (int h = tid; h < hs; h += NT) tab[h] = NUM_EMPTY;
  for (int u = tid; u < m; u += NT) {
    lc[u] = F.ci[base + pa + u];
#pragma unroll
    for (int g = 0; g < G; ++g) w[g * ld + u] = F.v[bg[g] + u];
  }
  lds_barrier();
  // one entry per node, from the node's first present column
  for (int u = tid; u < m; u += NT) {
    const int32_t c = lc[u];
    const uint32_t q = node_of(c);
    if (u > 0 && node_of(lc[u - 1]) == q) continue;
    uint32_t pm = 1u << ((uint32_t)c - 3u * q);
    if (u + 1 < m && node_of(lc[u + 1]) == q) pm |= 1u << ((uint32_t)lc[u + 1] - 3u * q);
    if (u + 2 < m && node_of(lc[u + 2]) == q) pm |= 1u << ((uint32_t)lc[u + 2] - 3u * q);
    const uint32_t hh = num_hash((int32_t)q);
    const uint32_t ent = ((hh & 0x7FFFu) << 16) | (pm << 13) | (uint32_t)u;
    for (uint32_t h = hh >> (32 - lg);; h = (h + 1) & mask)
      if (atomicCAS(&tab[h], NUM_EMPTY, ent) == NUM_EMPTY) break;
  }
  lds_barrier();
  int32_t ja[M], jb[M], jc[M];
  double va[3][M], vb[3][M], vc[3][M];
  for (int b0 = 0; b0 < nc; b0 += UPD_BATCH) {
    const int nb = min(UPD_BATCH, nc - b0);
    if (tid < nb) {
      const int32_t k = lc[b0 + tid];
      const int64_t kb = F.rp[k];
      const int kd = F.o_diag[k];
      const int nu_k = (int)(F.rp[k + 1] - kb - kd - 1);
      const int64_t q = kb + kd + 1;
      mq[tid] = q;
      mn[tid] = nu_k;
      mu[tid] = 1.0 / F.v[kb + kd];  // the pivot's reciprocal, taken once per column
      mf[tid] = F.v[nu_k > 0 ? q : kb + kd];
      ms[tid] = F.v[nu_k > 1 ? q + 1 : kb + kd];
      tf[tid] = tri[k] != 0 && tid + 2 < nb && lc[b0 + tid + 1] == k + 1 && lc[b0 + tid + 2] == k + 2;
    }
    lds_barrier();
    if (tid < 64) {  // wave 0: the batch's steps, in slot order
      const bool st0 = lane < nb && tf[lane] != 0;
      const bool mem = (lane >= 1 && lane - 1 < nb && tf[lane - 1] != 0) || (lane >= 2 && lane - 2 < nb && tf[lane - 2] != 0);
      const bool is_step = lane < nb && !mem;
      const uint64_t bal = __ballot(is_step);
      if (is_step) {
        // the step's record: its slot and width, the starts of the three
        // streams the chunks load (a single column reads row k's U part three
        // times, the extra values unused; a triple reads row k past k+1, k+2,
        // row k+1 past k+2 and row k+2's U part, which name the same columns)
        // and the stream's length
        const int c = __popcll(bal & ((1ull << lane) - 1ull));
        stp[c] = lane | ((st0 ? 3 : 1) << 8);
        sq[c] = st0 ? mq[lane] + 2 : mq[lane];
        sq[UPD_BATCH + c] = st0 ? mq[lane + 1] + 1 : mq[lane];
        sq[2 * UPD_BATCH + c] = st0 ? mq[lane + 2] : mq[lane];
        sn[c] = st0 ? mn[lane + 2] : mn[lane];
      }
      if (lane == 0) nsteps[0] = __popcll(bal);
    }
    lds_barrier();
    const int ns = nsteps[0];
    // The batch's steps as one stream of chunks of NT M entries, chunk
    // (c, off) = entries [off, off + NT M) of step c's U stream (a step with
    // an empty U stream still has its chunk: its multipliers). Loads run
    // three chunks ahead across step boundaries; a barrier precedes the
    // first chunk of every step (its pivot entries must have all earlier
    // steps' updates). Positions past the batch load from entry 0; every
    // chunk issues the loads of a triple, so the loads in flight are the same
    // on every path.
    constexpr int CHUNK = NT * M;
    // chunk p -> the next chunk; ln as issue() returned it for p: the
    // stream's length (low 16 bits, at most 8192) and the step's slot and
    // width (high 16 bits)
    auto adv = [&](int2 p, int ln) -> int2 {
      if (p.x >= ns) return p;
      p.y += CHUNK;
      if (p.y >= (ln & 0xffff)) {
        p.x += 1;
        p.y = 0;
      }
      return p;
    };
    auto issue = [&](int2 p, int32_t* j, double (*v)[M]) -> int {
      // the step's record (past the batch: step 0's, with no entries)
      const bool ok = p.x < ns;
      const int c = ok ? p.x : 0;
      const int64_t q0 = sq[c], q1 = sq[UPD_BATCH + c], q2 = sq[2 * UPD_BATCH + c];
      const int n = ok ? sn[c] : 0;
      const int sd = stp[c];
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int idx = p.y + tid + r * NT;
        const bool in = idx < n;
        j[r] = F.ci[in ? q2 + idx : 0];
        v[0][r] = F.v[in ? q0 + idx : 0];
        v[1][r] = F.v[in ? q1 + idx : 0];
        v[2][r] = F.v[in ? q2 + idx : 0];
      }
      return n | (sd << 16);
    };
    double m0[G], m1[G], m2[G];
#pragma unroll
    for (int g = 0; g < G; ++g) m0[g] = m1[g] = m2[g] = 0.0;
    auto process = [&](int2 p, const int32_t* j, double (*v)[M], int ln) {
      const int n = ln & 0xffff, sd = ln >> 16;
      const int sl = sd & 0xff;
      const bool t3 = (sd >> 8) == 3;
      if (p.y == 0) {  // a new step: every earlier step's updates are in
        lds_barrier();
        const int t = b0 + sl;
        if (t3) {
          const double u01 = mf[sl], u02 = ms[sl], u12 = mf[sl + 1];
          const double r0 = mu[sl], r1 = mu[sl + 1], r2 = mu[sl + 2];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            m0[g] = w[g * ld + t] * r0;
            m1[g] = fma(-m0[g], u01, w[g * ld + t + 1]) * r1;
            m2[g] = fma(-m1[g], u12, fma(-m0[g], u02, w[g * ld + t + 2])) * r2;
          }
          if (tid == 0) {  // one branch: the step's multipliers go to the factor
#pragma unroll
            for (int g = 0; g < G; ++g) {
              F.v[bg[g] + t] = m0[g];
              F.v[bg[g] + t + 1] = m1[g];
              F.v[bg[g] + t + 2] = m2[g];
            }
          }
        } else {
          const double rp = mu[sl];
#pragma unroll
          for (int g = 0; g < G; ++g) m0[g] = w[g * ld + t] * rp;
          if (tid == 0) {
#pragma unroll
            for (int g = 0; g < G; ++g) F.v[bg[g] + t] = m0[g];
          }
        }
      }
      uint32_t q[M], h[M], e[M];
#pragma unroll
      for (int r = 0; r < M; ++r) {
        q[r] = node_of(j[r]);
        h[r] = num_hash((int32_t)q[r]) >> (32 - lg);
        e[r] = tab[h[r]];
      }
      int pos[M];
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int pp = node_probe(tab, lc, mask, q[r], (uint32_t)j[r] - 3u * q[r],
                                  num_hash((int32_t)q[r]) & 0x7FFFu, h[r], e[r]);
        pos[r] = p.y + tid + r * NT < n && pp >= 0 ? pp : mcap;
      }
      if (t3) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          double o[M];
#pragma unroll
          for (int r = 0; r < M; ++r) o[r] = w[g * ld + pos[r]];
#pragma unroll
          for (int r = 0; r < M; ++r)
            w[g * ld + pos[r]] = fma(-m2[g], v[2][r], fma(-m1[g], v[1][r], fma(-m0[g], v[0][r], o[r])));
        }
      } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          double o[M];
#pragma unroll
          for (int r = 0; r < M; ++r) o[r] = w[g * ld + pos[r]];
#pragma unroll
          for (int r = 0; r < M; ++r) w[g * ld + pos[r]] = fma(-m0[g], v[0][r], o[r]);
        }
      }
    };
    int2 p0 = make_int2(0, 0);
    int na = issue(p0, ja, va);
    int2 p1 = adv(p0, na);
    int nb1 = issue(p1, jb, vb);
    int2 p2 = adv(p1, nb1);
    int nc1 = issue(p2, jc, vc);
    // One exit, at the loop test: every pass issues the three chunks' loads
    // (past the batch, loads of entry 0), so the loads in flight when a
    // chunk is used are the same on every path into it and the compiler's
    // memory-counter waits keep the other two chunks in flight (with early
    // exits in the body, a shared exit block made them wait for all loads).
    // Positions only advance, so a chunk past the batch is followed by
    // chunks past it.
    while (p0.x < ns) {
      process(p0, ja, va, na);
      const int2 p3 = adv(p2, nc1);
      na = issue(p3, ja, va);
      if (p1.x < ns) process(p1, jb, vb, nb1);
      const int2 p4 = adv(p3, na);
      nb1 = issue(p4, jb, vb);
      if (p2.x < ns) process(p2, jc, vc, nc1);
      const int2 p5 = adv(p4, nb1);
      nc1 = issue(p5, jc, vc);
      p0 = p3;
      p1 = p4;
      p2 = p5;
    }
    lds_barrier();  // the batch's last updates, before its metadata is replaced
  }
}

template <int NT, int M, int G>
__global__ __launch_bounds__(NT) void k_ilu_numeric_update(IluView<double> F, const int32_t* __restrict__ rows,
                                                           const uint8_t* __restrict__ tri, int nrows,
                                                           int32_t col_a, int32_t col_b, int mcap, int hs) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int item = xcd_contiguous(blockIdx.x, gridDim.x);
  if (item >= nrows) return;
  const int32_t i = rows[item];
  const int64_t base = F.rp[i];
  const int len = (int)(F.rp[i + 1] - base);
  const int nx = F.o_lint[i];
  // (every wave finds the same range)
  const int pa = wave_lower_bound(F.ci + base, nx, col_a, lane);
  const int pb = wave_lower_bound(F.ci + base, nx, col_b, lane);
  if (pa == pb) return;  // the whole workgroup: no barrier is passed
  const int m = len - pa;  // positions [pa, len): every column a U part of stage rows can name
  const int nc = pb - pa;
  ilu_group_columns<NT, M, G>(F, tri, i, pa, nc, m, mcap, hs, smem);
  const double* w = (const double*)smem;
  const int ld = mcap + 1;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t bg = F.rp[i + g] + pa;
    for (int u = nc + tid; u < m; u += NT) F.v[bg + u] = w[g * ld + u];
  }
}

// In-block columns and pivots of a staged factorisation (its external
// columns already applied): one workgroup per block of the stage, its row
// groups in order, each a node triple (tri) or a single row. A group applies
// its L columns of earlier rows of the block (ilu_group_columns, the sources
// finished by this workgroup), then for a triple the in-node columns: row
// r+1 eliminates column r, row r+2 columns r and r+1, each multiplier the
// quotient by the pivot and each update one fused multiply-add per entry of
// the pivot row's U part in LDS (the same order as the IKJ loop); then the
// rows go to the factor, pivots checked. Without TRI (a stage whose three
// rows do not fit in LDS together) every row is its own group: a triple's
// later rows then take its in-node columns as ordinary L columns of finished
// rows, in ascending order as in the IKJ loop.
template <int NT, int M, bool TRI>
__global__ __launch_bounds__(NT) void k_ilu_numeric_block(IluView<double> F, const uint8_t* __restrict__ tri,
                                                          int blk0, int blk1, int mcap, int hs,
                                                          int* __restrict__ err) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int tid = threadIdx.x;
  const int blk = blk0 + blockIdx.x;
  if (blk >= blk1) return;
  const int r0 = F.blk_ptr[blk], r1 = F.blk_ptr[blk + 1];
  const int ld = mcap + 1;
  double* w = (double*)smem;
  for (int r = r0; r < r1;) {
    const int G = TRI && tri[r] ? 3 : 1;
    const int64_t base = F.rp[r];
    const int len = (int)(F.rp[r + 1] - base);
    const int pa = F.o_lint[r];          // the first in-block column
    const int nc = F.o_diag[r] - pa;     // the row's in-block L columns (rows r0 .. r-1)
    const int m = len - pa;
    if (G == 3) ilu_group_columns<NT, M, 3>(F, tri, r, pa, nc, m, mcap, hs, smem);
    else ilu_group_columns<NT, M, 1>(F, tri, r, pa, nc, m, mcap, hs, smem);
    if (G == 3) {
      // rows r+1 and r+2 name columns r (position nc) and r + 1 (nc + 1)
      const int d = nc;
      const double m1 = w[ld + d] / w[d];
      const double m2a = w[2 * ld + d] / w[d];
      for (int u = d + 1 + tid; u < m; u += NT) {
        const double x0 = w[u];
        w[ld + u] = fma(-m1, x0, w[ld + u]);
        w[2 * ld + u] = fma(-m2a, x0, w[2 * ld + u]);
      }
      lds_barrier();
      const double m2b = w[2 * ld + d + 1] / w[ld + d + 1];
      for (int u = d + 2 + tid; u < m; u += NT) w[2 * ld + u] = fma(-m2b, w[ld + u], w[2 * ld + u]);
      lds_barrier();
      if (tid == 0) {
        w[ld + d] = m1;
        w[2 * ld + d] = m2a;
        w[2 * ld + d + 1] = m2b;
      }
      lds_barrier();
    }
    for (int g = 0; g < G; ++g) {
      const int64_t bg = F.rp[r + g] + pa;
      for (int u = nc + tid; u < m; u += NT) F.v[bg + u] = w[g * ld + u];
      if (tid == 0) {
        const double piv = w[g * ld + nc + g];
        if (!(piv != 0.0) || !isfinite(piv)) atomicExch(err, 1);
      }
    }
    __threadfence_block();
    __syncthreads();  // the group's rows are sources of the block's later rows
    r += G;
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
  BlockLds<T> L = carve_lds<T>(smem, a.max_rows, a.cap);
  unsigned int gen = __hip_atomic_load(&a.gb->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int nwg = gridDim.x;
  const int64_t tid = (int64_t)blockIdx.x * BLOCK + threadIdx.x, nthr = (int64_t)nwg * BLOCK;
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
      for (int blk = a.stage_ptr[s] + blockIdx.x; blk < a.stage_ptr[s + 1]; blk += nwg) {
        ilu_fwd_block<T, T, T, GF, true, BLOCK>(a.F, blk, src, dst, L, a.cap, threadIdx.x);
        __syncthreads();
      }
      ok = grid_sync(a.gb, gen, a.timeout);
    }
    for (int s = a.n_stages - 1; s >= 0 && ok; --s) {
      for (int blk = a.stage_ptr[s] + blockIdx.x; blk < a.stage_ptr[s + 1]; blk += nwg) {
        ilu_bwd_block<T, T, T, GF, false, BLOCK>(a.F, blk, dst, dst, L, a.cap, threadIdx.x);
        __syncthreads();
      }
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

// q = A p; partial[b] = sum over the workgroup's rows of p_i q_i
template <typename T, int G>
__global__ __launch_bounds__(BLOCK) void k_spmv_dot(int64_t nrows, const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const T* __restrict__ val,
                                                    const T* __restrict__ p, T* __restrict__ q,
                                                    double* __restrict__ partial) {
  __shared__ double sh[BLOCK / WAVE];
  const int lane = threadIdx.x % G;
  const int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G;
  double contrib = 0;
  if (row < nrows) {
    const T acc = row_dot<T, T, T, G, false>(ci, val, p, rp[row], rp[row + 1], lane, 0, 0);
    if (lane == 0) {
      q[row] = acc;
      contrib = (double)acc * (double)p[row];
    }
  }
  contrib = block_sum(contrib, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = contrib;
}

// sum of nb partials, identical in every workgroup
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nb, double* sh) {
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  s = block_sum(s, sh);
  __shared__ double tot;
  if (threadIdx.x == 0) tot = s;
  __syncthreads();
  return tot;
}

// ---------------------------------------------------------------------------
// The inner CG on M_p capped at one step (k_inner_cg's arithmetic with
// max_it = 1) as ordinary launches, without grid barriers: from x = 0 the
// residual is g = -b and the first direction d = -ILU_S^{-1} g = ILU_S^{-1} b
// (the factor's stage kernels), then
//   k_cg_step_spmv: q = M d and per-workgroup partials of d.q, b.d, b.b
//   k_cg_step_x:    alpha = b.d / d.q (= g.h / d.q), x = alpha d and partials
//                   of |g + alpha q|^2 with g = -b
//   k_cg_step_end:  the step count and the convergence flag
// Each kernel sums the previous one's partials in the same fixed order in
// every workgroup. At most CG_STEP_WG workgroups per kernel (grid-stride).
constexpr int CG_STEP_WG = 512;

template <typename T, int G>
__global__ __launch_bounds__(BLOCK) void k_cg_step_spmv(int64_t n, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci, const T* __restrict__ val,
                                                        const T* __restrict__ b, const T* __restrict__ d,
                                                        T* __restrict__ q, double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  const int lane = threadIdx.x % G;
  double dq = 0, bd = 0, bb = 0;
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < n;
       row += (int64_t)gridDim.x * (BLOCK / G)) {
    const T acc = row_dot<T, T, T, G, false>(ci, val, d, rp[row], rp[row + 1], lane, 0, 0);
    if (lane == 0) {
      q[row] = acc;
      const double dr = (double)d[row], br = (double)b[row];
      dq += (double)acc * dr;
      bd += br * dr;
      bb += br * br;
    }
  }
  dq = block_sum(dq, sh);
  bd = block_sum(bd, sh);
  bb = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = dq;
    part[CG_STEP_WG + blockIdx.x] = bd;
    part[2 * CG_STEP_WG + blockIdx.x] = bb;
  }
}

// the step's start: |b| and whether the CG stops before its first step
// (res <= reduce * res, i.e. b = 0 for reduce < 1, or res <= tol)
__device__ __forceinline__ bool cg_step_done0(const double* part, int nb1, double tol, double reduce, double* sh,
                                              double& res0) {
  res0 = sqrt(sum_partials(part + 2 * CG_STEP_WG, nb1, sh));
  return res0 <= res0 * reduce || res0 <= tol;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_cg_step_x(int64_t n, int nb1, const T* __restrict__ b,
                                                     const T* __restrict__ d, const T* __restrict__ q,
                                                     T* __restrict__ x, double* __restrict__ part, double tol,
                                                     double reduce) {
  __shared__ double sh[BLOCK / WAVE];
  double res0;
  const bool done0 = cg_step_done0(part, nb1, tol, reduce, sh, res0);
  const double dq = sum_partials(part, nb1, sh);
  const double bd = sum_partials(part + CG_STEP_WG, nb1, sh);
  const double alpha = bd / dq;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    if (done0) {
      x[i] = 0;
    } else {
      x[i] = (T)alpha * d[i];
      const T gi = -b[i] + (T)alpha * q[i];
      s += (double)gi * (double)gi;
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[3 * CG_STEP_WG + blockIdx.x] = s;
}

// one workgroup: iterations += 1 unless the CG stopped before its step; the
// failure flag when the step did not reach the tolerance
static __global__ __launch_bounds__(BLOCK) void k_cg_step_end(int nb1, int nb2, const double* __restrict__ part,
                                                              double tol, double reduce, int* __restrict__ iters,
                                                              int* __restrict__ fail) {
  __shared__ double sh[BLOCK / WAVE];
  double res0;
  const bool done0 = cg_step_done0(part, nb1, tol, reduce, sh, res0);
  const double res = sqrt(sum_partials(part + 3 * CG_STEP_WG, nb2, sh));
  if (threadIdx.x == 0 && !done0) {
    iters[0] += 1;
    if (!(res <= res0 * reduce || res <= tol)) fail[0] = 1;
  }
}

// partial[b] = sum a_i b_i
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

// alpha = rz / pq; x += alpha p (x = alpha p on the first step); r -= alpha q
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_pcg_update(int64_t n, const double* __restrict__ pq,
                                                      const double* __restrict__ rz, const T* __restrict__ p,
                                                      const T* __restrict__ q, T* __restrict__ x,
                                                      T* __restrict__ r, int first) {
  const double a = pq[0] != 0.0 ? rz[0] / pq[0] : 0.0;
  const T alpha = (T)a;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const T pi = p[i];
    x[i] = first ? alpha * pi : x[i] + alpha * pi;
    r[i] -= alpha * q[i];
  }
}

// rz_new = sum(rz partials); beta = rz_new / rz_old (0 on the first step);
// p = z + beta p; workgroup 0 stores rz_new (another slot than rz_old)
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

// The inner PCG's update step in Chronopoulos-Gear form (the per-row
// arithmetic the forward ILU stages apply when they carry it): p = z + beta p,
// x += alpha p (p = z, x = alpha p when first); with r: also q = s + beta q,
// r -= alpha q. Without r only x is formed (p is not stored).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_cg1_update(int64_t n, const double* __restrict__ alpha_p,
                                                      const double* __restrict__ beta_p, const T* __restrict__ z,
                                                      const T* __restrict__ s, T* __restrict__ p, T* __restrict__ q,
                                                      T* __restrict__ x, T* __restrict__ r, int first) {
  const T alpha = (T)alpha_p[0], beta = (T)beta_p[0];
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const T zi = z[i];
    const T pi = first ? zi : zi + beta * p[i];
    x[i] = first ? alpha * pi : x[i] + alpha * pi;
    if (r) {
      const T si = s[i];
      const T qi = first ? si : si + beta * q[i];
      p[i] = pi;
      q[i] = qi;
      r[i] -= alpha * qi;
    }
  }
}

// The scalars of one Chronopoulos-Gear CG step (one workgroup): gamma =
// sum(rz_part) = r.z, delta = sum(zs_part) = z.A z;
//   first: beta = 0, alpha = gamma / delta
//   else:  beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
// sc[0] = alpha, sc[1] = beta, sc[2] = gamma (read as alpha_old, gamma_old).
static __global__ __launch_bounds__(1024) void k_cg1_scalars(int nr, const double* __restrict__ rz_part, int nd,
                                                       const double* __restrict__ zs_part, double* __restrict__ sc,
                                                       int first) {
  __shared__ double sh[1024 / WAVE];
  // thread t sums entries t, t + 1024, ... of each array in that order, the
  // loads of 16 such entries of both arrays in flight together (unconditional
  // loads of a clamped index, nr, nd >= 1; entries past the end count 0)
  constexpr int NT = 1024, U = 16;
  double a = 0, b = 0;
  const int m = nr > nd ? nr : nd;
  for (int i = threadIdx.x; i < m; i += U * NT) {
    double v[U], w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = i + u * NT;
      v[u] = rz_part[j < nr ? j : nr - 1];
      w[u] = zs_part[j < nd ? j : nd - 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = i + u * NT;
      a += j < nr ? v[u] : 0.0;
      b += j < nd ? w[u] : 0.0;
    }
  }
  const double gamma = block_sum(a, sh);
  const double delta = block_sum(b, sh);
  if (threadIdx.x == 0) {
    const double a_old = sc[0], g_old = sc[2];
    double beta = 0, den = delta;
    if (!first && g_old != 0.0 && a_old != 0.0) {
      beta = gamma / g_old;
      den = delta - beta * gamma / a_old;
    }
    sc[0] = den != 0.0 ? gamma / den : 0.0;
    sc[1] = beta;
    sc[2] = gamma;
  }
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
