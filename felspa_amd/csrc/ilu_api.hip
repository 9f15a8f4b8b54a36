// Standalone SparseILU and SparseMatrix objects of include/fs_stokes.h.
//   fs_ilu_*    include/felspa/linear_algebra/ilu.h:25-238 (initialize, vmult,
//               Tvmult) on the GPU, caller's row order at the boundary
//   fs_matrix_* dealii::SparseMatrix::vmult (CSR SpMV) on the GPU
//   fs_ilu_symbolic  the host symbolic phase alone
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

#include "fs_stokes.h"
#include "solver_internal.h"

using namespace fsk;
using namespace fs;

struct fs_ilu {
  int k = 0;
  int device = 0;  // the GPU the factor lives on
  int64_t n = 0;
  hipStream_t st = nullptr;
  IluDev F;
  IluDev FT;                     // transposed factor, built on the first Tvmult
  std::vector<int64_t> tmap;     // F position -> FT position
  DevBuf tmap_d;
  bool ft_ready = false;
  DevCsr A;                      // source matrix in the factor's order
  std::vector<int64_t> amap;     // caller entry -> A position
  std::vector<double> hA;
  DevBuf perm, va, vb, vc, err;
  bool factored = false;
  Counters cnt;
  ~fs_ilu() {
    if (st) (void)hipStreamDestroy(st);
  }
  void sync() { hip_check(hipStreamSynchronize(st), "sync"); }
};

struct fs_matrix {
  DevCsr M;
  hipStream_t st = nullptr;
  DevBuf x, y;
  ~fs_matrix() {
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {
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

int grid1(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + BLOCK - 1) / BLOCK)); }
}  // namespace

extern "C" {

int fs_ilu_symbolic(const fs_csr* A, int32_t k, int64_t** row_ptr, int32_t** col_idx) {
  if (!row_ptr || !col_idx) { fs::set_error("fs_ilu_symbolic: null output"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  HostCsr h = csr_from_abi(A, "A", false);
  if (h.n_rows != h.n_cols) throw FsError(FS_ERR_BAD_INPUT, "matrix not square");
  if (k < 0) throw FsError(FS_ERR_BAD_INPUT, "negative level of fill");
  HostPattern F = symbolic_ilu(pattern_of(h), k, nullptr);
  *row_ptr = (int64_t*)std::malloc(8 * (F.n + 1));
  *col_idx = (int32_t*)std::malloc(4 * std::max<int64_t>(1, F.nnz()));
  if (!*row_ptr || !*col_idx) throw std::bad_alloc();
  std::memcpy(*row_ptr, F.rp.data(), 8 * (F.n + 1));
  std::memcpy(*col_idx, F.ci.data(), 4 * F.nnz());
  return FS_OK;
  FS_GUARD_END
}

int fs_block_color_ordering(const fs_csr* A, int32_t k, int32_t block_rows, const int32_t* block_hint,
                            int32_t* perm, int32_t* blk_ptr, int32_t* n_blocks, int32_t* stage_ptr,
                            int32_t* n_stages) {
  FS_GUARD_BEGIN
  HostCsr hA = csr_from_abi(A, "A", false);
  if (hA.n_rows != hA.n_cols) throw FsError(FS_ERR_BAD_INPUT, "matrix not square");
  if (k < 0 || k > 20) throw FsError(FS_ERR_BAD_INPUT, "level of fill out of range");
  Schedule s = schedule_block_color(pattern_of(hA), k, block_rows > 0 ? block_rows : 64, block_hint, 256);
  if (perm) std::copy(s.perm.begin(), s.perm.end(), perm);
  if (blk_ptr) std::copy(s.blk_ptr.begin(), s.blk_ptr.end(), blk_ptr);
  if (stage_ptr) std::copy(s.stage_ptr.begin(), s.stage_ptr.end(), stage_ptr);
  if (n_blocks) *n_blocks = s.n_blocks();
  if (n_stages) *n_stages = s.n_stages();
  return FS_OK;
  FS_GUARD_END
}

int fs_ilu_create(const fs_csr* A, int32_t k, int32_t ordering, int32_t block_rows, const int32_t* block_hint,
                  fs_ilu** out) {
  if (!out) { fs::set_error("fs_ilu_create: out is null"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  FS_GUARD_BEGIN
  ensure_device();
  std::unique_ptr<fs_ilu> h(new fs_ilu());
  HostCsr hA = csr_from_abi(A, "A", false);
  if (hA.n_rows != hA.n_cols) throw FsError(FS_ERR_BAD_INPUT, "matrix not square");
  if (k < 0 || k > 20) throw FsError(FS_ERR_BAD_INPUT, "level of fill out of range");
  h->k = k;
  h->n = hA.n_rows;
  hip_check(hipGetDevice(&h->device), "hipGetDevice");
  HostPattern pA = pattern_of(hA);
  Schedule s = ordering == FS_ORDER_REFERENCE
                 ? schedule_reference(h->n, 128)
                 : schedule_block_color(pA, k, block_rows > 0 ? block_rows : 64, block_hint, 256);
  std::vector<int64_t> map;
  HostCsr An = permute_square(hA, s.perm, s.iperm, &map);
  h->F.build(pattern_of(An), s, k, 8);
  h->A.upload_pattern(An);
  h->A.v64.alloc(8 * std::max<int64_t>(1, An.nnz()));
  std::vector<int64_t> c2s = caller_to_sorted(A);
  h->amap.resize(c2s.size());
  for (size_t e = 0; e < c2s.size(); ++e) h->amap[e] = map[c2s[e]];
  h->hA.assign(An.nnz(), 0.0);
  h->perm.upload(s.perm.data(), 4 * h->n);
  h->va.alloc(8 * h->n);
  h->vb.alloc(8 * h->n);
  h->vc.alloc(8 * h->n);
  h->err.alloc(16);
  hip_check(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking), "stream");
  *out = h.release();
  return FS_OK;
  FS_GUARD_END
}

int fs_ilu_factorize(fs_ilu* h, const double* A_val) {
  if (!h || !A_val) { fs::set_error("fs_ilu_factorize: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  for (size_t e = 0; e < h->amap.size(); ++e) h->hA[h->amap[e]] = A_val[e];
  hip_check(hipMemcpyAsync(h->A.v64.p, h->hA.data(), 8 * h->hA.size(), hipMemcpyHostToDevice, h->st), "H2D");
  hip_check(hipMemsetAsync(h->err.p, 0, 16, h->st), "memset");
  h->F.factorize(h->A.rp.as<int64_t>(), h->A.ci.as<int32_t>(), h->A.v64.as<double>(), h->n, h->st,
                 h->err.as<int>(), h->cnt);
  int e = 0;
  hip_check(hipMemcpyAsync(&e, h->err.p, 4, hipMemcpyDeviceToHost, h->st), "D2H");
  h->sync();
  h->ft_ready = false;
  h->factored = true;
  if (e) {
    fs::set_error("fs_ilu_factorize: zero or non-finite pivot");
    return FS_ERR_ZERO_PIVOT;
  }
  return FS_OK;
  FS_GUARD_END
}

static void build_transposed(fs_ilu* h) {
  if (h->FT.n == 0) {
    // pattern of F^T with the same schedule; tmap: F position -> F^T position
    const HostPattern& P = h->F.host_pattern;
    const int64_t n = P.n, nnz = P.nnz();
    HostPattern T;
    T.n = n;
    T.rp.assign(n + 1, 0);
    for (int64_t p = 0; p < nnz; ++p) T.rp[P.ci[p] + 1]++;
    for (int64_t r = 0; r < n; ++r) T.rp[r + 1] += T.rp[r];
    T.ci.resize(nnz);
    h->tmap.resize(nnz);
    std::vector<int64_t> pos(T.rp.begin(), T.rp.end() - 1);
    for (int64_t r = 0; r < n; ++r)
      for (int64_t p = P.rp[r]; p < P.rp[r + 1]; ++p) {
        const int64_t q = pos[P.ci[p]]++;
        T.ci[q] = (int32_t)r;
        h->tmap[p] = q;
      }
    // the transposed pattern is its own level-0 pattern: build with k = 0
    h->FT.build(T, h->F.sched, 0, 8);
    h->tmap_d.upload(h->tmap.data(), 8 * nnz);
  }
  hipLaunchKernelGGL(k_scatter_vals, dim3(grid1(h->F.nnz)), dim3(BLOCK), 0, h->st, h->F.nnz,
                     h->tmap_d.as<int64_t>(), h->F.v64.as<double>(), h->FT.v64.as<double>());
  if (!h->FT.v32.p) h->FT.v32.alloc(4 * h->FT.nnz);
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid1(h->FT.nnz)), dim3(BLOCK), 0, h->st, h->FT.nnz,
                     h->FT.v64.as<double>(), h->FT.v32.as<float>());
  h->FT.gather_inblock(h->st);
  h->ft_ready = true;
}

static int ilu_apply(fs_ilu* h, const double* src, double* dst, int32_t precision, bool transposed) {
  if (!h || !src || !dst) { fs::set_error("fs_ilu_vmult: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  if (!h->factored) throw FsError(FS_ERR_BAD_INPUT, "fs_ilu: factorize before vmult");
  const int64_t n = h->n;
  const int g = grid1(n);
  hip_check(hipMemcpyAsync(h->vc.p, src, 8 * n, hipMemcpyHostToDevice, h->st), "H2D");
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, h->perm.as<int32_t>(),
                     h->vc.as<double>(), h->va.as<double>(), 0);
  if (transposed && !h->ft_ready) build_transposed(h);
  if (precision == FS_PREC_MIXED) {
    float* a = (float*)h->vb.p;
    float* t = a + ((n + 63) & ~int64_t(63));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(BLOCK), 0, h->st, n, h->va.as<double>(), a);
    if (transposed) h->FT.apply_transposed<float>(a, t, a, h->st, h->cnt);
    else h->F.apply<float, float, float>(a, t, a, h->st, h->cnt);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(BLOCK), 0, h->st, n, (const float*)a,
                       h->va.as<double>());
  } else {
    if (transposed) h->FT.apply_transposed<double>(h->va.as<double>(), h->vb.as<double>(), h->va.as<double>(),
                                                   h->st, h->cnt);
    else h->F.apply<double, double, double>(h->va.as<double>(), h->vb.as<double>(), h->va.as<double>(), h->st, h->cnt);
  }
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, h->perm.as<int32_t>(),
                     h->va.as<double>(), h->vc.as<double>(), 1);
  hip_check(hipMemcpyAsync(dst, h->vc.p, 8 * n, hipMemcpyDeviceToHost, h->st), "D2H");
  h->sync();
  hip_check(hipGetLastError(), "ilu apply");
  return FS_OK;
  FS_GUARD_END
}

int fs_ilu_vmult(fs_ilu* h, const double* src, double* dst, int32_t precision) {
  return ilu_apply(h, src, dst, precision, false);
}

}  // extern "C"

// ---- ILU application without allocation ------------------------------------
// The two work vectors of fs_ilu_vmult_noalloc (the permuted right-hand side
// and result, and the forward sweep's intermediate) in a file-scope device
// array: every GPU the library runs on holds its own copy until the process
// ends. Host code uses the address hipGetSymbolAddress returns for the
// current device (taken once per device), never the array's name.
constexpr int64_t FS_NOALLOC_MAX_ROWS = int64_t(1) << 21;  // 2,097,152 rows: 32 MiB
static __device__ __align__(256) double scratch[2 * FS_NOALLOC_MAX_ROWS];
static double* g_scratch_dev[64];
static std::mutex g_scratch_mu;

static double* scratch_for(int dev) {
  if (dev < 0 || dev >= 64) throw FsError(FS_ERR_HIP, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw FsError(FS_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_dev[dev] = (double*)p;
  }
  return g_scratch_dev[dev];
}

extern "C" {

int fs_ilu_vmult_noalloc(fs_ilu* h, const double* d_src, double* d_dst, int32_t precision) {
  if (!h || !d_src || !d_dst) { fs::set_error("fs_ilu_vmult_noalloc: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  if (!h->factored) throw FsError(FS_ERR_BAD_INPUT, "fs_ilu: factorize before vmult");
  if (precision != FS_PREC_FP64 && precision != FS_PREC_MIXED)
    throw FsError(FS_ERR_BAD_INPUT, "fs_ilu_vmult_noalloc: precision must be FS_PREC_FP64 or FS_PREC_MIXED");
  const int64_t n = h->n;
  const int64_t ld = (n + 63) & ~int64_t(63);
  if (ld > FS_NOALLOC_MAX_ROWS)
    throw FsError(FS_ERR_UNSUPPORTED, "fs_ilu_vmult_noalloc: more than 2097152 rows (the static work area's size); "
                                      "use fs_ilu_vmult");
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != h->device) throw FsError(FS_ERR_BAD_INPUT, "fs_ilu_vmult_noalloc: the factor lives on another device");
  double* va = scratch_for(dev);  // permuted src, then the result
  double* vb = va + ld;           // the forward sweep's intermediate
  const int g = grid1(n);
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, h->perm.as<int32_t>(), d_src, va, 0);
  if (precision == FS_PREC_MIXED) {
    float* a = (float*)vb;
    float* t = a + ld;
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(BLOCK), 0, h->st, n, (const double*)va, a);
    h->F.apply<float, float, float>(a, t, a, h->st, h->cnt);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(BLOCK), 0, h->st, n, (const float*)a, va);
  } else {
    h->F.apply<double, double, double>(va, vb, va, h->st, h->cnt);
  }
  hipLaunchKernelGGL((k_permute<double>), dim3(g), dim3(BLOCK), 0, h->st, n, h->perm.as<int32_t>(),
                     (const double*)va, d_dst, 1);
  hip_check(hipGetLastError(), "ilu apply");
  h->sync();
  return FS_OK;
  FS_GUARD_END
}
int fs_ilu_Tvmult(fs_ilu* h, const double* src, double* dst, int32_t precision) {
  return ilu_apply(h, src, dst, precision, true);
}

int fs_ilu_get_factors(fs_ilu* h, int64_t* n, int64_t* nnz, int64_t* row_ptr, int32_t* col_idx, double* val,
                       int32_t* perm) {
  if (!h) { fs::set_error("fs_ilu_get_factors: null handle"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  const HostPattern& P = h->F.host_pattern;
  if (n) *n = P.n;
  if (nnz) *nnz = P.nnz();
  if (row_ptr) std::memcpy(row_ptr, P.rp.data(), 8 * (P.n + 1));
  if (col_idx) std::memcpy(col_idx, P.ci.data(), 4 * P.nnz());
  if (perm) std::memcpy(perm, h->F.sched.perm.data(), 4 * P.n);
  if (val) {
    if (!h->factored) throw FsError(FS_ERR_BAD_INPUT, "fs_ilu: not factorized");
    hip_check(hipMemcpy(val, h->F.v64.p, 8 * P.nnz(), hipMemcpyDeviceToHost), "D2H");
  }
  return FS_OK;
  FS_GUARD_END
}

void fs_ilu_destroy(fs_ilu* h) { delete h; }

int fs_matrix_create(const fs_csr* M, fs_matrix** out) {
  if (!out) { fs::set_error("fs_matrix_create: out is null"); return FS_ERR_BAD_INPUT; }
  *out = nullptr;
  FS_GUARD_BEGIN
  ensure_device();
  std::unique_ptr<fs_matrix> h(new fs_matrix());
  HostCsr hM = csr_from_abi(M, "M", true);
  h->M.upload_pattern(hM);
  h->M.v64.upload(hM.v.data(), 8 * std::max<int64_t>(1, hM.nnz()));
  h->M.v32.alloc(4 * std::max<int64_t>(1, hM.nnz()));
  hip_check(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking), "stream");
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid1(hM.nnz())), dim3(BLOCK), 0, h->st, hM.nnz(),
                     h->M.v64.as<double>(), h->M.v32.as<float>());
  h->x.alloc(8 * std::max<int64_t>(1, hM.n_cols));
  h->y.alloc(8 * std::max<int64_t>(1, hM.n_rows));
  hip_check(hipStreamSynchronize(h->st), "sync");
  *out = h.release();
  return FS_OK;
  FS_GUARD_END
}

int fs_matrix_vmult(fs_matrix* h, const double* x, double* y, int32_t precision) {
  if (!h || !x || !y) { fs::set_error("fs_matrix_vmult: null argument"); return FS_ERR_BAD_INPUT; }
  FS_GUARD_BEGIN
  const DevCsr& M = h->M;
  hip_check(hipMemcpyAsync(h->x.p, x, 8 * M.n_cols, hipMemcpyHostToDevice, h->st), "H2D");
  if (precision == FS_PREC_MIXED)
    spmv_launch<float, double, double, double, double, 0>(M.G, M.n_rows, 0, M.rp.as<int64_t>(), M.ci.as<int32_t>(),
                                                          M.v32.as<float>(), h->x.as<double>(), h->y.as<double>(),
                                                          (const double*)nullptr, 0, 0, h->st);
  else
    spmv_launch<double, double, double, double, double, 0>(M.G, M.n_rows, 0, M.rp.as<int64_t>(), M.ci.as<int32_t>(),
                                                           M.v64.as<double>(), h->x.as<double>(), h->y.as<double>(),
                                                           (const double*)nullptr, 0, 0, h->st);
  hip_check(hipMemcpyAsync(y, h->y.p, 8 * M.n_rows, hipMemcpyDeviceToHost, h->st), "D2H");
  hip_check(hipStreamSynchronize(h->st), "sync");
  return FS_OK;
  FS_GUARD_END
}

void fs_matrix_destroy(fs_matrix* h) { delete h; }

}  // extern "C"
