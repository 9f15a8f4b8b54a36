// Internal declarations shared by the solver translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "fs_common.h"
#include "fs_stokes.h"
#include "kernels.h"
#include "ordering.h"

namespace fs {

struct FsError : std::runtime_error {
  int code;
  FsError(int c, const std::string& m);
};

void hip_check(hipError_t e, const char* what);
void* dev_alloc(size_t bytes);
void ensure_device();
int device_cus();
uint64_t timeout_ticks();
int lanes_for(double avg_row);

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf();
  void reset();
  void alloc(size_t b);
  void upload(const void* src, size_t b);
  template <typename T>
  T* as() const { return (T*)p; }
};

struct Counters {
  double bytes = 0;
  int64_t launches = 0;
};

struct DevCsr {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int G = 16;
  DevBuf rp, ci, v64, v32;
  // padded copy for the streaming SpMV: rows start at multiples of 4 entries
  bool padded = false;
  int64_t nnz4 = 0;  // groups of 4
  int G4 = 4;
  DevBuf rp4, ci4, v4_64, v4_32;
  int64_t n_pad = 0;  // entries of the padded layout
  void upload_pattern(const HostCsr& h, bool with_padded = false);
  // values in the plain layout (host), both precisions and both layouts
  void upload_values(const double* vals, hipStream_t st);
  // v64 already holds the values (written on the device): the FP32 copy and
  // the padded layout from it (entry -> padded position, on the device)
  void derive_from_v64(hipStream_t st);
  DevBuf d_pad_pos;
};

// A CSR matrix stored by row blocks for the blocked SpMV (k_spmv_blk_dot):
// every block's entries reference the block's own column list (its distinct
// columns, ascending) through 1-byte indices (2 when a list exceeds 256
// columns); rows padded to chunks of 4 entries (value 0, index 0).
struct BlkCsr {
  int64_t n = 0, nnz = 0, n_ent = 0, n_cols = 0;  // rows, entries, padded entries, column-list entries
  int nb = 0, ib = 1, cmax = 0, cnmax = 0, rmax = 0;
  DevBuf desc, cinfo, rng, idx, cols, v64, v32;
  int64_t n_pad = 0;  // entries of the padded layout
  // with_f32: also an FP32 copy of the values
  void build(const HostCsr& h, const std::vector<int32_t>& blk_ptr, bool with_f32 = true);
  // values in the plain layout, in device memory
  void upload_values_device(const double* d_vals, hipStream_t st);
  DevBuf d_pos;  // entry -> padded position
};

struct IluDev {
  int64_t n = 0, nnz = 0;
  Schedule sched;
  HostPattern host_pattern;
  DevBuf rp, ci, v64, v32, o_lint, o_diag, o_uext, blk_ptr, stage_ptr;
  // in-block entries in column order (values gathered after factorisation)
  DevBuf lcp, lrow, lpos, lv64, lv32, ucp, urow, upos, uv64, uv32;
  int64_t n_lin = 0, n_uin = 0;
  int max_rows = 1, cap = 0, max_len = 1, G = 16;
  std::vector<int> stage_max_len, stage_max_nl, stage_max_u;  // per stage: longest row, L part, named U part
  // left-looking numeric factorisation by stages (rows with long external L
  // parts): rows with external L columns in bins of length, each bin's rows
  // ascending; first[s] = first row of the bin after stage s
  struct NumBin {
    int mcap = 0, hs = 0, G = 1;  // row cap, hash size, rows per group
    DevBuf rows;
    std::vector<int> first;
    int nrows = 0;
  };
  int piece = fsk::DENSE_PIECE_SMALL;  // external chunks per piece of a split block
  bool staged = false;
  int64_t ikj_updates = 0;  // numeric work: sum over L entries (i, k) of |U part of row k|
  bool need_v32 = true;      // FP32 CSR copy of the factor (substitution kernels)
  bool keep_pattern = true;  // keep the host pattern (factor queries, transposed factor)
  std::vector<NumBin> bins;
  DevBuf tri;  // [n] row starts a node triple (staged factorisation only)
  // explicit inverses of the in-block triangles (packed per block), used by
  // apply(); `dense` off keeps the in-block substitution kernels
  bool dense = true;
  int64_t inv_total = 0;
  int inv_cap = 0;                 // largest packed triangle of a block
  int64_t n_ext = 0;               // external entries of L and U
  int64_t n_xl = 0, n_xu = 0;      // external entries of L / U, padded to chunks of 4 per row
  int xl_max = 0, xu_max = 0;      // largest per-block count of chunks of external L / U entries
  bool fuse_last = false;          // last stage: forward and backward in one launch
  DevBuf inv_off, linv64, uinv64, linv32, uinv32;
  DevBuf xldesc, xlrng, xlidx, xlhcol, xlhalo, xlv64, xlv32;
  DevBuf xudesc, xurng, xuidx, xuhcol, xuhalo, xuv64, xuv32;
  int xl_ib = 1, xu_ib = 1;          // bytes per halo index of an external entry
  int xl_hmax = 0, xu_hmax = 0;      // largest halo of a block
  int xl_phmax = 0, xu_phmax = 0;    // largest halo of a piece of a split block
  int64_t n_xlh = 0, n_xuh = 0;      // halo columns stored (padded)
  DevBuf xlsplit, xusplit, xlpieces, xupieces, dpart;  // pieces of blocks with many external chunks
  std::vector<int> xl_pc_ptr, xu_pc_ptr;               // per stage: range of pieces
  void gather_inblock(hipStream_t st);
  void build(const HostPattern& src_pattern_new, const Schedule& sc, int k, int tsz_max);
  template <typename T>
  fsk::IluView<T> view() const;
  void factorize(const int64_t* srp, const int32_t* sci, const double* sv, int64_t col_limit, hipStream_t st,
                 int* d_err, Counters& cnt);
  template <typename TV, typename T, typename TB>
  void apply(const TB* b, T* tmp, T* x, hipStream_t st, Counters& cnt, const fsk::PcgFuse<T>* pf = nullptr) const;
  template <typename T>
  void apply_transposed(const T* b, T* tmp, T* x, hipStream_t st, Counters& cnt) const;
};

template <typename TV, typename TX, typename TY, typename TB, typename TACC, int MODE>
void spmv_launch(int G, int64_t nrows, int64_t row_off, const int64_t* rp, const int32_t* ci, const TV* v,
                 const TX* x, TY* y, const TB* b, int64_t col_limit, int64_t x_off, hipStream_t st);

template <typename TV, typename TX, typename TY, typename TB, typename TACC, int MODE>
void spmv4_launch(const DevCsr& M, int64_t nrows, int64_t row_off, const TX* x, TY* y, const TB* b,
                  int64_t col_limit, hipStream_t st);

HostCsr csr_from_abi(const fs_csr* m, const char* name, bool need_vals);
std::vector<int64_t> caller_to_sorted(const fs_csr* m);
HostPattern pattern_of(const HostCsr& m);

}  // namespace fs
