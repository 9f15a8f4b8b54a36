// MI355X Stokes saddle-point solver: device data, setup, Krylov drivers and
// the C-ABI of include/fs_stokes.h.
//
// Internally every vector lives in the solver's order: velocity rows permuted
// by the A-block ILU schedule, pressure rows by the M_p ILU schedule, so the
// preconditioner never gathers. The caller's order is restored at the
// boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fs_common.h"
#include "fs_stokes.h"
#include "kernels.h"
#include "ordering.h"
#include "solver_internal.h"

using namespace fsk;

namespace fs {

// ---------------------------------------------------------------------------
// errors and device memory
static int64_t g_mem_cap = 0;
static int64_t g_mem_used = 0;
}  // namespace fs

extern "C" void fs_set_device_memory_cap(int64_t bytes) { fs::g_mem_cap = bytes; }

namespace fs {

FsError::FsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
      throw FsError(FS_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    throw FsError(FS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
}

void* dev_alloc(size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (g_mem_cap > 0 && g_mem_used + (int64_t)bytes > g_mem_cap)
    throw FsError(FS_ERR_OUT_OF_MEMORY, "device memory cap exceeded");
  void* p = nullptr;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  g_mem_used += (int64_t)bytes;
  return p;
}

DevBuf::~DevBuf() { reset(); }
void DevBuf::reset() {
  if (p) {
    (void)hipFree(p);
    g_mem_used -= (int64_t)bytes;
  }
  p = nullptr;
  bytes = 0;
}
void DevBuf::alloc(size_t b) {
  reset();
  p = dev_alloc(b);
  bytes = b;
}
void DevBuf::upload(const void* src, size_t b) {
  if (bytes < b) alloc(b);
  hip_check(hipMemcpy(p, src, b, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void ensure_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw FsError(FS_ERR_NO_DEVICE, "no HIP device visible");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw FsError(FS_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library built for gfx950");
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
  }
  return cus;
}

uint64_t timeout_ticks() {
  static uint64_t t = 0;
  if (!t) {
    int dev = 0, khz = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (khz <= 0) khz = 100000;
    t = (uint64_t)khz * 1000ull * 5ull;  // 5 s
  }
  return t;
}

// ---------------------------------------------------------------------------
// host CSR input checks (caller layout; deal.II diagonal-first rows accepted)
HostCsr csr_from_abi(const fs_csr* m, const char* name, bool need_vals) {
  if (!m || !m->row_ptr || !m->col_idx || (need_vals && !m->val))
    throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": null CSR array");
  if (m->n_rows < 0 || m->n_cols < 0 || m->n_cols > INT32_MAX)
    throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": bad dimensions");
  HostCsr h;
  h.n_rows = m->n_rows;
  h.n_cols = m->n_cols;
  const int64_t nnz = m->row_ptr[m->n_rows];
  if (m->row_ptr[0] != 0 || nnz < 0) throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": bad row_ptr");
  h.rp.assign(m->row_ptr, m->row_ptr + m->n_rows + 1);
  h.ci.resize(nnz);
  h.v.resize(need_vals || m->val ? nnz : 0);
  std::vector<std::pair<int32_t, double>> tmp;
  for (int64_t r = 0; r < h.n_rows; ++r) {
    if (h.rp[r + 1] < h.rp[r]) throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": row_ptr not monotone");
    tmp.clear();
    for (int64_t p = h.rp[r]; p < h.rp[r + 1]; ++p) {
      const int32_t c = m->col_idx[p];
      if (c < 0 || c >= h.n_cols) throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": column out of range");
      tmp.emplace_back(c, m->val ? m->val[p] : 0.0);
    }
    // deal.II rows store the diagonal first; everything is re-sorted here
    std::sort(tmp.begin(), tmp.end(),
              [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) { return a.first < b.first; });
    for (size_t t = 1; t < tmp.size(); ++t)
      if (tmp[t].first == tmp[t - 1].first) throw FsError(FS_ERR_BAD_INPUT, std::string(name) + ": duplicate column");
    int64_t p = h.rp[r];
    for (auto& e : tmp) {
      h.ci[p] = e.first;
      if (!h.v.empty()) h.v[p] = e.second;
      ++p;
    }
  }
  return h;
}

// position map: sorted-layout entry order of the caller's arrays. The caller
// may pass diagonal-first rows; map_caller[e] = index into the sorted row.
std::vector<int64_t> caller_to_sorted(const fs_csr* m) {
  const int64_t nnz = m->row_ptr[m->n_rows];
  std::vector<int64_t> map(nnz);
  std::vector<std::pair<int32_t, int64_t>> tmp;
  for (int64_t r = 0; r < m->n_rows; ++r) {
    tmp.clear();
    for (int64_t p = m->row_ptr[r]; p < m->row_ptr[r + 1]; ++p) tmp.emplace_back(m->col_idx[p], p);
    std::sort(tmp.begin(), tmp.end());
    int64_t q = m->row_ptr[r];
    for (auto& e : tmp) map[e.second] = q++;
  }
  return map;
}

HostPattern pattern_of(const HostCsr& m) {
  HostPattern p;
  p.n = m.n_rows;
  p.rp = m.rp;
  p.ci = m.ci;
  return p;
}

// ---------------------------------------------------------------------------
// ILU factor on the device
static size_t dense_wg_lds_limit();

void IluDev::build(const HostPattern& src_pattern_new, const Schedule& sc, int k, int tsz_max) {
  sched = sc;
  n = src_pattern_new.n;
  bool conflict = false;
  HostPattern F = symbolic_ilu(src_pattern_new, k, &sched, &conflict);
  if (conflict || !schedule_is_valid(F, sched))
    throw FsError(FS_ERR_UNSUPPORTED, "ILU fill joins two blocks of one stage; use a larger block size");
  fs::phase_mark("ILU: symbolic phase + schedule check");
  nnz = F.nnz();
  std::vector<int32_t> ol(n), od(n), ou(n);
  std::vector<int64_t> lc(n + 1, 0), uc(n + 1, 0);
  int64_t sum_lext = 0, sum_uext = 0;
  int64_t cap_need = 0;
  // per-row offsets, and per-column counts of in-block entries (an in-block
  // column lies in the row's own block: blocks are independent)
  int mlen = 0, bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sum_lext, sum_uext) reduction(max : mlen) \
    reduction(| : bad)
  for (int b = 0; b < sched.n_blocks(); ++b) {
    const int32_t r0 = sched.blk_ptr[b], r1 = sched.blk_ptr[b + 1];
    // (the stage kernels read row r0 of every block unconditionally)
    if (r1 <= r0) { bad |= 1; continue; }
    for (int32_t r = r0; r < r1; ++r) {
      const int64_t base = F.rp[r], len = F.rp[r + 1] - base;
      const int32_t* c = F.ci.data() + base;
      const int32_t* e = c + len;
      const int32_t olint = (int32_t)(std::lower_bound(c, e, r0) - c);
      const int32_t odiag = (int32_t)(std::lower_bound(c, e, r) - c);
      const int32_t ouext = (int32_t)(std::lower_bound(c, e, r1) - c);
      if (odiag >= len || c[odiag] != r) { bad |= 2; break; }
      ol[r] = olint;
      od[r] = odiag;
      ou[r] = ouext;
      for (int32_t p = olint; p < odiag; ++p) lc[c[p] + 1]++;
      for (int32_t p = odiag + 1; p < ouext; ++p) uc[c[p] + 1]++;
      sum_lext += olint;
      sum_uext += len - ouext;
      mlen = std::max<int>(mlen, (int)len);
    }
  }
  if (bad & 1) throw FsError(FS_ERR_BAD_INPUT, "ILU schedule with an empty block");
  if (bad & 2) throw FsError(FS_ERR_BAD_INPUT, "ILU pattern lacks a diagonal");
  max_len = mlen;
  // numeric work of the IKJ factorisation: sum over L entries (i, k) of the
  // length of row k's U part (each such entry streams that U part once)
  {
    std::vector<int32_t> ulen(n);  // U entries per row (a small table the sum reads at random)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) ulen[i] = (int32_t)(F.rp[i + 1] - F.rp[i] - od[i] - 1);
    int64_t upd = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : upd)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t p = F.rp[i]; p < F.rp[i] + od[i]; ++p) upd += ulen[F.ci[p]];
    ikj_updates = upd;
  }
  fs::phase_mark("ILU: row offsets, IKJ work");
  // per stage: the longest row and L part (the numeric kernel's LDS), and
  // the longest U part of any row up to the stage's end (the rows its L
  // columns can name: the numeric kernel's loads in flight)
  stage_max_len.assign(sched.n_stages(), 1);
  stage_max_nl.assign(sched.n_stages(), 0);
  stage_max_u.assign(sched.n_stages(), 0);
  {
    int umax = 0;
    for (int s = 0; s < sched.n_stages(); ++s) {
      for (int32_t r = sched.blk_ptr[sched.stage_ptr[s]]; r < sched.blk_ptr[sched.stage_ptr[s + 1]]; ++r) {
        const int len = (int)(F.rp[r + 1] - F.rp[r]);
        stage_max_len[s] = std::max(stage_max_len[s], len);
        stage_max_nl[s] = std::max(stage_max_nl[s], od[r]);
        umax = std::max(umax, len - od[r] - 1);
      }
      stage_max_u[s] = umax;
      int hs = 0, win = 0;
      if (ilu_numeric_plan(stage_max_len[s], stage_max_nl[s], dense_wg_lds_limit(), hs, win) == 0)
        throw FsError(FS_ERR_UNSUPPORTED, "ILU row too long for the numeric kernel's LDS");
    }
  }
  // rows longer than 512 entries (3D): factorise by stages, the external L
  // columns of all later rows applied after each stage
  bins.clear();
  staged = max_len > 512 && sched.n_stages() > 1;
  if (staged) {
    const size_t lim = dense_wg_lds_limit();
    // groups: up to 3 consecutive rows of one block with the same columns
    // (the velocity components of a node), which share every lookup
    std::vector<uint8_t> same(n, 0);
    std::vector<int32_t> blk_of(n);
    for (int b2 = 0; b2 < sched.n_blocks(); ++b2)
      for (int32_t r = sched.blk_ptr[b2]; r < sched.blk_ptr[b2 + 1]; ++r) blk_of[r] = b2;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 1; r < n; ++r) {
      const int64_t l0 = F.rp[r] - F.rp[r - 1], l1 = F.rp[r + 1] - F.rp[r];
      same[r] = blk_of[r] == blk_of[r - 1] && l0 == l1 &&
                std::equal(F.ci.begin() + F.rp[r], F.ci.begin() + F.rp[r + 1], F.ci.begin() + F.rp[r - 1]);
    }
    // node triples as sources: rows k, k+1, k+2 that form one such group
    // (the stage updates apply their three L columns in one step)
    {
      std::vector<uint8_t> tr(n, 0);
      for (int64_t r = 0; r < n;) {
        int g = 1;
        while (g < 3 && r + g < n && same[r + g]) ++g;
        tr[r] = g == 3;
        r += g;
      }
      tri.upload(tr.data(), tr.size());
    }
    static const int caps[] = {256, 512, 768, 1024, 1536, 2048, 2560, 3072, 4096, 5120, 6144, 8192};
    constexpr int NC = sizeof(caps) / sizeof(caps[0]);
    std::vector<std::vector<int32_t>> lists(NC * 3);  // [cap][G - 1]
    for (int64_t r = 0; r < n;) {
      int g = 1;
      while (g < 3 && r + g < n && same[r + g]) ++g;
      if (ol[r] > 0) {
        const int len = (int)(F.rp[r + 1] - F.rp[r]);
        int c = 0;
        while (c + 1 < NC && caps[c] < len) ++c;
        if (caps[c] < len) throw FsError(FS_ERR_UNSUPPORTED, "ILU row too long for the stage update's LDS");
        lists[3 * c + g - 1].push_back((int32_t)r);
      }
      r += g;
    }
    for (int c = 0; c < NC; ++c)
      for (int g = 3; g >= 1; --g) {
        auto& lst = lists[3 * c + g - 1];
        if (lst.empty()) continue;
        NumBin nb;
        nb.mcap = caps[c];
        nb.G = g;
        // the map is keyed by node: a row names at most mcap nodes (usually
        // a third as many), so mcap + 1 slots always leave one free; 4 mcap
        // unless that costs workgroups per CU. Triples that do not fit split
        // into a pair and a single row (same columns), else into single rows
        int hmin = 64, hgood = 64;
        while (hmin < nb.mcap + 1) hmin *= 2;
        while (hgood < 4 * nb.mcap) hgood *= 2;
        const size_t wmin = ilu_update_lds_bytes(nb.mcap, hmin, g), wgood = ilu_update_lds_bytes(nb.mcap, hgood, g);
        if (wmin > lim) {
          if (g == 1) throw FsError(FS_ERR_UNSUPPORTED, "ILU row too long for the stage update's LDS");
          const bool two = g == 3 && ilu_update_lds_bytes(nb.mcap, hmin, 2) <= lim;
          auto& one = lists[3 * c];
          auto& pair = lists[3 * c + 1];
          for (int32_t r0 : lst) {
            if (two) {
              pair.push_back(r0);
              one.push_back(r0 + 2);
            } else {
              for (int k2 = 0; k2 < g; ++k2) one.push_back(r0 + k2);
            }
          }
          std::sort(one.begin(), one.end());
          std::sort(pair.begin(), pair.end());
          continue;
        }
        nb.hs = (wgood <= lim && lim / wgood >= lim / wmin) ? hgood : hmin;
        nb.nrows = (int)lst.size();
        nb.rows.upload(lst.data(), 4 * lst.size());
        nb.first.resize(sched.n_stages());
        for (int st2 = 0; st2 < sched.n_stages(); ++st2) {
          const int32_t after = sched.blk_ptr[sched.stage_ptr[st2 + 1]];
          nb.first[st2] = (int)(std::lower_bound(lst.begin(), lst.end(), after) - lst.begin());
        }
        bins.push_back(std::move(nb));
      }
  }
  for (int64_t i = 0; i < n; ++i) {
    lc[i + 1] += lc[i];
    uc[i + 1] += uc[i];
  }
  fs::phase_mark("ILU: stage bins");
  n_lin = lc[n];
  n_uin = uc[n];
  for (int b = 0; b < sched.n_blocks(); ++b) {
    const int32_t r0 = sched.blk_ptr[b], r1 = sched.blk_ptr[b + 1];
    cap_need = std::max<int64_t>(cap_need, std::max(lc[r1] - lc[r0], uc[r1] - uc[r0]));
  }
  // column-ordered lists: rows visited in ascending order fill each column in
  // ascending row order
  std::vector<uint16_t> lr(std::max<int64_t>(n_lin, 1)), ur(std::max<int64_t>(n_uin, 1));
  std::vector<int64_t> lpo(std::max<int64_t>(n_lin, 1)), upo(std::max<int64_t>(n_uin, 1));
  {
    std::vector<int64_t> fl(lc.begin(), lc.end() - 1), fu(uc.begin(), uc.end() - 1);
    for (int b = 0; b < sched.n_blocks(); ++b) {
      const int32_t r0 = sched.blk_ptr[b], r1 = sched.blk_ptr[b + 1];
      for (int32_t r = r0; r < r1; ++r) {
        const int64_t base = F.rp[r];
        for (int32_t p = ol[r]; p < od[r]; ++p) {
          const int32_t c = F.ci[base + p];
          lr[fl[c]] = (uint16_t)(r - r0);
          lpo[fl[c]++] = base + p;
        }
        for (int32_t p = od[r] + 1; p < ou[r]; ++p) {
          const int32_t c = F.ci[base + p];
          ur[fu[c]] = (uint16_t)(r - r0);
          upo[fu[c]++] = base + p;
        }
      }
    }
  }
  max_rows = std::max(1, (int)sched.max_block_rows);
  if (max_rows > 65535) throw FsError(FS_ERR_UNSUPPORTED, "ILU block larger than 65535 rows");
  fs::phase_mark("ILU: in-block column lists");
  // LDS budget: one block per workgroup, <= 48 KiB so several workgroups fit a CU
  const int per_wave_budget = 48 * 1024;
  const int fixed = (int)lds_bytes_per_wave(max_rows, 0, tsz_max);
  if (fixed > per_wave_budget) throw FsError(FS_ERR_UNSUPPORTED, "ILU block too large for LDS");
  const int cap_fit = std::max(0, (per_wave_budget - fixed) / (tsz_max + 2));
  // at least one slot, so the clamped prefetch of an empty column stays inside LDS
  cap = (int)std::max<int64_t>(1, std::min<int64_t>(cap_need, cap_fit));
  // lanes per row in the part of a block solve that reads other blocks: the
  // whole block in one pass of a 256-thread workgroup when it has few rows,
  // about four entries per lane otherwise
  const double avg = n ? (double)(sum_lext + sum_uext) / (2.0 * n) : 0;
  int g_rows = 4;
  while (g_rows < 64 && g_rows * 2 * max_rows <= 256) g_rows *= 2;
  int g_len = 4;
  while (g_len < 64 && g_len * 4 < avg) g_len *= 2;
  G = std::max(4, std::min(g_rows, g_len));
  rp.upload(F.rp.data(), 8 * (n + 1));
  ci.upload(F.ci.data(), 4 * std::max<int64_t>(1, nnz));
  v64.alloc(8 * std::max<int64_t>(1, nnz));
  o_lint.upload(ol.data(), 4 * n);
  o_diag.upload(od.data(), 4 * n);
  o_uext.upload(ou.data(), 4 * n);
  blk_ptr.upload(sched.blk_ptr.data(), 4 * sched.blk_ptr.size());
  stage_ptr.upload(sched.stage_ptr.data(), 4 * sched.stage_ptr.size());
  lcp.upload(lc.data(), 8 * (n + 1));
  ucp.upload(uc.data(), 8 * (n + 1));
  lrow.upload(lr.data(), 2 * lr.size());
  urow.upload(ur.data(), 2 * ur.size());
  lpos.upload(lpo.data(), 8 * lpo.size());
  upos.upload(upo.data(), 8 * upo.size());
  lv64.alloc(8 * lpo.size());
  uv64.alloc(8 * upo.size());
  lv32.alloc(4 * lpo.size());
  uv32.alloc(4 * upo.size());
  fs::phase_mark("ILU: CSR and in-block uploads");
  // packed in-block inverses and block-contiguous external entries for the
  // dense stage kernel (blocks of at most 256 rows: one block's right-hand
  // side sits in LDS)
  dense = dense && sched.max_block_rows <= 256;
  if (dense) {
    // packed triangles, each starting at a multiple of 4 elements
    std::vector<int64_t> off(sched.n_blocks() + 1, 0);
    int64_t cap_inv = 0;
    for (int b = 0; b < sched.n_blocks(); ++b) {
      const int64_t sz = (packed_tri_size(sched.blk_ptr[b + 1] - sched.blk_ptr[b]) + 3) & ~int64_t(3);
      off[b + 1] = off[b] + sz;
      cap_inv = std::max(cap_inv, sz);
    }
    inv_total = off.back();
    inv_cap = (int)cap_inv;
    inv_off.upload(off.data(), 8 * off.size());
    linv64.alloc(8 * std::max<int64_t>(4, inv_total));
    uinv64.alloc(8 * std::max<int64_t>(4, inv_total));
    linv32.alloc(4 * std::max<int64_t>(4, inv_total));
    uinv32.alloc(4 * std::max<int64_t>(4, inv_total));
    // pieces of split blocks: 3072 chunks when rows are long (3D ILU(2): about
    // 1,500 entries per row), else 1536; FS_ILU_PIECE=1536|3072 overrides
    piece = F.nnz() > 600 * n ? DENSE_PIECE_LARGE : DENSE_PIECE_SMALL;
    if (const char* e = std::getenv("FS_ILU_PIECE")) {
      const int v = std::atoi(e);
      if (v == DENSE_PIECE_SMALL || v == DENSE_PIECE_LARGE) piece = v;
    }
    // external entries, L part [0, o_lint) and U part [o_uext, len) of every
    // row, in chunks of 4: every row's list padded to a multiple of 4
    // (padding: position -1 = value 0, column = the row's first external
    // column, a finished value that is read and multiplied by 0), rows of a
    // block after each other. Per block {first row, rows, first chunk,
    // chunks}; per row its first chunk relative to the block (low 16 bits)
    // and its chunk count (high 16 bits).
    auto ext = [&](bool lower, DevBuf& desc, DevBuf& rng, DevBuf& didx, DevBuf& dhcol, DevBuf& dhalo,
                   DevBuf& dv64, DevBuf& dv32, int64_t& n_ent, int& c_max, std::vector<int32_t>& dsc, int& ib,
                   int& h_max, int& ph_max, int64_t& n_halo, std::vector<std::vector<int32_t>>& phl) {
      const int nb = sched.n_blocks();
      std::vector<int32_t> rg(n);
      dsc.assign(4 * (size_t)nb, 0);
      int64_t q = 0;  // chunks so far
      c_max = 0;
      for (int b = 0; b < nb; ++b) {
        const int32_t r0 = sched.blk_ptr[b], r1 = sched.blk_ptr[b + 1];
        const int64_t qb = q;
        for (int32_t r = r0; r < r1; ++r) {
          const int64_t cnt = lower ? ol[r] : F.rp[r + 1] - F.rp[r] - ou[r];
          const int64_t nc = (cnt + 3) / 4;
          if (q - qb > 65535 || nc > 65535) throw FsError(FS_ERR_UNSUPPORTED, "ILU block with too many external entries");
          rg[r] = (int32_t)((q - qb) | (nc << 16));
          q += nc;
        }
        if (q > INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "ILU factor with too many external entries");
        dsc[4 * b] = r0;
        dsc[4 * b + 1] = r1 - r0;
        dsc[4 * b + 2] = (int32_t)qb;
        dsc[4 * b + 3] = (int32_t)(q - qb);
        c_max = std::max<int>(c_max, (int)(q - qb));
      }
      n_ent = 4 * q;
      fs::phase_mark("  ext: chunk layout");
      // the block's halo: its distinct external columns, ascending (nearby
      // lanes gather nearby values); every entry stores its column's index in
      // the halo (1 byte when no halo exceeds 256 columns, else 2); halos
      // start at multiples of 4. A block of more than `piece` chunks has a
      // halo per piece of `piece` chunks instead, listed in phl.
      // Padding entries repeat the row's first external column. Blocks are
      // independent: halos are formed in parallel, then laid end to end.
      std::vector<uint16_t> loc(std::max<int64_t>(4, n_ent), 0);
      std::vector<std::vector<int32_t>> bh(nb);   // per block: its halo(s), padded
      std::vector<std::vector<int32_t>> bpc(nb);  // per block: {offset in bh, count} per halo
      int too_big = 0, hm = 0, phm = 0;
#pragma omp parallel reduction(| : too_big) reduction(max : hm, phm)
      {
        std::vector<int32_t> stamp(n, -1), slot(n, 0), cc, hb;
#pragma omp for schedule(dynamic, 16)
        for (int b = 0; b < nb; ++b) {
          const int nc = dsc[4 * b + 3];
          const int64_t a = 4 * (int64_t)dsc[4 * b + 2];
          cc.clear();
          for (int32_t r = sched.blk_ptr[b]; r < sched.blk_ptr[b + 1]; ++r) {
            const int64_t base = F.rp[r], len = F.rp[r + 1] - base;
            const int64_t p0 = lower ? 0 : ou[r], p1 = lower ? ol[r] : len;
            const size_t e0 = cc.size();
            for (int64_t p = p0; p < p1; ++p) cc.push_back(F.ci[base + p]);
            while ((cc.size() - e0) % 4) cc.push_back(cc[e0]);
          }
          const bool split = nc > piece;
          for (int c = 0; c < (split ? nc : std::min(nc, 1)); c += split ? piece : nc) {
            const size_t ea = 4 * (size_t)c, ez = 4 * (size_t)(split ? std::min(nc, c + piece) : nc);
            hb.clear();
            for (size_t e2 = ea; e2 < ez; ++e2)
              if (stamp[cc[e2]] != b) {
                stamp[cc[e2]] = b;
                hb.push_back(cc[e2]);
              }
            for (size_t e2 = ea; e2 < ez; ++e2) stamp[cc[e2]] = -1;
            std::sort(hb.begin(), hb.end());
            if (hb.size() > 65536) {
              too_big = 1;
              continue;
            }
            for (size_t i = 0; i < hb.size(); ++i) slot[hb[i]] = (int32_t)i;
            for (size_t e2 = ea; e2 < ez; ++e2) loc[a + e2] = (uint16_t)slot[cc[e2]];
            bpc[b].push_back((int32_t)bh[b].size());
            bpc[b].push_back((int32_t)hb.size());
            bh[b].insert(bh[b].end(), hb.begin(), hb.end());
            while (bh[b].size() % 4) bh[b].push_back(hb.empty() ? 0 : hb[0]);
            if (split) phm = std::max<int>(phm, (int)hb.size());
            else hm = std::max<int>(hm, (int)hb.size());
          }
        }
      }
      if (too_big) throw FsError(FS_ERR_UNSUPPORTED, "ILU block with too many external columns");
      fs::phase_mark("  ext: halos and indices (parallel)");
      h_max = hm;
      ph_max = std::max(ph_max, phm);
      std::vector<int64_t> hoff(nb + 1, 0);
      for (int b = 0; b < nb; ++b) hoff[b + 1] = hoff[b] + (int64_t)bh[b].size();
      if (hoff[nb] > (int64_t)INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "ILU factor with too many external columns");
      std::vector<int32_t> hc(std::max<int64_t>(4, hoff[nb]), 0), hl(2 * (size_t)nb, 0);
      phl.assign(nb, {});
      for (int b = 0; b < nb; ++b) {
        std::copy(bh[b].begin(), bh[b].end(), hc.begin() + hoff[b]);
        const bool split = dsc[4 * b + 3] > piece;
        for (size_t k2 = 0; k2 + 1 < bpc[b].size(); k2 += 2) {
          const int32_t off = (int32_t)(hoff[b] + bpc[b][k2]), cnt = bpc[b][k2 + 1];
          if (split) {
            phl[b].push_back(off);
            phl[b].push_back(cnt);
          } else {
            hl[2 * b] = off;
            hl[2 * b + 1] = cnt;
          }
        }
        std::vector<int32_t>().swap(bh[b]);
      }
      n_halo = (int64_t)hc.size();
      fs::phase_mark("  ext: halo list");
      ib = std::max(h_max, ph_max) <= 256 ? 1 : 2;
      if (ib == 1) {
        std::vector<uint8_t> ix(loc.size());
#pragma omp parallel for schedule(static)
        for (size_t e2 = 0; e2 < loc.size(); ++e2) ix[e2] = (uint8_t)loc[e2];
        didx.upload(ix.data(), ix.size());
      } else {
        didx.upload(loc.data(), 2 * loc.size());
      }
      desc.upload(dsc.data(), 4 * dsc.size());
      rng.upload(rg.data(), 4 * rg.size());
      dhcol.upload(hc.data(), 4 * hc.size());
      dhalo.upload(hl.data(), 4 * hl.size());
      dv64.alloc(8 * loc.size());
      dv32.alloc(4 * loc.size());
    };
    n_ext = sum_lext + sum_uext;
    std::vector<int32_t> dl, du;
    std::vector<std::vector<int32_t>> pl, pu;
    xl_phmax = xu_phmax = 0;
    ext(true, xldesc, xlrng, xlidx, xlhcol, xlhalo, xlv64, xlv32, n_xl, xl_max, dl, xl_ib, xl_hmax, xl_phmax, n_xlh,
        pl);
  fs::phase_mark("ILU: external L entries and halos");
    ext(false, xudesc, xurng, xuidx, xuhcol, xuhalo, xuv64, xuv32, n_xu, xu_max, du, xu_ib, xu_hmax, xu_phmax, n_xuh,
        pu);
  fs::phase_mark("ILU: external U entries and halos");
    // blocks with many external chunks are summed in pieces of `piece`
    // chunks by workgroups of their own; pieces listed stage by stage
    int64_t part_need = 0;
    auto split = [&](const std::vector<int32_t>& dsc, const std::vector<std::vector<int32_t>>& phl, DevBuf& dsplit,
                     DevBuf& dpieces, std::vector<int>& pc_ptr) {
      const int nb = sched.n_blocks();
      std::vector<int32_t> sp(2 * (size_t)nb), pcs;
      pc_ptr.assign(1, 0);
      int64_t off = 0;
      for (int st = 0; st < sched.n_stages(); ++st) {
        for (int b = sched.stage_ptr[st]; b < sched.stage_ptr[st + 1]; ++b) {
          const int nr = dsc[4 * b + 1], nc = dsc[4 * b + 3];
          const int S = nc > piece ? (nc + piece - 1) / piece : 1;
          sp[2 * b] = S;
          sp[2 * b + 1] = (int32_t)off;
          if (S > 1) {
            for (int q = 0; q < S; ++q) {  // {block, first chunk, end chunk, piece}, {halo, halo columns, 0, 0}
              pcs.push_back(b);
              pcs.push_back(q * piece);
              pcs.push_back(std::min(nc, (q + 1) * piece));
              pcs.push_back(q);
              pcs.push_back(phl[b][2 * q]);
              pcs.push_back(phl[b][2 * q + 1]);
              pcs.push_back(0);
              pcs.push_back(0);
            }
            off += (int64_t)S * nr;
            if (off > INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "ILU factor with too many external entries");
          }
        }
        pc_ptr.push_back((int)(pcs.size() / 8));
      }
      if (pcs.empty()) pcs.assign(8, 0);
      dsplit.upload(sp.data(), 4 * sp.size());
      dpieces.upload(pcs.data(), 4 * pcs.size());
      part_need = std::max(part_need, off);
    };
    split(dl, pl, xlsplit, xlpieces, xl_pc_ptr);
    split(du, pu, xusplit, xupieces, xu_pc_ptr);
  fs::phase_mark("ILU: pieces");
    dpart.alloc(8 * std::max<int64_t>(1, part_need));
    // the last stage has no external U entries when no two blocks of a stage
    // couple: its backward solve is then in-block only and runs in the same
    // launch as its forward solve
    const int ns = sched.n_stages();
    fuse_last = false;
    if (ns > 0) {
      int64_t cnt = 0;
      for (int32_t r = sched.blk_ptr[sched.stage_ptr[ns - 1]]; r < n; ++r) cnt += F.rp[r + 1] - F.rp[r] - ou[r];
      fuse_last = cnt == 0;
    }
  }
  // the FP32 copy of the factor in CSR form serves the substitution kernels
  // only: not needed when the dense kernels take every precision
  need_v32 = true;
  if (dense) {
    const size_t lim = dense_wg_lds_limit();
    const int hc = (std::max(xl_hmax, xu_hmax) + 3) & ~3;
    auto fits = [&](int tv, int t, int rc) { return dense_lds_bytes(inv_cap, hc, 512 * rc, tv, t, DS_FWD) <= lim; };
    need_v32 = !(fits(8, 8, dense_rc<double, double>()) && fits(4, 8, dense_rc<float, double>()) &&
                 fits(4, 4, dense_rc<float, float>()));
  }
  // every array the numeric and stage kernels dereference is on the device
  // (a launch on an unset one would fault the device, not report an error)
  if (!rp.p || !ci.p || !v64.p || !o_lint.p || !o_diag.p || !o_uext.p || !blk_ptr.p || !stage_ptr.p || !lcp.p ||
      !ucp.p || !lrow.p || !urow.p || !lpos.p || !upos.p || !lv64.p || !uv64.p ||
      (dense && (!inv_off.p || !linv64.p || !uinv64.p || !xldesc.p || !xudesc.p || !xlrng.p || !xurng.p ||
                 !xlv64.p || !xuv64.p || !xlv32.p || !xuv32.p || !dpart.p)) ||
      (staged && !tri.p))
    throw FsError(FS_ERR_HIP, "ILU factor: device arrays missing after the symbolic phase");
  if (keep_pattern) host_pattern = std::move(F);
}

template <typename T>
IluView<T> IluDev::view() const {
  const bool d = std::is_same<T, double>::value;
  IluView<T> v;
  v.rp = (const int64_t*)rp.p;
  v.ci = (const int32_t*)ci.p;
  v.v = d ? (T*)v64.p : (T*)v32.p;
  v.o_lint = (const int32_t*)o_lint.p;
  v.o_diag = (const int32_t*)o_diag.p;
  v.o_uext = (const int32_t*)o_uext.p;
  v.blk_ptr = (const int32_t*)blk_ptr.p;
  v.lcp = (const int64_t*)lcp.p;
  v.lrow = (const uint16_t*)lrow.p;
  v.lv = d ? (const T*)lv64.p : (const T*)lv32.p;
  v.ucp = (const int64_t*)ucp.p;
  v.urow = (const uint16_t*)urow.p;
  v.uv = d ? (const T*)uv64.p : (const T*)uv32.p;
  return v;
}
template IluView<double> IluDev::view<double>() const;
template IluView<float> IluDev::view<float>() const;

static int grid_for(int64_t n, int per_block = BLOCK, int cap = 4096) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

void IluDev::factorize(const int64_t* srp, const int32_t* sci, const double* sv, int64_t col_limit,
                       hipStream_t st, int* d_err, Counters& cnt) {
  IluView<double> V = view<double>();
  // FS_ILU_ROWWISE=1 runs every in-block stage one row at a time (the form
  // taken when three rows of a stage do not fit in LDS), e.g. to test it on
  // small systems
  const char* rw = std::getenv("FS_ILU_ROWWISE");
  const bool rowwise = rw && std::atoi(rw) == 1;
  if (staged) {
    hip_check(hipMemsetAsync(v64.p, 0, 8 * (size_t)nnz, st), "memset");
    hipLaunchKernelGGL(k_ilu_source_scatter, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, V, n, srp, sci, sv,
                       col_limit);
    cnt.launches++;
  }
  for (int s = 0; s < sched.n_stages(); ++s) {
    const int b0 = sched.stage_ptr[s], b1 = sched.stage_ptr[s + 1];
    if (b1 <= b0) continue;
    if (staged && s > 0) {
      // every external L column of the stage's rows (all in earlier,
      // finished stages), applied in ascending order by one workgroup per
      // row group before the stage's own kernel: each row is loaded, hashed
      // and stored once. The update is latency-bound and its LDS allows two
      // (long rows) to four workgroups per CU, so the workgroups are wide:
      // 768 threads x 1 entry per chunk for rows above 1024 entries, 256 x 2
      // below (at 48^3, ILU(2), one launch per finished stage for all later
      // rows: 256 x 8 took 11.0 s, 256 x 4 10.6 s, 128 x 8 15.8 s, 512 x 2
      // 8.0 s, 1024 x 1 7.2 s per factorisation; at 32^3 with node-triple
      // steps 1024 x 1 1.22 s, 512 x 2 1.17 s, 768 x 1 1.14 s)
      const int32_t cb = sched.blk_ptr[b0];
      for (const NumBin& nb : bins) {
        const int first = nb.first[s - 1], cnt_rows = nb.first[s] - first;
        if (cnt_rows <= 0) continue;
        const size_t lds = ilu_update_lds_bytes(nb.mcap, nb.hs, nb.G);
        const int32_t* rl = nb.rows.as<int32_t>() + first;
#define FS_UPD(NT, M, G)                                                                                       \
  {                                                                                                            \
    auto fn = &k_ilu_numeric_update<NT, M, G>;                                                                 \
    if (lds > 64 * 1024)                                                                                       \
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),    \
                "hipFuncSetAttribute");                                                                        \
    hipLaunchKernelGGL(fn, dim3(cnt_rows), dim3(NT), lds, st, V, rl, tri.as<uint8_t>(), cnt_rows, 0, cb,        \
                       nb.mcap, nb.hs);                                                                        \
  }
        if (nb.mcap <= 1024) {
          if (nb.G == 3) FS_UPD(256, 2, 3) else if (nb.G == 2) FS_UPD(256, 2, 2) else FS_UPD(256, 2, 1)
        } else {
          if (nb.G == 3) FS_UPD(768, 1, 3) else if (nb.G == 2) FS_UPD(768, 1, 2) else FS_UPD(768, 1, 1)
        }
#undef FS_UPD
        cnt.launches++;
      }
    }
    const int ml = stage_max_len[s];
    if (staged) {
      // the in-block columns by node groups (k_ilu_numeric_block), when the
      // stage's longest row fits with a node-keyed map of at most half load,
      // three rows at once or else one (a row names at most ml nodes, so a
      // map of more than ml slots is the last resort)
      int hsb = 64;
      while (hsb < 2 * ml) hsb *= 2;
      const bool tri3 = ilu_update_lds_bytes(ml, hsb, 3) <= dense_wg_lds_limit() && !rowwise;
      if (!tri3 && ilu_update_lds_bytes(ml, hsb, 1) > dense_wg_lds_limit()) {
        hsb = 64;
        while (hsb < ml + 1) hsb *= 2;
      }
      const size_t ldsb = ilu_update_lds_bytes(ml, hsb, tri3 ? 3 : 1);
      if (ldsb <= dense_wg_lds_limit() && ml <= 8192) {
#define FS_BLK(NT, M)                                                                                          \
  {                                                                                                            \
    auto fn = tri3 ? &k_ilu_numeric_block<NT, M, true> : &k_ilu_numeric_block<NT, M, false>;                   \
    if (ldsb > 64 * 1024)                                                                                      \
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsb),   \
                "hipFuncSetAttribute");                                                                        \
    hipLaunchKernelGGL(fn, dim3(b1 - b0), dim3(NT), ldsb, st, V, tri.as<uint8_t>(), b0, b1, ml, hsb, d_err);    \
  }
        if (ml <= 1024) FS_BLK(256, 2)
        else FS_BLK(768, 1)
#undef FS_BLK
        cnt.launches++;
        continue;
      }
    }
    // LDS sized by the stage's longest row; threads x loads per thread cover
    // the longest U part its L columns can name (one wave up to 256 entries,
    // four up to 1024, sixteen waves of 2 loads each beyond)
    int hs = 0, win = 0;
    const size_t lds = ilu_numeric_plan(ml, stage_max_nl[s], dense_wg_lds_limit(), hs, win);
    if (lds == 0) throw FsError(FS_ERR_UNSUPPORTED, "ILU row too long for the numeric kernel's LDS");
    const int um = stage_max_u[s];
#define FS_NUM(NT, M)                                                                                          \
  {                                                                                                            \
    auto fn = &k_ilu_numeric_stage<NT, M>;                                                                     \
    if (lds > 64 * 1024)                                                                                       \
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),    \
                "hipFuncSetAttribute");                                                                        \
    hipLaunchKernelGGL(fn, dim3(b1 - b0), dim3(NT), lds, st, V, b0, b1, srp, sci, sv, col_limit, ml, hs, win,   \
                       staged ? 1 : 0, d_err);                                                                 \
  }
    if (um <= 256) FS_NUM(64, 4)
    else if (um <= 1024) FS_NUM(256, 4)
    else FS_NUM(1024, 2)
#undef FS_NUM
    cnt.launches++;
  }
  cnt.bytes += (double)nnz * 12.0 * 2.0;
  hip_check(hipGetLastError(), "ilu numeric launch");
  if (need_v32) {
    if (!v32.p) v32.alloc(4 * std::max<int64_t>(1, nnz));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(nnz)), dim3(BLOCK), 0, st, nnz,
                       (const double*)v64.p, (float*)v32.p);
    cnt.launches++;
  }
  gather_inblock(st);
  if (dense) {
    const int nb = sched.n_blocks();
    const int gb = std::max(1, std::min(nb, 65535));
    if (n_xl > 0)
      hipLaunchKernelGGL(k_gather_ext, dim3(gb), dim3(BLOCK), 0, st, xldesc.as<i32x4>(), xlrng.as<int32_t>(),
                         rp.as<int64_t>(), o_lint.as<int32_t>(), o_uext.as<int32_t>(), 1, v64.as<double>(),
                         xlv64.as<double>(), xlv32.as<float>(), nb);
    if (n_xu > 0)
      hipLaunchKernelGGL(k_gather_ext, dim3(gb), dim3(BLOCK), 0, st, xudesc.as<i32x4>(), xurng.as<int32_t>(),
                         rp.as<int64_t>(), o_lint.as<int32_t>(), o_uext.as<int32_t>(), 0, v64.as<double>(),
                         xuv64.as<double>(), xuv32.as<float>(), nb);
    // the block in LDS as two dense tiles when they fit (at most 99 rows)
    const int ldi = (std::max(1, (int)sched.max_block_rows) + 1) & ~1;
    const size_t ldsi = (size_t)2 * sched.max_block_rows * ldi * sizeof(double);
    if (ldsi <= dense_wg_lds_limit()) {
      auto fn = &k_ilu_block_inverse_lds<256>;
      if (ldsi > 64 * 1024)
        hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsi),
                  "hipFuncSetAttribute");
      hipLaunchKernelGGL(fn, dim3(std::max(1, std::min(nb, 65535))), dim3(256), ldsi, st, V, nb, inv_off.as<int64_t>(),
                         linv64.as<double>(), uinv64.as<double>(), ldi);
    } else {
      hipLaunchKernelGGL(k_ilu_block_inverse, dim3(std::max(1, std::min(nb, 65535))), dim3(BLOCK), 0, st, V, nb,
                         inv_off.as<int64_t>(), linv64.as<double>(), uinv64.as<double>());
    }
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(inv_total)), dim3(BLOCK), 0, st, inv_total,
                       linv64.as<double>(), linv32.as<float>());
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(inv_total)), dim3(BLOCK), 0, st, inv_total,
                       uinv64.as<double>(), uinv32.as<float>());
    cnt.launches += 3;
    cnt.bytes += (double)nnz * 12.0 + 2.0 * inv_total * 12.0;
  }
}

// in-block entries in column order, both precisions
void IluDev::gather_inblock(hipStream_t st) {
  if (n_lin > 0) {
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(n_lin)), dim3(BLOCK), 0, st, n_lin, lpos.as<int64_t>(),
                       v64.as<double>(), lv64.as<double>(), lv32.as<float>());
  }
  if (n_uin > 0) {
    hipLaunchKernelGGL(k_gather_vals, dim3(grid_for(n_uin)), dim3(BLOCK), 0, st, n_uin, upos.as<int64_t>(),
                       v64.as<double>(), uv64.as<double>(), uv32.as<float>());
  }
}

template <typename T, typename TB, typename TY, int G, bool UNIT, bool FWD>
static void launch_stage(const IluDev& F, int b0, int b1, const TB* b, TY* y, hipStream_t st) {
  const size_t lds = lds_bytes_per_wave(F.max_rows, F.cap, sizeof(T));
  const int nb = b1 - b0;
  if (nb <= 0) return;
  const int g = std::max(1, std::min(nb, 65535));
  hipLaunchKernelGGL((k_ilu_stage<T, TB, TY, G, UNIT, FWD>), dim3(g), dim3(BLOCK), lds, st, F.view<T>(), b0, b1,
                     b, y, F.max_rows, F.cap);
}

template <typename T, typename TB, typename TY, bool UNIT, bool FWD>
static void launch_stage_g(const IluDev& F, int b0, int b1, const TB* b, TY* y, hipStream_t st) {
  switch (F.G) {
    case 4: launch_stage<T, TB, TY, 4, UNIT, FWD>(F, b0, b1, b, y, st); break;
    case 8: launch_stage<T, TB, TY, 8, UNIT, FWD>(F, b0, b1, b, y, st); break;
    case 16: launch_stage<T, TB, TY, 16, UNIT, FWD>(F, b0, b1, b, y, st); break;
    case 32: launch_stage<T, TB, TY, 32, UNIT, FWD>(F, b0, b1, b, y, st); break;
    default: launch_stage<T, TB, TY, 64, UNIT, FWD>(F, b0, b1, b, y, st); break;
  }
}

// LDS available to one workgroup of the dense stage kernel (4 waves)
static size_t dense_wg_lds_limit() {
  static size_t lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    lim = v > 0 ? (size_t)v : 64 * 1024;
    lim = std::min<size_t>(lim, 160 * 1024);
  }
  return lim;
}

// the dense kernels' view of one triangle (forward: L, backward: U)
template <typename TV>
static DenseTri<TV> dense_view(const IluDev& F, bool fwd) {
  const bool d = std::is_same<TV, double>::value;
  DenseTri<TV> D;
  D.desc = fwd ? F.xldesc.as<i32x4>() : F.xudesc.as<i32x4>();
  D.rng = fwd ? F.xlrng.as<int32_t>() : F.xurng.as<int32_t>();
  D.xidx = fwd ? F.xlidx.p : F.xuidx.p;
  D.xv = fwd ? (d ? (const TV*)F.xlv64.p : (const TV*)F.xlv32.p) : (d ? (const TV*)F.xuv64.p : (const TV*)F.xuv32.p);
  D.hcol = fwd ? F.xlhcol.as<int32_t>() : F.xuhcol.as<int32_t>();
  D.halo = fwd ? F.xlhalo.as<int2>() : F.xuhalo.as<int2>();
  D.inv_off = F.inv_off.as<int64_t>();
  D.inv = fwd ? (d ? (const TV*)F.linv64.p : (const TV*)F.linv32.p) : (d ? (const TV*)F.uinv64.p : (const TV*)F.uinv32.p);
  D.split = fwd ? F.xlsplit.as<int2>() : F.xusplit.as<int2>();
  D.part = F.dpart.p;
  return D;
}

// LDS of a dense stage launch: halo values of the largest halo, and the
// partial sums of every chunk of the largest block when the workgroup stays
// within 40 KiB (several workgroups per CU), else chunks in passes of at
// least the first round's dense_rc * NT
template <typename TV, typename T, typename TY, int NT, int MODE>
static void dense_lds_plan(const IluDev& F, int& hcap, int& part_cap, size_t& lds) {
  const bool fwd = MODE != DS_BWD;
  hcap = ((fwd ? F.xl_hmax : F.xu_hmax) + 3) & ~3;
  const size_t base = dense_lds_bytes(F.inv_cap, hcap, 0, sizeof(TV), sizeof(T), MODE);
  const int need = ((fwd ? F.xl_max : F.xu_max) + 3) & ~3;
  const int first = dense_rc<TV, TY>() * NT;
  const size_t target = std::max<size_t>(40 * 1024, base + (size_t)first * sizeof(T));
  part_cap = need;
  if (base + (size_t)need * sizeof(T) > target)
    part_cap = std::max<int>(first, (int)(((target - base) / sizeof(T)) & ~(size_t)3));
  lds = dense_lds_bytes(F.inv_cap, hcap, part_cap, sizeof(TV), sizeof(T), MODE);
}

template <typename TV, typename T, typename TB, typename TY, int NT, int G, int MODE, int IB>
static void launch_dense(const IluDev& F, int b0, int b1, const TB* b, const TY* yin, TY* y, const PcgFuse<T>& pf,
                         hipStream_t st) {
  const int nb = b1 - b0;
  if (nb <= 0) return;
  const bool d = std::is_same<TV, double>::value;
  const DenseTri<TV> D = dense_view<TV>(F, MODE != DS_BWD);
  const TV* uinv = MODE == DS_BOTH ? (d ? (const TV*)F.uinv64.p : (const TV*)F.uinv32.p) : nullptr;
  int hcap = 0, part_cap = 0;
  size_t lds = 0;
  dense_lds_plan<TV, T, TY, NT, MODE>(F, hcap, part_cap, lds);
  if (lds > dense_wg_lds_limit()) throw FsError(FS_ERR_UNSUPPORTED, "ILU block too large for the dense stage kernel");
  auto fn = &k_ilu_stage_dense<TV, T, TB, TY, NT, G, MODE, IB>;
  if (lds > 64 * 1024)
    hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
              "hipFuncSetAttribute");
  hipLaunchKernelGGL(fn, dim3(nb), dim3(NT), lds, st, D, uinv, b0, F.inv_cap, hcap, part_cap, b, yin, y, pf);
}

template <typename TV, typename T, typename TB, typename TY, int MODE, int IB>
static void launch_dense_ib(const IluDev& F, int b0, int b1, const TB* b, const TY* yin, TY* y,
                            const PcgFuse<T>& pf, hipStream_t st) {
  const bool fwd = MODE != DS_BWD;
  // pieces of the stage's split blocks first
  const std::vector<int>& pp = fwd ? F.xl_pc_ptr : F.xu_pc_ptr;
  int s = 0;
  while (F.sched.stage_ptr[s] != b0) ++s;
  if (pp[s + 1] > pp[s]) {
    const int hcap = ((fwd ? F.xl_phmax : F.xu_phmax) + 3) & ~3;
    const size_t lds = sizeof(T) * ((size_t)F.piece + hcap + DENSE_SPLIT_NT);
    auto fn = F.piece == DENSE_PIECE_LARGE ? &k_ilu_ext_partials<TV, T, TY, IB, DENSE_PIECE_LARGE>
                                           : &k_ilu_ext_partials<TV, T, TY, IB, DENSE_PIECE_SMALL>;
    if (lds > dense_wg_lds_limit()) throw FsError(FS_ERR_UNSUPPORTED, "ILU block halo too large");
    if (lds > 64 * 1024)
      hip_check(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                "hipFuncSetAttribute");
    hipLaunchKernelGGL(fn, dim3(pp[s + 1] - pp[s]), dim3(DENSE_SPLIT_NT), lds, st, dense_view<TV>(F, fwd),
                       fwd ? F.xlpieces.as<i32x4>() : F.xupieces.as<i32x4>(), pp[s], hcap, yin);
  }
  // 128 threads for blocks of up to 64 rows (more workgroups resident per
  // CU), 512 above (larger blocks carry many more external entries: more of
  // them in flight per block); G threads per row in the in-block product
  const int mr = F.sched.max_block_rows;
  if (mr <= 16) launch_dense<TV, T, TB, TY, 128, 8, MODE, IB>(F, b0, b1, b, yin, y, pf, st);
  else if (mr <= 32) launch_dense<TV, T, TB, TY, 128, 4, MODE, IB>(F, b0, b1, b, yin, y, pf, st);
  else if (mr <= 64) launch_dense<TV, T, TB, TY, 128, 2, MODE, IB>(F, b0, b1, b, yin, y, pf, st);
  else if (mr <= 128) launch_dense<TV, T, TB, TY, 512, 4, MODE, IB>(F, b0, b1, b, yin, y, pf, st);
  else launch_dense<TV, T, TB, TY, 512, 2, MODE, IB>(F, b0, b1, b, yin, y, pf, st);
}

template <typename TV, typename T, typename TB, typename TY, int MODE>
static void launch_dense_g(const IluDev& F, int b0, int b1, const TB* b, const TY* yin, TY* y,
                           const PcgFuse<T>& pf, hipStream_t st) {
  // one byte per halo index when no block's halo exceeds 256 columns
  if ((MODE != DS_BWD ? F.xl_ib : F.xu_ib) == 1) launch_dense_ib<TV, T, TB, TY, MODE, 1>(F, b0, b1, b, yin, y, pf, st);
  else launch_dense_ib<TV, T, TB, TY, MODE, 2>(F, b0, b1, b, yin, y, pf, st);
}

template <typename TV, typename T, typename TB>
static void apply_dense(const IluDev& F, bool fuse, const TB* b, T* tmp, T* x, const PcgFuse<T>& pf,
                        hipStream_t st) {
  const Schedule& sc = F.sched;
  const int ns = sc.n_stages();
  const int nf = fuse ? ns - 1 : ns;  // stages with a forward launch of their own
  const PcgFuse<T> none{};
  for (int s = 0; s < nf; ++s)
    launch_dense_g<TV, T, TB, T, DS_FWD>(F, sc.stage_ptr[s], sc.stage_ptr[s + 1], b, tmp, tmp, pf, st);
  if constexpr (sizeof(TV) == 4) {
    if (fuse) launch_dense_g<TV, T, TB, T, DS_BOTH>(F, sc.stage_ptr[ns - 1], sc.stage_ptr[ns], b, tmp, x, pf, st);
  }
  for (int s = nf - 1; s >= 0; --s)
    launch_dense_g<TV, T, T, T, DS_BWD>(F, sc.stage_ptr[s], sc.stage_ptr[s + 1], tmp, x, x, none, st);
}

// x = (LU)^{-1} b with TV factor values and T vectors / arithmetic: forward
// into tmp, backward into x (tmp may alias x). With pf, the PCG update step
// that produces b (= pf->r) runs first: inside the forward stage kernels, or
// as its own kernel ahead of the substitution kernels.
template <typename TV, typename T, typename TB>
void IluDev::apply(const TB* b, T* tmp, T* x, hipStream_t st, Counters& cnt, const PcgFuse<T>* pf) const {
  const PcgFuse<T> none{};
  const int ns = sched.n_stages();
  // the dense kernels keep a block's packed triangle (the fused last stage
  // both of them), its halo values and its chunk partial sums in LDS (room
  // for the first round's partial sums of 512 threads)
  const size_t lim = dense_wg_lds_limit();
  const int pc = 512 * dense_rc<TV, T>();
  const int hc = (std::max(xl_hmax, xu_hmax) + 3) & ~3;
  if (dense && dense_lds_bytes(inv_cap, hc, pc, sizeof(TV), sizeof(T), DS_FWD) <= lim) {
    // (FP32 factors only: with FP64 ones both triangles' registers cost more
    // occupancy than the launch saves)
    const bool fuse =
      sizeof(TV) == 4 && fuse_last && dense_lds_bytes(inv_cap, hc, pc, sizeof(TV), sizeof(T), DS_BOTH) <= lim;
    apply_dense<TV, T, TB>(*this, fuse, b, tmp, x, pf ? *pf : none, st);
    cnt.launches += fuse ? 2 * ns - 1 : 2 * ns;
    // external entries (value + halo index), the halo columns, both packed
    // triangles, the per-row chunk ranges, b read and y written in both sweeps
    cnt.bytes += (double)n_ext * sizeof(TV) + (double)n_xl * xl_ib + (double)n_xu * xu_ib +
                 4.0 * (double)(n_xlh + n_xuh) + 2.0 * inv_total * sizeof(TV) + 8.0 * n + 4.0 * n * sizeof(T);
    return;
  }
  if (!std::is_same<TV, T>::value)
    throw FsError(FS_ERR_UNSUPPORTED, "ILU application with FP32 factors and FP64 vectors needs the dense kernels");
  if (std::is_same<T, float>::value && !v32.p)
    throw FsError(FS_ERR_UNSUPPORTED, "ILU substitution kernels without the FP32 factor copy");
  if (pf && pf->r) {
    hipLaunchKernelGGL((k_cg1_update<T>), dim3(256), dim3(BLOCK), 0, st, (int64_t)n, pf->alpha, pf->beta, pf->z,
                       pf->s, pf->p, pf->q, pf->x, pf->r, pf->first);
    cnt.launches++;
  }
  using TS = T;  // substitution kernels: factor values and vectors of one type
  for (int s = 0; s < ns; ++s)
    launch_stage_g<TS, TB, TS, true, true>(*this, sched.stage_ptr[s], sched.stage_ptr[s + 1], b, tmp, st);
  for (int s = ns - 1; s >= 0; --s)
    launch_stage_g<TS, TS, TS, false, false>(*this, sched.stage_ptr[s], sched.stage_ptr[s + 1], tmp, x, st);
  cnt.launches += 2 * ns;
  cnt.bytes += (double)nnz * (sizeof(T) + 4) + 16.0 * (n + 1) + 3.0 * n * sizeof(T) + 20.0 * n;
}
template void IluDev::apply<double, double, double>(const double*, double*, double*, hipStream_t, Counters&, const PcgFuse<double>*) const;
template void IluDev::apply<float, double, double>(const double*, double*, double*, hipStream_t, Counters&, const PcgFuse<double>*) const;
template void IluDev::apply<float, float, double>(const double*, float*, float*, hipStream_t, Counters&, const PcgFuse<float>*) const;
template void IluDev::apply<float, float, float>(const float*, float*, float*, hipStream_t, Counters&, const PcgFuse<float>*) const;

// transposed solve with a transposed factor (lower part = U^T with diagonal,
// upper part = L^T unit)
template <typename T>
void IluDev::apply_transposed(const T* b, T* tmp, T* x, hipStream_t st, Counters& cnt) const {
  const int ns = sched.n_stages();
  for (int s = 0; s < ns; ++s)
    launch_stage_g<T, T, T, false, true>(*this, sched.stage_ptr[s], sched.stage_ptr[s + 1], b, tmp, st);
  for (int s = ns - 1; s >= 0; --s)
    launch_stage_g<T, T, T, true, false>(*this, sched.stage_ptr[s], sched.stage_ptr[s + 1], tmp, x, st);
  cnt.launches += 2 * ns;
}
template void IluDev::apply_transposed<double>(const double*, double*, double*, hipStream_t, Counters&) const;
template void IluDev::apply_transposed<float>(const float*, float*, float*, hipStream_t, Counters&) const;

// ---------------------------------------------------------------------------
// SpMV launch on a CSR (rows [row_off, row_off + nrows))
template <typename TV, typename TX, typename TY, typename TB, typename TACC, int MODE>
void spmv_launch(int G, int64_t nrows, int64_t row_off, const int64_t* rp, const int32_t* ci, const TV* v,
                 const TX* x, TY* y, const TB* b, int64_t col_limit, int64_t x_off, hipStream_t st) {
  const int rows_per_block = BLOCK / G;
  const int g = (int)std::max<int64_t>(1, (nrows + rows_per_block - 1) / rows_per_block);
#define FS_SPMV_CASE(GG)                                                                                   \
  case GG:                                                                                                 \
    hipLaunchKernelGGL((k_spmv<TV, TX, TY, TB, TACC, GG, MODE>), dim3(g), dim3(BLOCK), 0, st, nrows, row_off, \
                       rp, ci, v, x, y, b, col_limit, x_off);                                              \
    break;
  switch (G) {
    FS_SPMV_CASE(4)
    FS_SPMV_CASE(8)
    FS_SPMV_CASE(16)
    FS_SPMV_CASE(32)
    default:
      FS_SPMV_CASE(64)
  }
#undef FS_SPMV_CASE
}

// lanes per row so that each lane streams about four entries
int lanes_for(double avg_row) {
  return avg_row > 192 ? 64 : avg_row > 96 ? 32 : avg_row > 48 ? 16 : avg_row > 20 ? 8 : 4;
}

void DevCsr::upload_pattern(const HostCsr& h, bool with_padded) {
  n_rows = h.n_rows;
  n_cols = h.n_cols;
  nnz = h.nnz();
  rp.upload(h.rp.data(), 8 * (n_rows + 1));
  ci.upload(h.ci.data(), 4 * std::max<int64_t>(nnz, 1));
  G = lanes_for(n_rows ? (double)nnz / n_rows : 1);
  padded = with_padded;
  if (!padded) return;
  std::vector<int64_t> r4(n_rows + 1, 0);
  for (int64_t r = 0; r < n_rows; ++r) r4[r + 1] = r4[r] + (h.rp[r + 1] - h.rp[r] + 3) / 4;
  nnz4 = r4[n_rows];
  std::vector<int32_t> c4(4 * std::max<int64_t>(nnz4, 1), 0);
  std::vector<int64_t> pad_pos(std::max<int64_t>(nnz, 1), 0);  // entry -> padded position
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t b = h.rp[r], e = h.rp[r + 1];
    int64_t o = 4 * r4[r];
    for (int64_t p = b; p < e; ++p) {
      c4[o] = h.ci[p];
      pad_pos[p] = o++;
    }
    const int32_t fill = e > b ? h.ci[e - 1] : 0;  // padding reads a column the row already uses
    for (; o < 4 * r4[r + 1]; ++o) c4[o] = fill;
  }
  rp4.upload(r4.data(), 8 * (n_rows + 1));
  ci4.upload(c4.data(), 4 * c4.size());
  d_pad_pos.upload(pad_pos.data(), 8 * pad_pos.size());
  n_pad = (int64_t)c4.size();
  v4_64.alloc(8 * n_pad);
  v4_32.alloc(4 * n_pad);
  hip_check(hipMemset(v4_64.p, 0, 8 * n_pad), "memset");  // padding entries stay 0
  const double g = n_rows ? (double)nnz4 / n_rows : 1;  // groups per row, about two per lane
  G4 = g > 64 ? 32 : g > 32 ? 16 : g > 12 ? 8 : 4;
}

void BlkCsr::build(const HostCsr& h, const std::vector<int32_t>& blk_ptr, bool with_f32) {
  n = h.n_rows;
  nnz = h.nnz();
  nb = (int)blk_ptr.size() - 1;
  std::vector<int32_t> dsc(4 * (size_t)nb), ci(2 * (size_t)nb), rg(std::max<int64_t>(1, n));
  int64_t q = 0;
  cnmax = rmax = 0;
  for (int b = 0; b < nb; ++b) {
    const int32_t r0 = blk_ptr[b], r1 = blk_ptr[b + 1];
    // (the SpMV reads row r0 and column-list entry 0 of every block unconditionally)
    if (r1 <= r0) throw FsError(FS_ERR_BAD_INPUT, "SpMV blocking with an empty block");
    const int64_t qb = q;
    for (int32_t r = r0; r < r1; ++r) {
      const int64_t nc = (h.rp[r + 1] - h.rp[r] + 3) / 4;
      if (q - qb > 65535 || nc > 65535) throw FsError(FS_ERR_UNSUPPORTED, "matrix block with too many entries");
      rg[r] = (int32_t)((q - qb) | (nc << 16));
      q += nc;
    }
    if (q > INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "matrix with too many entries");
    dsc[4 * b] = r0;
    dsc[4 * b + 1] = r1 - r0;
    dsc[4 * b + 2] = (int32_t)qb;
    dsc[4 * b + 3] = (int32_t)(q - qb);
    cnmax = std::max<int>(cnmax, (int)(q - qb));
    rmax = std::max<int>(rmax, r1 - r0);
  }
  n_ent = 4 * q;
  // the blocks' column lists and every entry's index into its block's list
  std::vector<int32_t> cl, loc(std::max<int64_t>(4, n_ent), 0);
  std::vector<int64_t> pos(std::max<int64_t>(1, nnz), 0);  // entry -> padded position
  std::vector<std::vector<int32_t>> bl(nb);               // per block: its column list
  int bad = 0;
  // blocks are independent: column lists and indices in parallel
#pragma omp parallel reduction(| : bad)
  {
    std::vector<int32_t> stamp(std::max<int64_t>(1, h.n_cols), -1), slot(std::max<int64_t>(1, h.n_cols), 0);
#pragma omp for schedule(dynamic, 16)
    for (int b = 0; b < nb; ++b) {
      const int32_t r0 = blk_ptr[b], r1 = blk_ptr[b + 1];
      std::vector<int32_t>& hb = bl[b];
      for (int64_t e = h.rp[r0]; e < h.rp[r1]; ++e)
        if (stamp[h.ci[e]] != b) {
          stamp[h.ci[e]] = b;
          hb.push_back(h.ci[e]);
        }
      std::sort(hb.begin(), hb.end());
      if (hb.empty()) bad |= 1;
      if (hb.size() > 65536) bad |= 2;
      if (bad) continue;
      for (size_t i = 0; i < hb.size(); ++i) slot[hb[i]] = (int32_t)i;
      int64_t o = 4 * (int64_t)dsc[4 * b + 2];
      for (int32_t r = r0; r < r1; ++r) {
        const int64_t o0 = o;
        for (int64_t e = h.rp[r]; e < h.rp[r + 1]; ++e, ++o) {
          loc[o] = slot[h.ci[e]];
          pos[e] = o;
        }
        for (; (o - o0) % 4; ++o) loc[o] = 0;
      }
    }
  }
  if (bad & 1) throw FsError(FS_ERR_BAD_INPUT, "SpMV blocking: a block without entries");
  if (bad & 2) throw FsError(FS_ERR_UNSUPPORTED, "matrix block with too many columns");
  cmax = 0;
  for (int b = 0; b < nb; ++b) {
    const std::vector<int32_t>& hb = bl[b];
    ci[2 * b] = (int32_t)cl.size();
    ci[2 * b + 1] = (int32_t)hb.size();
    cmax = std::max<int>(cmax, (int)hb.size());
    cl.insert(cl.end(), hb.begin(), hb.end());
    while (cl.size() % 4) cl.push_back(hb[0]);
    std::vector<int32_t>().swap(bl[b]);
  }
  if (cl.empty()) cl.assign(4, 0);
  if (cl.size() > (size_t)INT32_MAX) throw FsError(FS_ERR_UNSUPPORTED, "matrix with too many block columns");
  // one row per thread of a 512-thread workgroup, the block's column values
  // and chunk sums in LDS (FP64 at most): otherwise the caller keeps CSR4
  if (rmax > 512 || 8.0 * (((cmax + 3) & ~3) + (double)cnmax) > 160.0 * 1024) {
    nb = 0;
    return;
  }
  n_cols = (int64_t)cl.size();
  ib = cmax <= 256 ? 1 : 2;
  if (ib == 1) {
    std::vector<uint8_t> ix(loc.size());
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < loc.size(); ++e) ix[e] = (uint8_t)loc[e];
    idx.upload(ix.data(), ix.size());
  } else {
    std::vector<uint16_t> ix(loc.size());
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < loc.size(); ++e) ix[e] = (uint16_t)loc[e];
    idx.upload(ix.data(), 2 * ix.size());
  }
  desc.upload(dsc.data(), 4 * dsc.size());
  cinfo.upload(ci.data(), 4 * ci.size());
  rng.upload(rg.data(), 4 * rg.size());
  cols.upload(cl.data(), 4 * cl.size());
  d_pos.upload(pos.data(), 8 * pos.size());
  n_pad = std::max<int64_t>(4, n_ent);
  v64.alloc(8 * n_pad);
  if (with_f32) v32.alloc(4 * n_pad);
  hip_check(hipMemset(v64.p, 0, 8 * n_pad), "memset");  // padding entries stay 0
}

// dst[map[e]] = src[e]
static __global__ __launch_bounds__(BLOCK) void k_scatter_map(int64_t n, const int64_t* __restrict__ map,
                                                              const double* __restrict__ src,
                                                              double* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < n; e += (int64_t)gridDim.x * BLOCK)
    dst[map[e]] = src[e];
}

void BlkCsr::upload_values_device(const double* d_vals, hipStream_t st) {
  if (nb == 0) return;
  hipLaunchKernelGGL(k_scatter_map, dim3(grid_for(nnz)), dim3(BLOCK), 0, st, nnz, d_pos.as<int64_t>(), d_vals,
                     v64.as<double>());
  if (v32.p)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n_pad)), dim3(BLOCK), 0, st, n_pad,
                       v64.as<double>(), v32.as<float>());
}

void DevCsr::derive_from_v64(hipStream_t st) {
  if (!v32.p) v32.alloc(4 * std::max<int64_t>(1, nnz));
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(nnz)), dim3(BLOCK), 0, st, nnz, v64.as<double>(),
                     v32.as<float>());
  if (!padded) return;
  hipLaunchKernelGGL(k_scatter_map, dim3(grid_for(nnz)), dim3(BLOCK), 0, st, nnz, d_pad_pos.as<int64_t>(),
                     (const double*)v64.as<double>(), v4_64.as<double>());
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n_pad)), dim3(BLOCK), 0, st, n_pad,
                     v4_64.as<double>(), v4_32.as<float>());
}

void DevCsr::upload_values(const double* vals, hipStream_t st) {
  if (!v64.p) v64.alloc(8 * std::max<int64_t>(1, nnz));
  hip_check(hipMemcpyAsync(v64.p, vals, 8 * nnz, hipMemcpyHostToDevice, st), "H2D");
  derive_from_v64(st);
  hip_check(hipStreamSynchronize(st), "sync");
}

// padded SpMV launch (rows [row_off, row_off + nrows))
template <typename TV, typename TX, typename TY, typename TB, typename TACC, int MODE>
void spmv4_launch(const DevCsr& M, int64_t nrows, int64_t row_off, const TX* x, TY* y, const TB* b,
                  int64_t col_limit, hipStream_t st) {
  const int G = M.G4;
  const int rows_per_block = BLOCK / G;
  const int g = (int)std::max<int64_t>(1, (nrows + rows_per_block - 1) / rows_per_block);
  using V4 = typename Vec4<TV>::type;
  const V4* vv = std::is_same<TV, double>::value ? (const V4*)M.v4_64.p : (const V4*)M.v4_32.p;
#define FS_SPMV4_CASE(GG)                                                                                      \
  case GG:                                                                                                     \
    hipLaunchKernelGGL((k_spmv4<TV, TX, TY, TB, TACC, GG, MODE>), dim3(g), dim3(BLOCK), 0, st, nrows, row_off, \
                       M.rp4.as<int64_t>(), M.ci4.as<i32x4>(), vv, x, y, b, col_limit);                        \
    break;
  switch (G) {
    FS_SPMV4_CASE(4)
    FS_SPMV4_CASE(8)
    FS_SPMV4_CASE(16)
    default:
      FS_SPMV4_CASE(32)
  }
#undef FS_SPMV4_CASE
}
template void spmv4_launch<double, double, double, double, double, 0>(const DevCsr&, int64_t, int64_t, const double*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<double, double, double, double, double, 1>(const DevCsr&, int64_t, int64_t, const double*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<double, double, double, double, double, 2>(const DevCsr&, int64_t, int64_t, const double*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<float, float, double, double, float, 0>(const DevCsr&, int64_t, int64_t, const float*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<float, float, float, double, float, 2>(const DevCsr&, int64_t, int64_t, const float*, float*, const double*, int64_t, hipStream_t);
template void spmv4_launch<float, double, double, double, double, 0>(const DevCsr&, int64_t, int64_t, const double*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<double, float, double, double, double, 0>(const DevCsr&, int64_t, int64_t, const float*, double*, const double*, int64_t, hipStream_t);
template void spmv4_launch<float, double, double, double, double, 2>(const DevCsr&, int64_t, int64_t, const double*, double*, const double*, int64_t, hipStream_t);

}  // namespace fs

// explicit instantiations used by solver_krylov.hip
namespace fs {
template void spmv_launch<double, double, double, double, double, 0>(int, int64_t, int64_t, const int64_t*, const int32_t*, const double*, const double*, double*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<double, double, double, double, double, 1>(int, int64_t, int64_t, const int64_t*, const int32_t*, const double*, const double*, double*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<double, double, double, double, double, 2>(int, int64_t, int64_t, const int64_t*, const int32_t*, const double*, const double*, double*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<float, float, double, double, float, 0>(int, int64_t, int64_t, const int64_t*, const int32_t*, const float*, const float*, double*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<float, float, float, double, float, 2>(int, int64_t, int64_t, const int64_t*, const int32_t*, const float*, const float*, float*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<float, double, double, double, double, 0>(int, int64_t, int64_t, const int64_t*, const int32_t*, const float*, const double*, double*, const double*, int64_t, int64_t, hipStream_t);
template void spmv_launch<float, float, float, float, float, 0>(int, int64_t, int64_t, const int64_t*, const int32_t*, const float*, const float*, float*, const float*, int64_t, int64_t, hipStream_t);
}  // namespace fs
