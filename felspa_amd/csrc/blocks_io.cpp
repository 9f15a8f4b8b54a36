// Binary interchange of an assembled Stokes system (include/fs_blocks_io.h).
#include "fs_blocks_io.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "fs_common.h"

struct fs_blocks {
  int64_t n_u = 0, n_p = 0;
  struct Blk {
    int64_t n_rows = 0, n_cols = 0;
    int32_t diag_first = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
  } b[5];
  std::vector<double> rhs;
};

namespace {

struct Hash {  // FNV-1a, 64 bit
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
};

struct Writer {
  FILE* f;
  Hash hash;
  bool ok = true;
  void put(const void* p, size_t n, bool hashed = true) {
    if (n && std::fwrite(p, 1, n, f) != n) ok = false;
    if (hashed) hash.add(p, n);
  }
};

struct Reader {
  FILE* f;
  Hash hash;
  bool ok = true;
  void get(void* p, size_t n, bool hashed = true) {
    if (n && std::fread(p, 1, n, f) != n) ok = false;
    if (hashed && ok) hash.add(p, n);
  }
};

const int64_t kRows[5] = {0, 0, 1, 1, 1};  // 0: n_u rows, 1: n_p rows
const int64_t kCols[5] = {0, 1, 0, 1, 1};

int fail(const std::string& m, int code = FS_ERR_BAD_INPUT) {
  fs::set_error("%s", m.c_str());
  return code;
}

}  // namespace

extern "C" {

int fs_blocks_save(const char* path, const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11,
                   const fs_csr* Mp, const double* rhs) {
  const fs_csr* m[5] = {A, Bt, B, K11, Mp};
  if (!path) return fail("fs_blocks_save: null path");
  for (int w = 0; w < 5; ++w)
    if (!m[w] || !m[w]->row_ptr || !m[w]->col_idx || !m[w]->val) return fail("fs_blocks_save: null block");
  const int64_t nn[2] = {A->n_rows, Mp->n_rows};
  for (int w = 0; w < 5; ++w)
    if (m[w]->n_rows != nn[kRows[w]] || m[w]->n_cols != nn[kCols[w]])
      return fail("fs_blocks_save: block shapes do not match");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("fs_blocks_save: cannot write ") + path);
  Writer wr{f};
  const uint32_t ver = 1, has = rhs ? 1 : 0, zero = 0;
  wr.put("FSBK", 4, false);
  wr.put(&ver, 4, false);
  wr.put(&nn[0], 8, false);
  wr.put(&nn[1], 8, false);
  wr.put(&has, 4, false);
  wr.put(&zero, 4, false);
  for (int w = 0; w < 5; ++w) {
    const int64_t nnz = m[w]->row_ptr[m[w]->n_rows];
    const int32_t df = m[w]->diag_first;
    wr.put(&m[w]->n_rows, 8);
    wr.put(&m[w]->n_cols, 8);
    wr.put(&nnz, 8);
    wr.put(&df, 4);
    wr.put(&zero, 4);
    wr.put(m[w]->row_ptr, 8 * (size_t)(m[w]->n_rows + 1));
    wr.put(m[w]->col_idx, 4 * (size_t)nnz);
    wr.put(m[w]->val, 8 * (size_t)nnz);
  }
  if (rhs) wr.put(rhs, 8 * (size_t)(nn[0] + nn[1]));
  const uint64_t h = wr.hash.h;
  wr.put(&h, 8, false);
  const bool closed = std::fclose(f) == 0;
  if (!wr.ok || !closed) return fail(std::string("fs_blocks_save: short write to ") + path);
  return FS_OK;
}

int fs_blocks_load(const char* path, fs_blocks** out) {
  if (!path || !out) return fail("fs_blocks_load: null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("fs_blocks_load: cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  try {
    std::unique_ptr<fs_blocks> b(new fs_blocks());
    Reader rd{f};
    char magic[4];
    uint32_t ver = 0, has = 0, zero = 0;
    rd.get(magic, 4, false);
    rd.get(&ver, 4, false);
    rd.get(&b->n_u, 8, false);
    rd.get(&b->n_p, 8, false);
    rd.get(&has, 4, false);
    rd.get(&zero, 4, false);
    if (!rd.ok || std::memcmp(magic, "FSBK", 4) != 0 || ver != 1 || b->n_u < 0 || b->n_p < 0 ||
        b->n_u + b->n_p > INT32_MAX || has > 1)
      return fail("fs_blocks_load: not a version-1 block file");
    const int64_t nn[2] = {b->n_u, b->n_p};
    for (int w = 0; w < 5; ++w) {
      auto& k = b->b[w];
      int64_t nnz = 0;
      int32_t pad = 0;
      rd.get(&k.n_rows, 8);
      rd.get(&k.n_cols, 8);
      rd.get(&nnz, 8);
      rd.get(&k.diag_first, 4);
      rd.get(&pad, 4);
      if (!rd.ok || k.n_rows != nn[kRows[w]] || k.n_cols != nn[kCols[w]] || nnz < 0 ||
          nnz > k.n_rows * std::max<int64_t>(k.n_cols, 1))
        return fail("fs_blocks_load: bad block header");
      k.rp.resize(k.n_rows + 1);
      k.ci.resize(nnz);
      k.v.resize(nnz);
      rd.get(k.rp.data(), 8 * k.rp.size());
      rd.get(k.ci.data(), 4 * k.ci.size());
      rd.get(k.v.data(), 8 * k.v.size());
      if (!rd.ok) return fail("fs_blocks_load: truncated file");
      if (k.rp[0] != 0 || k.rp[k.n_rows] != nnz) return fail("fs_blocks_load: bad row pointers");
      for (int64_t r = 0; r < k.n_rows; ++r)
        if (k.rp[r + 1] < k.rp[r]) return fail("fs_blocks_load: bad row pointers");
      for (int64_t e = 0; e < nnz; ++e)
        if (k.ci[e] < 0 || k.ci[e] >= k.n_cols) return fail("fs_blocks_load: column index out of range");
    }
    if (has) {
      b->rhs.resize(b->n_u + b->n_p);
      rd.get(b->rhs.data(), 8 * b->rhs.size());
    }
    uint64_t h = 0;
    const uint64_t want = rd.hash.h;
    rd.get(&h, 8, false);
    if (!rd.ok) return fail("fs_blocks_load: truncated file");
    if (h != want) return fail("fs_blocks_load: checksum mismatch (corrupted file)");
    *out = b.release();
    return FS_OK;
  } catch (const std::bad_alloc&) {
    return fail("fs_blocks_load: host allocation failed", FS_ERR_OUT_OF_MEMORY);
  }
}

int fs_blocks_get(const fs_blocks* f, int32_t which, fs_csr* out) {
  if (!f || !out || which < 0 || which > 4) return fail("fs_blocks_get: bad argument");
  const auto& k = f->b[which];
  out->n_rows = k.n_rows;
  out->n_cols = k.n_cols;
  out->row_ptr = k.rp.data();
  out->col_idx = k.ci.data();
  out->val = k.v.data();
  out->diag_first = k.diag_first;
  return FS_OK;
}

const double* fs_blocks_rhs(const fs_blocks* f, int64_t* n_u, int64_t* n_p) {
  if (!f) return nullptr;
  if (n_u) *n_u = f->n_u;
  if (n_p) *n_p = f->n_p;
  return f->rhs.empty() ? nullptr : f->rhs.data();
}

void fs_blocks_free(fs_blocks* f) { delete f; }

}  // extern "C"
