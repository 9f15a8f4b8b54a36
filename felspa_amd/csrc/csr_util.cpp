// Host CSR utilities and the per-thread error string behind fs_last_error.
#include <chrono>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "fs_common.h"

namespace fs {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_err; }

int g_verbose = 0;
static std::chrono::steady_clock::time_point g_phase_t = std::chrono::steady_clock::now();

void phase_reset() { g_phase_t = std::chrono::steady_clock::now(); }

void phase_mark(const char* what) {
  const auto t = std::chrono::steady_clock::now();
  if (g_verbose)
    std::fprintf(stderr, "  set-up: %-44s %8.3f s\n", what, std::chrono::duration<double>(t - g_phase_t).count());
  g_phase_t = t;
}

HostCsr transpose(const HostCsr& a) {
  HostCsr t;
  t.n_rows = a.n_cols;
  t.n_cols = a.n_rows;
  t.rp.assign(t.n_rows + 1, 0);
  const int64_t nnz = a.nnz();
  for (int64_t k = 0; k < nnz; ++k) t.rp[a.ci[k] + 1]++;
  for (int64_t r = 0; r < t.n_rows; ++r) t.rp[r + 1] += t.rp[r];
  t.ci.resize(nnz);
  t.v.resize(nnz);
  std::vector<int64_t> pos(t.rp.begin(), t.rp.end() - 1);
  for (int64_t r = 0; r < a.n_rows; ++r)
    for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k) {
      int64_t p = pos[a.ci[k]]++;
      t.ci[p] = (int32_t)r;
      t.v[p] = a.v.empty() ? 0.0 : a.v[k];
    }
  return t;
}

}  // namespace fs

extern "C" void fs_free(void* p) { std::free(p); }
