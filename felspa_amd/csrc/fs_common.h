// Internal host-side helpers shared by the generator, the symbolic ILU, the
// orderings and the HIP solver library.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace fs {

// Host CSR matrix: 64-bit row pointers (ILU(k) fill of 3D problems exceeds
// 2^31 entries), 32-bit column indices, FP64 values.
struct HostCsr {
  int64_t n_rows = 0, n_cols = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
  int64_t nnz() const { return rp.empty() ? 0 : rp.back(); }
};

// Pattern-only CSR.
struct HostPattern {
  int64_t n = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  int64_t nnz() const { return rp.empty() ? 0 : rp.back(); }
};

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* get_error();

// Set-up phase times: with g_verbose set, phase_mark prints the time since
// the previous mark (or since phase_reset) to stderr.
extern int g_verbose;
void phase_reset();
void phase_mark(const char* what);

// transpose (pattern + values)
HostCsr transpose(const HostCsr& a);

}  // namespace fs
