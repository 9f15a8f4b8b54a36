/* A plain C caller of the library's C-ABI (include/fs_stokes.h and
 * include/fs_assemble.h), as a deal.II-free host would use it.
 *
 *   fs_c_smoke          host part only: default options, a generated Q2-Q1
 *                       system, its symbolic ILU(1) and block-colour schedule
 *   fs_c_smoke --gpu    also fs_stokes_create / fs_stokes_solve /
 *                       fs_stokes_destroy on that system, and the FP64
 *                       residual of the returned solution checked here
 *
 * Prints "fs_c_smoke ok" and exits 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fs_assemble.h"
#include "fs_stokes.h"

#define CHECK(cond, msg)                                            \
  do {                                                              \
    if (!(cond)) {                                                  \
      fprintf(stderr, "fs_c_smoke: %s (%s)\n", msg, fs_last_error(NULL)); \
      return 1;                                                     \
    }                                                               \
  } while (0)

/* y += M x for a CSR block with rows [0, m) */
static void csr_gemv(int64_t m, const int64_t* rp, const int32_t* ci, const double* v, const double* x,
                     double* y) {
  for (int64_t r = 0; r < m; ++r)
    for (int64_t p = rp[r]; p < rp[r + 1]; ++p) y[r] += v[p] * x[ci[p]];
}

int main(int argc, char** argv) {
  const int gpu = argc > 1 && strcmp(argv[1], "--gpu") == 0;
  fs_opts o;
  fs_opts_default(&o);
  CHECK(o.method == FS_METHOD_FGMRES && o.rtol == 1e-8 && o.ilu_k_A == 1 && o.ilu_k_S == 0, "default options");

  /* 2D Rayleigh-Taylor layers, free slip, 8x8 Q2-Q1 */
  fs_problem p;
  memset(&p, 0, sizeof(p));
  p.dim = 2;
  p.n_cells = 8;
  p.field = FS_FIELD_LAYERS;
  p.bc = FS_BC_FREE_SLIP;
  p.eta_matrix = 1.0;
  p.eta_incl = 100.0;
  p.rho_matrix = 1.0;
  p.rho_incl = 1.1;
  p.amp = 0.02;
  p.gravity = 1.0;
  p.shear_velocity = 1.0;
  fs_system* sys = NULL;
  CHECK(fs_system_create(&p, &sys) == FS_OK, "fs_system_create");
  fs_system_info info;
  CHECK(fs_system_get_info(sys, &info) == FS_OK, "fs_system_get_info");
  const int64_t nu = info.n_u, np = info.n_p, n = nu + np;
  const int64_t rows[5] = {nu, nu, np, np, np}, cols[5] = {nu, np, nu, np, np};
  fs_csr blk[5];
  for (int w = 0; w < 5; ++w) {
    const int64_t* rp;
    const int32_t* ci;
    const double* v;
    CHECK(fs_system_csr(sys, w, &rp, &ci, &v) == FS_OK, "fs_system_csr");
    blk[w].n_rows = rows[w];
    blk[w].n_cols = cols[w];
    blk[w].row_ptr = rp;
    blk[w].col_idx = ci;
    blk[w].val = v;
    blk[w].diag_first = 0;
  }

  /* host symbolic phase: ILU(1) of A contains A's pattern */
  int64_t* frp = NULL;
  int32_t* fci = NULL;
  CHECK(fs_ilu_symbolic(&blk[0], 1, &frp, &fci) == FS_OK, "fs_ilu_symbolic");
  CHECK(frp[nu] >= blk[0].row_ptr[nu], "ILU(1) smaller than A");
  fs_free(frp);
  fs_free(fci);

  /* block-colour schedule: a permutation of the velocity rows */
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * nu);
  int32_t* bp = (int32_t*)malloc(sizeof(int32_t) * (nu + 1));
  int32_t* sp = (int32_t*)malloc(sizeof(int32_t) * (nu + 1));
  int32_t* hint = (int32_t*)malloc(sizeof(int32_t) * nu);
  char* seen = (char*)calloc(nu, 1);
  int32_t nb = 0, ns = 0;
  CHECK(fs_system_block_hint(sys, 5, hint) == FS_OK, "fs_system_block_hint");
  CHECK(fs_block_color_ordering(&blk[0], 1, 0, hint, perm, bp, &nb, sp, &ns) == FS_OK, "fs_block_color_ordering");
  CHECK(nb >= ns && ns >= 1 && bp[nb] == nu, "schedule sizes");
  for (int64_t i = 0; i < nu; ++i) {
    CHECK(perm[i] >= 0 && perm[i] < nu && !seen[perm[i]], "permutation");
    seen[perm[i]] = 1;
  }
  free(perm);
  free(bp);
  free(sp);
  free(hint);
  free(seen);
  if (!gpu) {
    fs_system_destroy(sys);
    printf("fs_c_smoke ok (host part: n_u %lld, n_p %lld, %d stages)\n", (long long)nu, (long long)np, ns);
    return 0;
  }

  /* the solve, from a zero guess, host buffers */
  fs_handle* h = NULL;
  CHECK(fs_stokes_create(&blk[0], &blk[1], &blk[2], &blk[3], &blk[4], &o, &h) == FS_OK, "fs_stokes_create");
  const double* rhs = fs_system_rhs(sys);
  double* x = (double*)calloc(n, sizeof(double));
  double* r = (double*)calloc(n, sizeof(double));
  fs_stats st;
  CHECK(fs_stokes_solve(h, rhs, x, &st) == FS_OK, "fs_stokes_solve");
  /* r = K x, K = [[A, Bt], [B, K11]] from the caller's arrays */
  csr_gemv(nu, blk[0].row_ptr, blk[0].col_idx, blk[0].val, x, r);
  csr_gemv(nu, blk[1].row_ptr, blk[1].col_idx, blk[1].val, x + nu, r);
  csr_gemv(np, blk[2].row_ptr, blk[2].col_idx, blk[2].val, x, r + nu);
  csr_gemv(np, blk[3].row_ptr, blk[3].col_idx, blk[3].val, x + nu, r + nu);
  double rn = 0, bn = 0;
  for (int64_t i = 0; i < n; ++i) {
    rn += (rhs[i] - r[i]) * (rhs[i] - r[i]);
    bn += rhs[i] * rhs[i];
  }
  const double rel = sqrt(rn / bn);
  CHECK(rel <= 1e-8 * 1.0001 && fabs(rel - st.true_rel_residual) <= 1e-9, "residual of the returned solution");
  fs_stokes_destroy(h);
  free(x);
  free(r);
  fs_system_destroy(sys);
  printf("fs_c_smoke ok (solve: %d outer iterations, residual %.3e)\n", st.outer_iters, rel);
  return 0;
}
