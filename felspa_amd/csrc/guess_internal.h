// Learned warm start: model state shared between guess.hip (network,
// training) and stokes.hip (inference inside the solve).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fs_guess.h"
#include "solver_internal.h"

struct fs_guess_model {
  fs_guess_config cfg{};
  int cin = 0, cout = 0;
  int64_t n_params = 0;
  struct LayerOff {
    int64_t R, W, b;
  };
  int64_t off_Wlift = 0, off_blift = 0, off_Wp1 = 0, off_bp1 = 0, off_Wp2 = 0, off_bp2 = 0;
  std::vector<LayerOff> lay;
  fs::DevBuf p32, pbf, grad, adam_m, adam_v;  // FP32 master, BF16 copy, gradient, Adam moments
  int64_t adam_t = 0;
  float mean[4] = {0, 0, 0, 0}, stdv[4] = {1, 1, 1, 1};
  hipStream_t st = nullptr;
  // workspace for one (n_cells, batch); grows on demand
  int ws_nc = -1, ws_batch = -1;
  bool ws_train = false;
  int64_t N1 = 0, P = 0, M = 0, Mm = 0;
  fs::DevBuf ws, phi, part;
  // pointers into ws
  std::vector<float*> H, Z, T2;  // H[0..L], Z[1..L] (Z[0] unused), T2[1..L]
  std::vector<float*> S;         // S[1..dim-1] spectral intermediates
  float *O = nullptr, *P1 = nullptr, *G = nullptr, *Y = nullptr, *Xin = nullptr;
  float *dHa = nullptr, *dHb = nullptr, *dG = nullptr, *dO = nullptr, *dT2 = nullptr, *dY = nullptr;
  std::vector<float*> dS;
  // cosine matrices (BF16 bits): analysis Af [M][N1], its transpose AfT
  // [N1][M], synthesis Sb [N1][M], its transpose SbT [M][N1]
  const uint16_t *Af = nullptr, *AfT = nullptr, *Sb = nullptr, *SbT = nullptr;

  ~fs_guess_model() {
    if (st) (void)hipStreamDestroy(st);
  }
  void layout();
  void ensure_workspace(int nc, int batch, bool train);
  void forward(const float* X, int batch, hipStream_t s);
  void backward(const float* X, int batch, hipStream_t s);
};

namespace fs {
// Network forward for `batch` fields on stream s: log10(eta) on the (N+1)^d
// lattice in, normalised outputs on the (2N+1)^d lattice written to the
// model's Y buffer (returned).
const float* guess_infer(fs_guess_model* m, int n_cells, const float* d_log_eta, hipStream_t s);
}  // namespace fs
