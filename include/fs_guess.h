/*
 * fs_guess.h -- learned warm start for the Stokes solve (MI355X, gfx950).
 *
 * No counterpart in the reference: FELSPA starts GMRES from the previous
 * time step's solution or from zero (the `soln` argument of
 * StokesLinearSystem::solve_gmres, src/pde/stokes_ilu.cc:523-524, zeroed at
 * first allocation :114-124). This header adds a neural operator that maps
 * the log10-viscosity field of a Q2-Q1 problem to an initial (u, p); it plugs
 * into that same initial-guess slot through FS_GUESS_LEARNED and
 * fs_stokes_set_guess_field (fs_stokes.h). The solve always converges from
 * the guess to the same FP64 true-residual tolerance; a guess whose residual
 * exceeds the zero guess's (||b||) is replaced by zero.
 *
 * Network (cosine-basis spectral neural operator, all tensors channel-major
 * [batch][channel][lattice], lattice = (2N+1)^dim Q2 nodes, x fastest):
 *   input  X   = [log10(eta)/3, x_0, .., x_{dim-1}]          (dim+1 channels)
 *   lift   H_0 = W_lift X + b_lift                            (width channels)
 *   layer  H_l = act(S_l(H_{l-1}) + W_l H_{l-1} + b_l),  l = 1..layers,
 *               act = GELU except identity on the last layer;
 *               S_l = synthesis o (per-mode channel mixing R_l) o analysis,
 *               analysis  c_m = int_0^1 f(x) cos(pi m x) dx (trapezoid rule),
 *               synthesis f(x) = sum_m c'_m (m ? 2 : 1) cos(pi m x),
 *               `modes` cosine modes per axis (resolution independent)
 *   head   Y   = W_p2 GELU(W_p1 H_L + b_p1) + b_p2             (dim+1 channels)
 *   output u_a = Y_a * std_a + mean_a,  p = Y_dim * std_dim + mean_dim.
 * Every GEMM-shaped product (lift, pointwise mixing, the per-axis cosine
 * transforms, the per-mode mixing, the head, and all their gradients) runs
 * on one BF16 MFMA kernel (v_mfma_f32_16x16x32_bf16) with FP32 accumulation;
 * the master weights and Adam moments are FP32, the GEMM operands BF16.
 *
 * Parameter order (get/set_params, gradients, the weight file): W_lift
 * [width][dim+1], b_lift[width], then per layer R_l [modes^dim][width][width]
 * (mode index with the x mode fastest, then c_in, then c_out), W_l
 * [width][width] (out, in), b_l [width]; then W_p1 [proj][width], b_p1
 * [proj], W_p2 [dim+1][proj], b_p2 [dim+1].
 *
 * Device pointers passed here are plain addresses of device (HBM) memory on
 * the current HIP device; the model uses its own stream and synchronises it
 * before returning unless stated otherwise, so the caller's writes to the
 * buffers it passes must be complete (its stream synchronised) first.
 */
#ifndef FS_GUESS_H
#define FS_GUESS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs_guess_config {
  int32_t dim;        /* 2 or 3 */
  int32_t width;      /* channels of the hidden layers (32) */
  int32_t modes;      /* cosine modes kept per axis */
  int32_t layers;     /* spectral layers (4) */
  int32_t proj;       /* hidden width of the output head */
  uint64_t seed;      /* initialisation seed */
} fs_guess_config;

typedef struct fs_guess_model fs_guess_model;

int fs_guess_model_create(const fs_guess_config* cfg, fs_guess_model** out);
void fs_guess_model_destroy(fs_guess_model* m);
int fs_guess_model_get_config(const fs_guess_model* m, fs_guess_config* cfg);

/* Versioned weight file: magic "FSGM", version, config, normalisation,
 * parameter count, FNV-1a hash of the FP32 parameters, then the parameters.
 * Loading checks the hash and the sizes. */
int fs_guess_model_save(const fs_guess_model* m, const char* path);
int fs_guess_model_load(const char* path, fs_guess_model** out);

int64_t fs_guess_model_num_params(const fs_guess_model* m);
/* FP32 master parameters / last gradients (host arrays of num_params). */
int fs_guess_model_get_params(const fs_guess_model* m, float* out);
int fs_guess_model_set_params(fs_guess_model* m, const float* in);
int fs_guess_model_get_grads(const fs_guess_model* m, float* out);
/* output normalisation (dim+1 values each), see the header comment */
int fs_guess_model_set_normalization(fs_guess_model* m, const float* mean, const float* std);
int fs_guess_model_get_normalization(const fs_guess_model* m, float* mean, float* std);

/* Network input on the (2N+1)^dim lattice from log10(eta) on the (N+1)^dim
 * pressure-node lattice (multilinear interpolation = the Q1 field at the Q2
 * nodes), for `batch` fields stored one after another. Device pointers:
 * d_log_eta [batch][(N+1)^dim], d_in [batch][dim+1][(2N+1)^dim]. */
int fs_guess_prepare_input(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_log_eta,
                           float* d_in);

/* Forward pass. d_in as above, d_out [batch][dim+1][(2N+1)^dim] receives the
 * normalised outputs Y (denormalise with mean/std). */
int fs_guess_model_forward(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_in, float* d_out);

/* Mean-squared error against normalised targets d_target (same layout as
 * d_out), loss = sum (Y - T)^2 / (batch (dim+1) (2N+1)^dim), and its gradient
 * with respect to every parameter (kept in the model, see get_grads). */
int fs_guess_model_loss_grad(fs_guess_model* m, int32_t n_cells, int32_t batch, const float* d_in,
                             const float* d_target, double* loss);

/* One Adam step (decoupled weight decay) on the FP32 master parameters with
 * the last gradients; refreshes the BF16 copies. */
int fs_guess_model_adam_step(fs_guess_model* m, float lr, float beta1, float beta2, float eps,
                             float weight_decay);

/* The MFMA GEMM every layer uses, exported for tests: for b < batch
 *   C_b = alpha * A_b B_b + beta * C_b,
 * A_b(i,k) = A[b*sAb + i*sAm + k*sAk] (FP32, or BF16 bits when a_bf16),
 * B_b(k,j) = B[b*sBb + k*sBk + j*sBn], C_b(i,j) = C[b*sCb + i*sCm + j*sCn]
 * (FP32). Operands are rounded to BF16 (nearest even), products summed in
 * FP32. Device pointers; runs on the null stream and synchronises. */
int fs_gemm_bf16(int32_t M, int32_t N, int32_t K, int32_t batch, const void* A, int32_t a_bf16, int64_t sAb,
                 int64_t sAm, int64_t sAk, const void* B, int32_t b_bf16, int64_t sBb, int64_t sBk,
                 int64_t sBn, float* C, int64_t sCb, int64_t sCm, int64_t sCn, float alpha, float beta);

/* Kernel timing for the roofline of the learned part: mean time of one
 * forward pass (ms) and its GEMM flops, on the model's stream. */
int fs_guess_model_time_forward(fs_guess_model* m, int32_t n_cells, int32_t batch, int32_t reps,
                                double* ms, double* flops);

#ifdef __cplusplus
}
#endif
#endif
