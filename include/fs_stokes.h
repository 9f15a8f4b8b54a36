/*
 * fs_stokes.h -- C-ABI of the MI355X (gfx950) Stokes saddle-point solver.
 *
 * Drop-in for FELSPA's StokesLinearSystem<dim,double>::solve(soln, rhs, ctrl)
 * (include/felspa/pde/stokes_ilu.h:232-236, src/pde/stokes_ilu.cc:442-561).
 * The caller keeps the assembled blocks of the BlockSparseMatrix
 *     K = [[A, B^T], [B, K11]],  preconditioner(1,1) = M_p(1/eta)
 * in host memory as CSR arrays (deal.II's diagonal-first row layout is
 * accepted as-is); the library copies them to the GPU once per sparsity
 * pattern and owns every device buffer. No C++ types and no exceptions cross
 * this boundary: every entry point returns an FS_* status code and
 * fs_last_error() describes the last failure on the calling thread.
 * Calls on one handle must be serialised by the caller (the reference
 * solve() is not re-entrant either).
 */
#ifndef FS_STOKES_H
#define FS_STOKES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  FS_OK = 0,
  FS_NOT_CONVERGED = 1,      /* deal.II SolverControl::NoConvergence */
  FS_ERR_BAD_INPUT = -1,     /* malformed CSR, size mismatch, NaN in rhs/guess */
  FS_ERR_ZERO_PIVOT = -2,    /* ILU pivot zero or non-finite (DEBUG-only check in the reference, ilu.implement.h:498-499) */
  FS_ERR_OUT_OF_MEMORY = -3, /* device or host allocation failed (or FS_MAX_DEVICE_BYTES exceeded) */
  FS_ERR_HIP = -4,           /* HIP runtime error */
  FS_ERR_UNSUPPORTED = -5,   /* option combination not implemented */
  FS_ERR_TIMEOUT = -6,       /* a device-side bounded wait expired */
  FS_ERR_NO_DEVICE = -7      /* no gfx950 device visible */
};

/* ---- options ----------------------------------------------------------- */
enum {
  FS_METHOD_GMRES_REF = 0, /* deal.II SolverGMRES semantics: left preconditioning,
                              restart, ReductionControl(tol, reduce) on the
                              preconditioned residual (stokes_ilu.cc:533-542) */
  FS_METHOD_FGMRES = 1,    /* flexible right-preconditioned GMRES, stops on the
                              FP64 true residual ||b-Kx|| <= rtol ||b|| */
  FS_METHOD_SCR = 2        /* Schur-complement reduction (stokes_ilu.cc:564-710) */
};
enum {
  FS_PREC_FP64 = 0,              /* everything in FP64 */
  FS_PREC_MIXED = 1,             /* preconditioner in FP32: matrix blocks and ILU factors stored in
                                    FP32, inner PCG / ILU application / inner CG computed in FP32;
                                    outer FGMRES in FP64, including its products with K (the
                                    restarts make it iterative refinement) */
  FS_PREC_MIXED_FP64_VECTORS = 2 /* as FS_PREC_MIXED with the preconditioner's vectors and
                                    arithmetic in FP64 (stored values FP32) */
};
enum {
  FS_ORDER_REFERENCE = 0,   /* ILU on the caller's row order (sequential stages) */
  FS_ORDER_BLOCK_COLOR = 1  /* ILU on a block-coloured symmetric permutation */
};
enum { FS_GUESS_USER = 0, FS_GUESS_ZERO = 1, FS_GUESS_LEARNED = 2 };

typedef struct fs_csr {
  int64_t n_rows, n_cols;
  const int64_t* row_ptr; /* n_rows+1 */
  const int32_t* col_idx; /* row_ptr[n_rows] */
  const double* val;      /* row_ptr[n_rows]; may be NULL for pattern-only calls */
  int32_t diag_first;     /* 1: square block stored with its diagonal first (deal.II) */
} fs_csr;

typedef struct fs_opts {
  int32_t method;        /* FS_METHOD_* */
  int32_t precision;     /* FS_PREC_* */
  int32_t ordering;      /* FS_ORDER_* */
  int32_t ilu_k_A;       /* Control::set_level_of_fill A (stokes_ilu.cc:139-147), default 1 */
  int32_t ilu_k_S;       /* ... S, default 0 */
  int32_t restart;       /* Krylov basis size, at most 255; 0 (default): the method's own, 23 for
                            FS_METHOD_GMRES_REF (deal.II max_n_tmp_vectors 25 - 2,
                            stokes_ilu.cc:538) and 60 for FS_METHOD_FGMRES */
  int32_t max_outer;     /* 0: n_u + n_p (stokes_ilu.cc:534) */
  double rtol;           /* tol = rtol ||rhs|| (stokes_ilu.cc:533), default 1e-8 */
  double reduce;         /* ReductionControl reduce (stokes_common.h:343-348), default 1e-6 */
  int32_t inner_cg_max;  /* 0: n_p (linear_operators.implement.h:76); 1 (a single ILU-PCG step,
                            the failure flag set when it misses the tolerance) runs as ordinary
                            kernel launches, more steps as one cooperative kernel */
  double inner_cg_tol;   /* InverseMatrixCG tol = reduce (linear_operators.h:44-45), default 1e-12;
                            FS_METHOD_FGMRES in FP64 uses it as a relative reduction only */
  double inner_cg_tol_mixed; /* FP32 inner CG relative tolerance in FS_PREC_MIXED */
  int32_t inner_cg_max_mixed; /* its step cap (1: as for inner_cg_max) */
  int32_t apply_diagonal_scaling; /* StokesSolverControl::apply_diagonal_scaling (stokes_common.h:360) */
  int32_t guess;         /* FS_GUESS_* */
  int32_t block_rows;    /* target rows per block for FS_ORDER_BLOCK_COLOR (0: default) */
  const int32_t* block_hint; /* optional [n_u] block id per velocity row (e.g. geometric boxes) */
  int32_t verbose;
  int32_t inner_a_iters; /* A^{-1} in the block preconditioner of FS_METHOD_FGMRES:
                            0 = one ILU(k) application (the reference's
                            APreconditioner), >0 = that many ILU(k)-preconditioned
                            CG steps on A (default 60) */
} fs_opts;

void fs_opts_default(fs_opts* opts);

typedef struct fs_stats {
  int32_t status;            /* FS_OK / FS_NOT_CONVERGED / error */
  int32_t outer_iters;       /* Krylov iterations (SolverControl::last_step) */
  int32_t restarts;          /* restart cycles */
  int32_t guess_used;        /* FS_GUESS_* actually used (learned falls back to zero) */
  int64_t inner_iters;       /* total inner CG iterations */
  double last_value;         /* SolverControl::last_value: value seen by the stopping rule */
  double initial_rel_residual; /* ||b - K x0|| / ||b|| */
  double true_rel_residual;  /* ||b - K x|| / ||b|| in FP64 after the solve */
  double t_setup_ms;         /* numeric factorisation + scaling */
  double t_krylov_ms;        /* Krylov iterations */
  double t_total_ms;
  double bytes_algorithmic;  /* sum over kernels of their algorithmic HBM bytes */
  int64_t n_kernel_launches;
  double t_guess_ms;         /* learned guess: network inference + scatter + residual check
                                (part of t_krylov_ms) */
  double guess_rel_residual; /* ||b - K x0|| / ||b|| of the learned guess before the fallback test */
} fs_stats;

typedef struct fs_handle fs_handle;

/* StokesLinearSystem::setup_constraints_and_system + SparseILU symbolic phase
 * (stokes_ilu.cc:226-257, ilu.implement.h:58-87): copies the blocks to the
 * GPU, builds the orderings and the ILU(k) patterns of A and M_p. */
int fs_stokes_create(const fs_csr* A, const fs_csr* Bt, const fs_csr* B, const fs_csr* K11,
                     const fs_csr* Mp, const fs_opts* opts, fs_handle** out);

/* New values on the same patterns (use_previous_sparsity = true,
 * stokes_ilu.cc:547-548): numeric factorisation happens in the next solve. */
int fs_stokes_refactor(fs_handle* h, const double* A_val, const double* Bt_val,
                       const double* B_val, const double* K11_val, const double* Mp_val);
/* The same with the values already in device memory (FP64, the caller's
 * layouts given at create), e.g. written by fs_assembler_assemble: device
 * copies of the value maps (about 8 bytes per entry of K and M_p, allocated
 * on the first call) scatter them into the solver's layouts; nothing passes
 * through the host. */
int fs_stokes_refactor_device(fs_handle* h, const double* d_A, const double* d_Bt,
                              const double* d_B, const double* d_K11, const double* d_Mp);

/* StokesLinearSystem::solve (stokes_ilu.cc:442-468 -> solve_gmres :509-561):
 * rhs [n_u+n_p] host, x [n_u+n_p] host: guess in, solution out. The caller
 * zeroes the constrained entries of the guess before the call and distributes
 * the constraints after it, as the reference does around its GMRES call
 * (stokes_ilu.cc:523-524, :545). Returns FS_OK, FS_NOT_CONVERGED or an error. */
int fs_stokes_solve(fs_handle* h, const double* rhs, double* x, fs_stats* stats);

/* Same, with rhs and x already resident in device memory (FP64). */
int fs_stokes_solve_device(fs_handle* h, const double* d_rhs, double* d_x, fs_stats* stats);

/* K x (BlockSparseMatrix::vmult), host vectors [n_u+n_p]. */
int fs_stokes_vmult(fs_handle* h, const double* src, double* dst, int32_t precision);

/* BlockSchurPreconditionerBase::vmult (linear_operators.implement.h:200-209):
 * dst = P^{-1} src with the current factorisation, host vectors [n_u+n_p].
 * The inner CG starts from zero. */
int fs_stokes_precond_vmult(fs_handle* h, const double* src, double* dst, int32_t precision,
                            fs_stats* stats);

/* Initial-guess slot of StokesLinearSystem::solve_gmres (stokes_ilu.cc:523-524)
 * for FS_GUESS_LEARNED: the network `model` (fs_guess.h, not owned; it must
 * outlive the handle's solves) maps log10(eta) on the pressure-node lattice
 * ((N+1)^dim values, x fastest, host array) to (u, p) on the node lattices;
 * vnode/vcomp [n_u] and pnode [n_p] place every dof of the caller's order on
 * them (fs_system_dof_maps in fs_assemble.h). In each solve the guess is the
 * network output at the dofs, with every row that holds only its diagonal
 * (constrained rows) set to rhs/diagonal; if its residual is not below
 * ||rhs|| the solve starts from zero instead (fs_stats.guess_used says which). */
struct fs_guess_model;
int fs_stokes_set_guess_field(fs_handle* h, struct fs_guess_model* model, const float* log_eta,
                              int32_t n_cells, const int64_t* vnode, const int8_t* vcomp,
                              const int64_t* pnode);

/* Velocity hand-off to the level-set side (VelocityExtractor::extract,
 * include/felspa/pde/src/stokes_common.implement.h:582-591, which evaluates the
 * velocity of the solution at the points of an FEValues object):
 * fs_stokes_set_dof_lattice places the caller's velocity dofs on the Q2 node
 * lattice of the uniform mesh of n_cells^dim cells of [0,1]^dim (vnode [n_u]
 * node index, x fastest, on the (2 n_cells + 1)^dim lattice; vcomp [n_u] its
 * component; every node and component exactly once, e.g. fs_system_dof_maps);
 * fs_stokes_velocity_lattice writes the velocity of a device solution vector
 * d_x [n_u + n_p] (caller's order, constraints distributed) to d_u
 * [dim][(2 n_cells + 1)^dim] in device memory; fs_stokes_velocity_at_points
 * evaluates that Q2 field at n_points device points d_points [n_points][dim]
 * (clamped to the unit cube) into d_vel [n_points][dim]. The velocity stays in
 * HBM for a device-side advection step. */
int fs_stokes_set_dof_lattice(fs_handle* h, int32_t dim, int32_t n_cells, const int64_t* vnode,
                              const int8_t* vcomp);
int fs_stokes_velocity_lattice(fs_handle* h, const double* d_x, double* d_u);
int fs_stokes_velocity_at_points(fs_handle* h, const double* d_u, int64_t n_points, const double* d_points,
                                 double* d_vel);

/* Switch the solution method (fs_opts.method) of an existing handle between
 * solves: StokesSolutionMethod FC / SCR (stokes_common.h:323-328) on the same
 * factorisation; CompareTest (stokes_ilu.cc:471-505) is two solves and a
 * difference, done by the caller. */
int fs_stokes_set_method(fs_handle* h, int32_t method);

/* Switch the precision (fs_opts.precision, FS_PREC_*) of an existing handle
 * between solves: every handle holds FP64 and FP32 copies of its matrices and
 * factors, so one symbolic phase and one factor serve the plain FP64 and the
 * mixed solve (Control::precision has no counterpart in the reference, whose
 * solve is FP64 only). Reference-mode GMRES and SCR stay FP64-only. */
int fs_stokes_set_precision(fs_handle* h, int32_t precision);

/* Switch the guess mode (fs_opts.guess) of an existing handle between solves,
 * e.g. to compare FS_GUESS_LEARNED with FS_GUESS_ZERO on the same matrices. */
int fs_stokes_set_guess_mode(fs_handle* h, int32_t guess);

/* Change the Krylov parameters of an existing handle between solves (the
 * reference fixes them when it builds its solver objects: SolverGMRES's
 * max_n_tmp_vectors, stokes_ilu.cc:538, and InverseMatrixCG's control,
 * linear_operators.implement.h:76): restart (at most the basis allocated at
 * fs_stokes_create, i.e. the restart given there or the methods' defaults),
 * inner_a_iters, and the inner CG's tolerance and step cap in FP64 and in
 * FS_PREC_MIXED. A negative value keeps the current setting. */
int fs_stokes_set_krylov(fs_handle* h, int32_t restart, int32_t inner_a_iters, double inner_cg_tol,
                         int32_t inner_cg_max, double inner_cg_tol_mixed, int32_t inner_cg_max_mixed);

/* Row permutation used by FS_ORDER_BLOCK_COLOR for the A (which=0) and
 * M_p (which=1) blocks: perm[new] = old. */
int fs_stokes_get_ordering(fs_handle* h, int32_t which, int32_t* perm, int32_t* n_stages);

/* The block / stage schedule of that ordering (solver order): blk_ptr
 * [n_blocks+1] row ranges of the blocks, stage_ptr [n_stages+1] block ranges
 * of the colour stages. Pass NULL arrays to query the sizes. */
int fs_stokes_get_schedule(fs_handle* h, int32_t which, int32_t* blk_ptr, int32_t* n_blocks, int32_t* stage_ptr,
                           int32_t* n_stages);

/* Size of the ILU(k) factor of A (which=0) or M_p (which=1): rows, stored
 * entries (diagonal included), the numeric factorisation's work (the sum over
 * L entries (i, k) of the length of row k's U part, which IKJ streams once
 * per such entry) and the number of colour stages. Query only; no device
 * work. */
int fs_stokes_get_factor_info(fs_handle* h, int32_t which, int64_t* n, int64_t* nnz, int64_t* ikj_updates,
                              int32_t* n_stages);

/* Kernel timing for roofline reports: runs `reps` launches of one kernel
 * class on the handle's own stream between HIP events, with the current
 * factorisation and solver-order data, and returns the mean duration of one
 * launch (ms) and its algorithmic HBM bytes. kind: 0 = A-block SpMV of the
 * inner PCG (q = A p with the p.q and r.p partials), 1 = K SpMV with FP64
 * values as the outer Arnoldi step runs it (the input vector in FP32 for
 * FS_PREC_MIXED, FP64 otherwise), 2 = one ILU(k) application on A as the inner PCG runs it (all
 * stage launches, the forward stages carrying the PCG update step), 3 = one
 * inner CG on M_p, 4 = one numeric ILU(k) factorisation of A (FP64, all its
 * launches and the block inverses; precision ignored). */
int fs_stokes_time_kernel(fs_handle* h, int32_t kind, int32_t precision, int32_t reps, double* ms_per_launch,
                          double* bytes_per_launch);

const char* fs_last_error(const fs_handle* h);
void fs_stokes_destroy(fs_handle* h);

/* ---- SparseILU (include/felspa/linear_algebra/ilu.h:25-238) ------------- */
typedef struct fs_ilu fs_ilu;

/* SparseILU::initialize symbolic phase (ilu.implement.h:29-87) on the GPU
 * object; ordering as above (FS_ORDER_REFERENCE keeps the caller's order). */
int fs_ilu_create(const fs_csr* A, int32_t level_of_fill, int32_t ordering, int32_t block_rows,
                  const int32_t* block_hint, fs_ilu** out);
/* numeric_factorize (ilu.implement.h:438-541) on the GPU, values of A in the
 * caller's order/pattern. */
int fs_ilu_factorize(fs_ilu* ilu, const double* A_val);
/* vmult (ilu.implement.h:544-595) / Tvmult (:626-680), host vectors in the
 * caller's order; precision FS_PREC_FP64 or FS_PREC_MIXED (FP32 factors). */
int fs_ilu_vmult(fs_ilu* ilu, const double* src, double* dst, int32_t precision);
/* fs_ilu_vmult on device vectors (FP64, caller's order, on the GPU that holds
 * the factor) that allocates nothing: no reference counterpart; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops). Its two work vectors live in a static device area of
 * 2 x 2,097,152 FP64 values (32 MiB per GPU, held until the process ends):
 * factors of more rows return FS_ERR_UNSUPPORTED. The work area is shared by
 * every call on a GPU, so calls that overlap on one GPU race; the call runs
 * on the factor's stream and returns when it is done. */
int fs_ilu_vmult_noalloc(fs_ilu* ilu, const double* d_src, double* d_dst, int32_t precision);
int fs_ilu_Tvmult(fs_ilu* ilu, const double* src, double* dst, int32_t precision);
/* Factor pattern and values in the factor's (permuted) order, rows sorted,
 * diagonal in place. Pass NULL arrays to query n / nnz. */
int fs_ilu_get_factors(fs_ilu* ilu, int64_t* n, int64_t* nnz, int64_t* row_ptr, int32_t* col_idx,
                       double* val, int32_t* perm);
void fs_ilu_destroy(fs_ilu* ilu);

/* Host symbolic ILU(k) with the reference's level-of-fill rules
 * (ilu.implement.h:90-435); returns a malloc'd pattern the caller frees with
 * fs_free. */
int fs_ilu_symbolic(const fs_csr* A, int32_t level_of_fill, int64_t** row_ptr, int32_t** col_idx);
void fs_free(void* p);

/* Host-only: the block-colour ordering the GPU ILU uses for a square
 * pattern (no device needed). perm[n] (perm[new] = old), blk_ptr[n+1] and
 * stage_ptr[n+1] receive the schedule; *n_blocks / *n_stages its sizes. */
int fs_block_color_ordering(const fs_csr* A, int32_t level_of_fill, int32_t block_rows,
                            const int32_t* block_hint, int32_t* perm, int32_t* blk_ptr,
                            int32_t* n_blocks, int32_t* stage_ptr, int32_t* n_stages);

/* ---- SparseMatrix::vmult on the GPU (CSR SpMV) -------------------------- */
typedef struct fs_matrix fs_matrix;
int fs_matrix_create(const fs_csr* M, fs_matrix** out);
int fs_matrix_vmult(fs_matrix* M, const double* x, double* y, int32_t precision);
void fs_matrix_destroy(fs_matrix* M);

/* device memory cap used for FS_ERR_OUT_OF_MEMORY tests (0: none) */
void fs_set_device_memory_cap(int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
