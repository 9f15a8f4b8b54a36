/*
 * fs_assemble.h -- deterministic Q2-Q1 Stokes system generator (CPU).
 *
 * Produces the same blocks that FELSPA's StokesLinearSystem owns when it calls
 * solve(): A = K(0,0), B^T = K(0,1), B = K(1,0), K11 = K(1,1) (diagonal-only
 * pattern) and the Schur approximation M_p(1/eta) = preconditioner(1,1), plus
 * the right-hand side, for uniform quadrilateral / hexahedral meshes of the
 * unit square / cube.
 *
 * Reference behaviour restated here (FELSPA paths, file:line):
 *   local assembly  2 eta eps(u):eps(v) - q div v - p div u ; M_p = int pq/eta ;
 *                   rhs rho g.v                      src/pde/stokes_ilu.cc:739-821
 *   quadrature      QGauss(degree_v + 1)             include/felspa/pde/src/stokes_common.implement.h:168
 *   block sparsity  all couplings except p-p          include/felspa/pde/src/linear_systems.implement.h:185-220
 *   precond pattern p-p only                          src/pde/stokes_ilu.cc:260-292
 *   constraints     Dirichlet, periodic, no-normal-flux, first pressure dof
 *                   pinned                            src/pde/stokes_ilu.cc:310-403
 *   dof order       Cuthill-McKee, then velocity block / pressure block
 *                                                     include/felspa/pde/src/stokes_common.implement.h:336-340
 *   inclusion eta   SphericalInclusion::scalar_values include/felspa/physics/src/viscous_flow.implement.h:42-72
 *   simple shear    LinearSimpleShear::value          src/pde/boundary_conditions.cc:663-676
 *   top/bottom      TopBottomSimpleShearBC            examples/example_porphyroclast_Bons1997.h:31-84
 *
 * All entry points return 0 on success and a negative FS_* code on failure;
 * fs_last_error() (declared in fs_stokes.h) describes the last failure.
 */
#ifndef FS_ASSEMBLE_H
#define FS_ASSEMBLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* viscosity / density field */
enum {
  FS_FIELD_INCLUSION = 0, /* sphere/disk radius r at centre, eta_incl inside (SphericalInclusion) */
  FS_FIELD_LAYERS = 1,    /* two layers split at y = 0.5 + amp cos(2 pi x) (Rayleigh-Taylor) */
  FS_FIELD_ELLIPSES = 2,  /* 1..n seeded ellipses / ellipsoids, eta = 10^c, overlaps take max c */
  FS_FIELD_CONSTANT = 3   /* eta = 1 everywhere */
};

/* boundary conditions */
enum {
  FS_BC_TOPBOTTOM_SHEAR_PERIODIC = 0, /* u=(+U,0) top, (-U,0) bottom, periodic in x */
  FS_BC_FREE_SLIP = 1,                /* no normal flux on every wall */
  FS_BC_LINEAR_SHEAR = 2              /* u_0 = (x_{d-1}-0.5)/0.5 U, other comps 0, all faces */
};

typedef struct fs_problem {
  int32_t dim;          /* 2 or 3 */
  int32_t n_cells;      /* cells per direction */
  int32_t field;        /* FS_FIELD_* */
  int32_t bc;           /* FS_BC_* */
  double eta_matrix;    /* background viscosity */
  double eta_incl;      /* inclusion / upper-layer viscosity */
  double rho_matrix;    /* background (lower-layer) density */
  double rho_incl;      /* inclusion (upper-layer) density */
  double radius;        /* inclusion radius */
  double amp;           /* layer interface amplitude */
  double gravity;       /* body force = rho * (0,..,-gravity); 0 for none */
  double shear_velocity;/* U */
  uint64_t seed;        /* FS_FIELD_ELLIPSES */
  int32_t n_ellipses_min, n_ellipses_max;
  double axis_min, axis_max;   /* semi-axes U[axis_min, axis_max] */
  double logc_min, logc_max;   /* log10 contrast U[logc_min, logc_max] */
} fs_problem;

typedef struct fs_system fs_system;

/* constraint kinds reported per dof by fs_system_constraints */
enum { FS_DOF_FREE = 0, FS_DOF_DIRICHLET = 1, FS_DOF_PERIODIC = 2, FS_DOF_PINNED = 3 };

typedef struct fs_system_info {
  int32_t dim, n_cells;
  int64_t n_u, n_p;
  int64_t nnz_A, nnz_Bt, nnz_B, nnz_K11, nnz_Mp;
  int64_t n_vnodes_1d, n_pnodes_1d; /* 2N+1, N+1 */
} fs_system_info;

int fs_system_create(const fs_problem* prob, fs_system** out);
void fs_system_destroy(fs_system* sys);
int fs_system_get_info(const fs_system* sys, fs_system_info* info);

/* Borrowed pointers into the system (valid until fs_system_destroy).
 * which: 0=A 1=Bt 2=B 3=K11 4=Mp. Rows sorted by column, diagonal in place. */
int fs_system_csr(const fs_system* sys, int which, const int64_t** row_ptr,
                  const int32_t** col_idx, const double** val);
/* rhs [n_u + n_p] */
const double* fs_system_rhs(const fs_system* sys);
/* per-dof constraint kind / master dof / value, each [n_u + n_p] */
int fs_system_constraints(const fs_system* sys, const int8_t** kind,
                          const int64_t** master, const double** value);
/* AffineConstraints::distribute equivalent: sets constrained entries of x. */
int fs_system_distribute(const fs_system* sys, double* x);
/* Grid maps for fields that live on node lattices:
 *   vnode[i] = lattice index of velocity dof i on the (2N+1)^d lattice,
 *   vcomp[i] = its component, pnode[j] = lattice index of pressure dof j on
 *   the (N+1)^d lattice. */
int fs_system_dof_maps(const fs_system* sys, const int64_t** vnode,
                       const int8_t** vcomp, const int64_t** pnode);
/* log10(eta) sampled at the (N+1)^d pressure-node lattice (x fastest). */
const float* fs_system_log_eta_grid(const fs_system* sys);
/* geometric block hint for the velocity rows: block id of each velocity dof
 * when the node lattice is cut into cubes of `nodes_per_block` nodes per
 * direction. Written into out[n_u]. */
int fs_system_block_hint(const fs_system* sys, int32_t nodes_per_block, int32_t* out);
/* Viscosity and density at the quadrature points the assembly uses
 * (viscous_flow.implement.h:42-72 evaluated at the points as stokes_ilu.cc:762-772 does):
 * eta_q, rho_q [n_cells^dim * 3^dim], cell-major (cells x fastest, then the
 * cell's Gauss points with index (q%3, q/3%3, q/9%3)). */
int fs_system_quadrature_fields(const fs_system* sys, double* eta_q, double* rho_q);

/* ---------------------------------------------------------------------------
 * Device assembly (GPU). StokesAssembler::local_assembly (stokes_ilu.cc:739-821)
 * and copy_local_to_global (stokes_common.implement.h:658-669) with the
 * constraint condensation of the host generator, on the patterns of an
 * existing system: values of A, B^T, B, K11, M_p and the rhs are written to
 * caller device arrays laid out exactly like fs_system_csr's blocks, and are
 * bit-identical to the host generator's for the same quadrature samples.
 * Every time step that changes the viscosity only re-runs
 * fs_assembler_assemble (then fs_stokes_refactor_device and
 * fs_stokes_solve_device): no host assembly and no host-device copy.
 * Calls on one assembler are serialised by the caller; each call completes
 * before it returns. */
typedef struct fs_assembler fs_assembler;
/* uploads the patterns, dof maps and constraints of `sys` (which may be
 * destroyed afterwards) to the current device */
int fs_assembler_create(const fs_system* sys, fs_assembler** out);
/* the system's own viscosity / density field at the quadrature points, as
 * fs_system_quadrature_fields lays them out, into device arrays */
int fs_assembler_sample_field(fs_assembler* a, double* d_eta_q, double* d_rho_q);
/* assembles from device samples d_eta_q, d_rho_q [n_cells^dim * 3^dim] into
 * device value arrays of nnz(A), nnz(Bt), nnz(B), nnz(K11), nnz(Mp) and
 * d_rhs [n_u + n_p] */
int fs_assembler_assemble(fs_assembler* a, const double* d_eta_q, const double* d_rho_q, double* d_A,
                          double* d_Bt, double* d_B, double* d_K11, double* d_Mp, double* d_rhs);
/* GPU time of the last fs_assembler_assemble (ms, HIP events on its stream) */
double fs_assembler_last_ms(const fs_assembler* a);
void fs_assembler_destroy(fs_assembler* a);

#ifdef __cplusplus
}
#endif
#endif
