"""ctypes binding of felspa_amd/libfelspa_amd.so (include/fs_stokes.h,
include/fs_assemble.h, include/fs_guess.h). The library is built in-tree by
felspa_amd.build.build(); nothing here falls back to another implementation:
a missing library raises."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfelspa_amd.so")

P = C.POINTER
i32, i64, f64, f32 = C.c_int32, C.c_int64, C.c_double, C.c_float

FS_OK, FS_NOT_CONVERGED = 0, 1
FS_ERR_BAD_INPUT, FS_ERR_ZERO_PIVOT, FS_ERR_OUT_OF_MEMORY = -1, -2, -3
FS_ERR_HIP, FS_ERR_UNSUPPORTED, FS_ERR_TIMEOUT, FS_ERR_NO_DEVICE = -4, -5, -6, -7
FS_METHOD_GMRES_REF, FS_METHOD_FGMRES, FS_METHOD_SCR = 0, 1, 2
FS_PREC_FP64, FS_PREC_MIXED, FS_PREC_MIXED_FP64_VECTORS = 0, 1, 2
FS_ORDER_REFERENCE, FS_ORDER_BLOCK_COLOR = 0, 1
FS_GUESS_USER, FS_GUESS_ZERO, FS_GUESS_LEARNED = 0, 1, 2
FS_FIELD_INCLUSION, FS_FIELD_LAYERS, FS_FIELD_ELLIPSES, FS_FIELD_CONSTANT = 0, 1, 2, 3
FS_BC_TOPBOTTOM_SHEAR_PERIODIC, FS_BC_FREE_SLIP, FS_BC_LINEAR_SHEAR = 0, 1, 2
FS_DOF_FREE, FS_DOF_DIRICHLET, FS_DOF_PERIODIC, FS_DOF_PINNED = 0, 1, 2, 3


class FsCsr(C.Structure):
    _fields_ = [("n_rows", i64), ("n_cols", i64), ("row_ptr", P(i64)), ("col_idx", P(i32)),
                ("val", P(f64)), ("diag_first", i32)]


class FsOpts(C.Structure):
    _fields_ = [("method", i32), ("precision", i32), ("ordering", i32), ("ilu_k_A", i32),
                ("ilu_k_S", i32), ("restart", i32), ("max_outer", i32), ("rtol", f64),
                ("reduce", f64), ("inner_cg_max", i32), ("inner_cg_tol", f64),
                ("inner_cg_tol_mixed", f64), ("inner_cg_max_mixed", i32),
                ("apply_diagonal_scaling", i32), ("guess", i32), ("block_rows", i32),
                ("block_hint", P(i32)), ("verbose", i32), ("inner_a_iters", i32)]


class FsStats(C.Structure):
    _fields_ = [("status", i32), ("outer_iters", i32), ("restarts", i32), ("guess_used", i32),
                ("inner_iters", i64), ("last_value", f64), ("initial_rel_residual", f64),
                ("true_rel_residual", f64), ("t_setup_ms", f64), ("t_krylov_ms", f64),
                ("t_total_ms", f64), ("bytes_algorithmic", f64), ("n_kernel_launches", i64),
                ("t_guess_ms", f64), ("guess_rel_residual", f64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FsProblem(C.Structure):
    _fields_ = [("dim", i32), ("n_cells", i32), ("field", i32), ("bc", i32),
                ("eta_matrix", f64), ("eta_incl", f64), ("rho_matrix", f64), ("rho_incl", f64),
                ("radius", f64), ("amp", f64), ("gravity", f64), ("shear_velocity", f64),
                ("seed", C.c_uint64), ("n_ellipses_min", i32), ("n_ellipses_max", i32),
                ("axis_min", f64), ("axis_max", f64), ("logc_min", f64), ("logc_max", f64)]


class FsGuessConfig(C.Structure):
    _fields_ = [("dim", i32), ("width", i32), ("modes", i32), ("layers", i32), ("proj", i32),
                ("seed", C.c_uint64)]


class FsSystemInfo(C.Structure):
    _fields_ = [("dim", i32), ("n_cells", i32), ("n_u", i64), ("n_p", i64), ("nnz_A", i64),
                ("nnz_Bt", i64), ("nnz_B", i64), ("nnz_K11", i64), ("nnz_Mp", i64),
                ("n_vnodes_1d", i64), ("n_pnodes_1d", i64)]


# every symbol the public headers declare (checked by the CPU test suite)
EXPORTS = [
    "fs_opts_default", "fs_stokes_create", "fs_stokes_refactor", "fs_stokes_refactor_device", "fs_stokes_solve",
    "fs_stokes_solve_device", "fs_stokes_vmult", "fs_stokes_precond_vmult",
    "fs_stokes_set_guess_field", "fs_stokes_set_guess_mode", "fs_stokes_set_krylov", "fs_stokes_set_precision", "fs_stokes_set_dof_lattice", "fs_stokes_velocity_lattice", "fs_stokes_velocity_at_points", "fs_stokes_set_method", "fs_stokes_get_ordering", "fs_stokes_get_factor_info", "fs_stokes_get_schedule", "fs_stokes_time_kernel", "fs_last_error", "fs_stokes_destroy",
    "fs_ilu_create", "fs_ilu_factorize", "fs_ilu_vmult", "fs_ilu_vmult_noalloc", "fs_ilu_Tvmult", "fs_ilu_get_factors",
    "fs_ilu_destroy", "fs_ilu_symbolic", "fs_block_color_ordering", "fs_free", "fs_matrix_create", "fs_matrix_vmult",
    "fs_matrix_destroy", "fs_set_device_memory_cap",
    "fs_system_create", "fs_system_destroy", "fs_system_get_info", "fs_system_csr",
    "fs_system_rhs", "fs_system_constraints", "fs_system_distribute", "fs_system_dof_maps",
    "fs_system_log_eta_grid", "fs_system_block_hint", "fs_system_quadrature_fields",
    "fs_assembler_create", "fs_assembler_sample_field", "fs_assembler_assemble", "fs_assembler_last_ms",
    "fs_assembler_destroy",
    "fs_blocks_save", "fs_blocks_load", "fs_blocks_get", "fs_blocks_rhs", "fs_blocks_free",
    "fs_guess_model_create", "fs_guess_model_destroy", "fs_guess_model_get_config",
    "fs_guess_model_save", "fs_guess_model_load", "fs_guess_model_num_params",
    "fs_guess_model_get_params", "fs_guess_model_set_params", "fs_guess_model_get_grads",
    "fs_guess_model_set_normalization", "fs_guess_model_get_normalization", "fs_guess_prepare_input",
    "fs_guess_model_forward", "fs_guess_model_loss_grad", "fs_guess_model_adam_step", "fs_gemm_bf16",
    "fs_guess_model_time_forward",
]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("felspa_amd: %s is missing; run felspa_amd.build.build() (or "
                           "__graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    V = C.c_void_p
    sig = {
        "fs_opts_default": (None, [P(FsOpts)]),
        "fs_stokes_create": (i32, [P(FsCsr)] * 5 + [P(FsOpts), P(V)]),
        "fs_stokes_refactor": (i32, [V] + [P(f64)] * 5),
        "fs_stokes_refactor_device": (i32, [V] * 6),
        "fs_stokes_solve": (i32, [V, P(f64), P(f64), P(FsStats)]),
        "fs_stokes_solve_device": (i32, [V, V, V, P(FsStats)]),
        "fs_stokes_vmult": (i32, [V, P(f64), P(f64), i32]),
        "fs_stokes_precond_vmult": (i32, [V, P(f64), P(f64), i32, P(FsStats)]),
        "fs_stokes_set_guess_field": (i32, [V, V, P(f32), i32, P(i64), P(C.c_int8), P(i64)]),
        "fs_stokes_set_guess_mode": (i32, [V, i32]),
        "fs_stokes_set_krylov": (i32, [V, i32, i32, f64, i32, f64, i32]),
        "fs_stokes_set_precision": (i32, [V, i32]),
        "fs_stokes_set_dof_lattice": (i32, [V, i32, i32, P(i64), P(C.c_int8)]),
        "fs_stokes_velocity_lattice": (i32, [V, V, V]),
        "fs_stokes_velocity_at_points": (i32, [V, V, i64, V, V]),
        "fs_stokes_set_method": (i32, [V, i32]),
        "fs_stokes_get_ordering": (i32, [V, i32, P(i32), P(i32)]),
        "fs_stokes_get_factor_info": (i32, [V, i32, P(i64), P(i64), P(i64), P(i32)]),
        "fs_stokes_get_schedule": (i32, [V, i32, P(i32), P(i32), P(i32), P(i32)]),
        "fs_stokes_time_kernel": (i32, [V, i32, i32, i32, P(f64), P(f64)]),
        "fs_last_error": (C.c_char_p, [V]),
        "fs_stokes_destroy": (None, [V]),
        "fs_ilu_create": (i32, [P(FsCsr), i32, i32, i32, P(i32), P(V)]),
        "fs_ilu_factorize": (i32, [V, P(f64)]),
        "fs_ilu_vmult": (i32, [V, P(f64), P(f64), i32]),
        "fs_ilu_vmult_noalloc": (i32, [V, V, V, i32]),
        "fs_ilu_Tvmult": (i32, [V, P(f64), P(f64), i32]),
        "fs_ilu_get_factors": (i32, [V, P(i64), P(i64), P(i64), P(i32), P(f64), P(i32)]),
        "fs_ilu_destroy": (None, [V]),
        "fs_ilu_symbolic": (i32, [P(FsCsr), i32, P(P(i64)), P(P(i32))]),
        "fs_free": (None, [V]),
        "fs_block_color_ordering": (i32, [P(FsCsr), i32, i32, P(i32), P(i32), P(i32), P(i32),
                                          P(i32), P(i32)]),
        "fs_matrix_create": (i32, [P(FsCsr), P(V)]),
        "fs_matrix_vmult": (i32, [V, P(f64), P(f64), i32]),
        "fs_matrix_destroy": (None, [V]),
        "fs_set_device_memory_cap": (None, [i64]),
        "fs_system_create": (i32, [P(FsProblem), P(V)]),
        "fs_system_destroy": (None, [V]),
        "fs_system_get_info": (i32, [V, P(FsSystemInfo)]),
        "fs_system_csr": (i32, [V, i32, P(P(i64)), P(P(i32)), P(P(f64))]),
        "fs_system_rhs": (P(f64), [V]),
        "fs_system_constraints": (i32, [V, P(P(C.c_int8)), P(P(i64)), P(P(f64))]),
        "fs_system_distribute": (i32, [V, P(f64)]),
        "fs_system_dof_maps": (i32, [V, P(P(i64)), P(P(C.c_int8)), P(P(i64))]),
        "fs_system_log_eta_grid": (P(f32), [V]),
        "fs_system_block_hint": (i32, [V, i32, P(i32)]),
        "fs_system_quadrature_fields": (i32, [V, P(f64), P(f64)]),
        "fs_assembler_create": (i32, [V, P(V)]),
        "fs_assembler_sample_field": (i32, [V, V, V]),
        "fs_assembler_assemble": (i32, [V] * 9),
        "fs_assembler_last_ms": (C.c_double, [V]),
        "fs_assembler_destroy": (None, [V]),
        "fs_blocks_save": (i32, [C.c_char_p] + [P(FsCsr)] * 5 + [P(f64)]),
        "fs_blocks_load": (i32, [C.c_char_p, P(V)]),
        "fs_blocks_get": (i32, [V, i32, P(FsCsr)]),
        "fs_blocks_rhs": (P(f64), [V, P(i64), P(i64)]),
        "fs_blocks_free": (None, [V]),
        "fs_guess_model_create": (i32, [P(FsGuessConfig), P(V)]),
        "fs_guess_model_destroy": (None, [V]),
        "fs_guess_model_get_config": (i32, [V, P(FsGuessConfig)]),
        "fs_guess_model_save": (i32, [V, C.c_char_p]),
        "fs_guess_model_load": (i32, [C.c_char_p, P(V)]),
        "fs_guess_model_num_params": (i64, [V]),
        "fs_guess_model_get_params": (i32, [V, P(f32)]),
        "fs_guess_model_set_params": (i32, [V, P(f32)]),
        "fs_guess_model_get_grads": (i32, [V, P(f32)]),
        "fs_guess_model_set_normalization": (i32, [V, P(f32), P(f32)]),
        "fs_guess_model_get_normalization": (i32, [V, P(f32), P(f32)]),
        "fs_guess_prepare_input": (i32, [V, i32, i32, V, V]),
        "fs_guess_model_forward": (i32, [V, i32, i32, V, V]),
        "fs_guess_model_loss_grad": (i32, [V, i32, i32, V, V, P(f64)]),
        "fs_guess_model_adam_step": (i32, [V, f32, f32, f32, f32, f32]),
        "fs_gemm_bf16": (i32, [i32, i32, i32, i32, V, i32, i64, i64, i64, V, i32, i64, i64, i64,
                               V, i64, i64, i64, f32, f32]),
        "fs_guess_model_time_forward": (i32, [V, i32, i32, i32, P(f64), P(f64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    e = lib().fs_last_error(None)
    return e.decode() if e else ""


class FsException(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__("%s failed with code %d: %s" % (where, code, last_error()))


def check(rc, where):
    if rc < 0:
        raise FsException(rc, where)
    return rc
