"""Host-side mirror of FELSPA's Stokes solver interface over the C-ABI.

  StokesSolverControl   include/felspa/pde/stokes_common.h:338-383
  StokesLinearSystem    include/felspa/pde/stokes_ilu.h:130-335 (solve path)
  SparseILU             include/felspa/linear_algebra/ilu.h:25-283
  SparseMatrix.vmult    dealii::SparseMatrix::vmult

Same names, argument meaning and error behaviour as the reference:
solve(soln, rhs, control) updates soln in place and appends the solver
statistics to the control; a solve that misses its tolerance raises
NoConvergence (deal.II SolverControl::NoConvergence)."""
import ctypes as C
import time

import numpy as np

from . import _lib as L

FC, SCR, CompareTest = 0, 1, 2  # StokesSolutionMethod (stokes_common.h:323-328)


class NoConvergence(RuntimeError):
    def __init__(self, last_step, last_value):
        self.last_step, self.last_value = last_step, last_value
        super().__init__("Iterative method reported convergence failure in step %d. The residual "
                         "in the last step was %g." % (last_step, last_value))


class StokesSolverControl:
    """ReductionControl(n, tol, reduce) + the statistics FELSPA records."""

    def __init__(self, n=100, tol=1.0e-10, reduce=1.0e-6):
        self.max_steps, self.tol, self.reduce = n, tol, reduce
        self.apply_diagonal_scaling = False
        self.log_cg = False
        self.log_gmres = False
        self.cg_error, self.cg_timer, self.n_cg_inner_iter, self.n_cg_outer_iter = [], [], [], []
        self.gmres_error, self.gmres_timer, self.n_gmres_iter = [], [], []
        self.soln_diff_l2, self.soln_diff_linfty = [], []
        self.last_step, self.last_value = 0, 0.0
        self.last_stats = None

    def write_statistics(self, filename):
        """StokesSolverControl::write_statistics (src/pde/stokes_common.cc:53-90)."""
        n = max(len(self.cg_error), len(self.gmres_error))
        with open(filename + ".csv", "w") as f:
            if self.log_cg:
                f.write("cg_error,cg_timer,n_cg_inner_iter,n_cg_outer_iter,")
            if self.log_gmres:
                f.write("gmres_error,gmres_timer,n_gmres_iter,")
            if self.log_cg and self.log_gmres:
                f.write("soln_diff_l2,soln_diff_linfty,")
            f.write("\n")
            for i in range(n):
                if self.log_cg:
                    f.write("%g,%g,%d,%d," % (self.cg_error[i], self.cg_timer[i],
                                              self.n_cg_inner_iter[i], self.n_cg_outer_iter[i]))
                if self.log_gmres:
                    f.write("%g,%g,%d," % (self.gmres_error[i], self.gmres_timer[i],
                                           self.n_gmres_iter[i]))
                if self.log_cg and self.log_gmres:
                    f.write("%g,%g," % (self.soln_diff_l2[i], self.soln_diff_linfty[i]))
                f.write("\n")


def _csr_struct(m, keep):
    rp = np.ascontiguousarray(m.indptr, dtype=np.int64)
    ci = np.ascontiguousarray(m.indices, dtype=np.int32)
    v = np.ascontiguousarray(m.data, dtype=np.float64)
    keep.extend([rp, ci, v])
    return L.FsCsr(m.shape[0], m.shape[1], rp.ctypes.data_as(L.P(L.i64)),
                   ci.ctypes.data_as(L.P(L.i32)), v.ctypes.data_as(L.P(L.f64)), 0)


def default_opts():
    o = L.FsOpts()
    L.lib().fs_opts_default(C.byref(o))
    return o


class StokesLinearSystem:
    """The blocks of K = [[A, Bt], [B, K11]] and preconditioner(1,1) = Mp on
    the GPU, with the block-Schur preconditioner and its ILU objects.

    blocks: dict of scipy CSR A, Bt, B, K11, Mp (caller's dof order).
    constraints: optional boolean mask of constrained dofs (their guess entries
    are zeroed before the solve, stokes_ilu.cc:523-524)."""

    def __init__(self, blocks, A_lof=1, S_lof=0, method=L.FS_METHOD_FGMRES,
                 precision=L.FS_PREC_FP64, ordering=L.FS_ORDER_BLOCK_COLOR, block_hint=None,
                 block_rows=0, restart=23, rtol=1e-8, reduce=1e-6, inner_cg_tol=1e-12,
                 inner_cg_tol_mixed=1e-2, inner_cg_max_mixed=50, inner_cg_max=0, guess=L.FS_GUESS_USER,
                 constraints=None, max_outer=0, inner_a_iters=30, apply_diagonal_scaling=False, verbose=0):
        self._lib = L.lib()
        self.n_u = blocks["A"].shape[0]
        self.n_p = blocks["Mp"].shape[0]
        self.n = self.n_u + self.n_p
        self.constraints = constraints
        o = default_opts()
        o.method, o.precision, o.ordering = method, precision, ordering
        o.ilu_k_A, o.ilu_k_S, o.restart, o.rtol, o.reduce = A_lof, S_lof, restart, rtol, reduce
        o.inner_cg_tol, o.inner_cg_tol_mixed, o.inner_cg_max_mixed = (inner_cg_tol, inner_cg_tol_mixed,
                                                                       inner_cg_max_mixed)
        o.guess, o.block_rows, o.max_outer = guess, block_rows, max_outer
        o.inner_a_iters = inner_a_iters
        o.inner_cg_max = inner_cg_max
        o.apply_diagonal_scaling = int(apply_diagonal_scaling)
        o.verbose = int(verbose)
        self._hint = None
        if block_hint is not None:
            self._hint = np.ascontiguousarray(block_hint, dtype=np.int32)
            o.block_hint = self._hint.ctypes.data_as(L.P(L.i32))
        self.opts = o
        keep = []
        cs = [_csr_struct(blocks[k], keep) for k in ("A", "Bt", "B", "K11", "Mp")]
        self.h = C.c_void_p()
        t0 = time.perf_counter()
        L.check(self._lib.fs_stokes_create(*[C.byref(c) for c in cs], C.byref(o), C.byref(self.h)),
                "fs_stokes_create")
        self.t_create_s = time.perf_counter() - t0

    def close(self):
        """Release the handle's device memory now (also done on garbage
        collection); the object is unusable afterwards."""
        if getattr(self, "h", None):
            self._lib.fs_stokes_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def refactor(self, blocks):
        """New values on the same sparsity (use_previous_sparsity = true)."""
        arrs = [np.ascontiguousarray(blocks[k].data, dtype=np.float64)
                for k in ("A", "Bt", "B", "K11", "Mp")]
        L.check(self._lib.fs_stokes_refactor(self.h, *[a.ctypes.data_as(L.P(L.f64)) for a in arrs]),
                "fs_stokes_refactor")

    def refactor_device(self, d_ptrs):
        """New values already on the device: d_ptrs = raw pointers of the A,
        Bt, B, K11, Mp value arrays (the layouts given at create)."""
        L.check(self._lib.fs_stokes_refactor_device(self.h, *[C.c_void_p(int(p)) for p in d_ptrs]),
                "fs_stokes_refactor_device")

    def solve(self, soln, rhs, control=None, solution_method=None):
        """StokesLinearSystem::solve(soln, rhs, solver_control) (stokes_ilu.cc:442-468).

        solution_method (StokesSolutionMethod, stokes_common.h:323-328): None
        keeps the handle's method; FC runs the handle's GMRES method (reference
        mode or FGMRES), SCR the Schur-complement reduction, CompareTest both
        and records their difference in the control (stokes_ilu.cc:471-505)."""
        if solution_method is None:
            return self._solve(soln, rhs, control)
        if solution_method == CompareTest:
            return self.solve_compare_test(soln, rhs, control)
        fc = self._fc_method()
        self.set_method(L.FS_METHOD_SCR if solution_method == SCR else fc)
        try:
            return self._solve(soln, rhs, control)
        finally:
            self.set_method(fc)

    def _fc_method(self):
        m = getattr(self, "_fc", None)
        if m is None:
            m = self._fc = (self.opts.method if self.opts.method != L.FS_METHOD_SCR else L.FS_METHOD_GMRES_REF)
        return m

    def set_method(self, method):
        L.check(self._lib.fs_stokes_set_method(self.h, int(method)), "fs_stokes_set_method")
        self.opts.method = method

    def solve_compare_test(self, soln, rhs, control=None):
        """CG Schur-complement reduction into soln, the GMRES block solve from
        the same guess, and the relative L2 / L-infinity difference appended to
        the control (stokes_ilu.cc:471-505)."""
        fc = self._fc_method()
        soln_gmres = soln.copy()
        self.set_method(L.FS_METHOD_SCR)
        try:
            st_cg = self._solve(soln, rhs, control)
        finally:
            self.set_method(fc)
        st_gm = self._solve(soln_gmres, rhs, control)
        diff = soln_gmres - soln
        l2 = float(np.linalg.norm(diff) / np.linalg.norm(soln))
        linf = float(np.abs(diff).max())
        if control is not None:
            control.soln_diff_l2.append(l2)
            control.soln_diff_linfty.append(linf)
        return dict(cg=st_cg, gmres=st_gm, soln_diff_l2=l2, soln_diff_linfty=linf)

    def _solve(self, soln, rhs, control=None):
        assert soln.dtype == np.float64 and soln.flags.c_contiguous and soln.size == self.n
        if self.constraints is not None:
            soln[self.constraints] = 0.0
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        st = L.FsStats()
        t0 = time.perf_counter()
        rc = self._lib.fs_stokes_solve(self.h, rhs.ctypes.data_as(L.P(L.f64)),
                                       soln.ctypes.data_as(L.P(L.f64)), C.byref(st))
        wall = time.perf_counter() - t0
        L.check(rc, "fs_stokes_solve")
        if control is not None:
            if self.opts.method == L.FS_METHOD_SCR:
                control.n_cg_outer_iter.append(st.outer_iters)
                control.n_cg_inner_iter.append(st.inner_iters)
                control.cg_error.append(st.last_value)
                control.cg_timer.append(wall)
                control.log_cg = True
            else:
                control.n_gmres_iter.append(st.outer_iters)
                control.gmres_error.append(st.last_value)
                control.gmres_timer.append(wall)
                control.log_gmres = True
            control.last_step, control.last_value = st.outer_iters, st.last_value
            control.last_stats = st.as_dict()
        self.last_stats = st.as_dict()
        if rc == L.FS_NOT_CONVERGED:
            raise NoConvergence(st.outer_iters, st.last_value)
        return self.last_stats

    def solve_device(self, d_rhs_ptr, d_x_ptr):
        """Solve with FP64 rhs / x already resident on the GPU (raw device pointers)."""
        st = L.FsStats()
        rc = self._lib.fs_stokes_solve_device(self.h, C.c_void_p(d_rhs_ptr), C.c_void_p(d_x_ptr),
                                              C.byref(st))
        L.check(rc, "fs_stokes_solve_device")
        self.last_stats = st.as_dict()
        return rc, self.last_stats

    def set_guess_field(self, model, system):
        """FS_GUESS_LEARNED input: the network `model` (felspa_amd.guess.GuessModel,
        kept alive here) and the system's log10(eta) field and dof lattice maps."""
        le = np.ascontiguousarray(system.log_eta, dtype=np.float32)
        vn = np.ascontiguousarray(system.vnode, dtype=np.int64)
        vc = np.ascontiguousarray(system.vcomp, dtype=np.int8)
        pn = np.ascontiguousarray(system.pnode, dtype=np.int64)
        L.check(self._lib.fs_stokes_set_guess_field(self.h, model.h, le.ctypes.data_as(L.P(L.f32)),
                                                    int(system.n_cells), vn.ctypes.data_as(L.P(L.i64)),
                                                    vc.ctypes.data_as(L.P(C.c_int8)),
                                                    pn.ctypes.data_as(L.P(L.i64))),
                "fs_stokes_set_guess_field")
        self._guess_model = model

    def set_guess_mode(self, guess):
        L.check(self._lib.fs_stokes_set_guess_mode(self.h, int(guess)), "fs_stokes_set_guess_mode")
        self.opts.guess = guess

    def set_krylov(self, restart=-1, inner_a_iters=-1, inner_cg_tol=-1.0, inner_cg_max=-1,
                   inner_cg_tol_mixed=-1.0, inner_cg_max_mixed=-1):
        """Krylov parameters between solves (negative: unchanged); restart is
        bounded by the basis allocated when the handle was made."""
        L.check(self._lib.fs_stokes_set_krylov(self.h, int(restart), int(inner_a_iters), float(inner_cg_tol),
                                               int(inner_cg_max), float(inner_cg_tol_mixed),
                                               int(inner_cg_max_mixed)), "fs_stokes_set_krylov")
        for k, v in (("restart", restart), ("inner_a_iters", inner_a_iters), ("inner_cg_tol", inner_cg_tol),
                     ("inner_cg_max", inner_cg_max), ("inner_cg_tol_mixed", inner_cg_tol_mixed),
                     ("inner_cg_max_mixed", inner_cg_max_mixed)):
            if v >= 0:
                setattr(self.opts, k, v)

    def factor_info(self, which=0):
        """ILU factor of A (0) or M_p (1): dict n, nnz, ikj_updates, n_stages."""
        n, nnz, upd, ns = L.i64(), L.i64(), L.i64(), L.i32()
        L.check(self._lib.fs_stokes_get_factor_info(self.h, which, C.byref(n), C.byref(nnz), C.byref(upd),
                                                    C.byref(ns)), "fs_stokes_get_factor_info")
        return dict(n=n.value, nnz=nnz.value, ikj_updates=upd.value, n_stages=ns.value)

    def schedule(self, which=0):
        """(blk_ptr, stage_ptr) of the ILU schedule of A (0) or M_p (1)."""
        nb, ns = L.i32(), L.i32()
        L.check(self._lib.fs_stokes_get_schedule(self.h, which, None, C.byref(nb), None, C.byref(ns)),
                "fs_stokes_get_schedule")
        blk = np.zeros(nb.value + 1, np.int32)
        stg = np.zeros(ns.value + 1, np.int32)
        L.check(self._lib.fs_stokes_get_schedule(self.h, which, blk.ctypes.data_as(L.P(L.i32)), None,
                                                 stg.ctypes.data_as(L.P(L.i32)), None), "fs_stokes_get_schedule")
        return blk, stg

    def set_dof_lattice(self, system):
        """The caller's velocity dofs on the Q2 node lattice of `system`
        (felspa_amd.StokesSystem: its dimension, cells and dof maps)."""
        self._vnode = np.ascontiguousarray(system.vnode, np.int64)
        self._vcomp = np.ascontiguousarray(system.vcomp, np.int8)
        self.lattice = (system.dim, system.n_cells)
        L.check(self._lib.fs_stokes_set_dof_lattice(self.h, system.dim, system.n_cells,
                                                    self._vnode.ctypes.data_as(L.P(L.i64)),
                                                    self._vcomp.ctypes.data_as(L.P(C.c_int8))),
                "fs_stokes_set_dof_lattice")

    def velocity_lattice(self, d_x_ptr, d_u_ptr):
        """VelocityExtractor: the velocity of device solution d_x on the node lattice (device d_u)."""
        L.check(self._lib.fs_stokes_velocity_lattice(self.h, C.c_void_p(d_x_ptr), C.c_void_p(d_u_ptr)),
                "fs_stokes_velocity_lattice")

    def velocity_at_points(self, d_u_ptr, n_points, d_points_ptr, d_vel_ptr):
        """The Q2 velocity field d_u at n_points device points (dim values each)."""
        L.check(self._lib.fs_stokes_velocity_at_points(self.h, C.c_void_p(d_u_ptr), int(n_points),
                                                       C.c_void_p(d_points_ptr), C.c_void_p(d_vel_ptr)),
                "fs_stokes_velocity_at_points")

    def set_precision(self, precision):
        """FS_PREC_* for the next solves on the same factorisation."""
        L.check(self._lib.fs_stokes_set_precision(self.h, int(precision)), "fs_stokes_set_precision")
        self.opts.precision = precision

    def vmult(self, x, precision=L.FS_PREC_FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        L.check(self._lib.fs_stokes_vmult(self.h, x.ctypes.data_as(L.P(L.f64)),
                                          y.ctypes.data_as(L.P(L.f64)), precision), "fs_stokes_vmult")
        return y

    def precondition(self, src, precision=L.FS_PREC_FP64):
        """BlockSchurPreconditionerBase::vmult with the current factorisation."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.n)
        st = L.FsStats()
        rc = self._lib.fs_stokes_precond_vmult(self.h, src.ctypes.data_as(L.P(L.f64)),
                                               dst.ctypes.data_as(L.P(L.f64)), precision,
                                               C.byref(st))
        L.check(rc, "fs_stokes_precond_vmult")
        return dst, st.as_dict()

    def time_kernel(self, kind, precision=L.FS_PREC_FP64, reps=20):
        """(ms per launch, algorithmic bytes per launch) of one kernel class,
        timed with HIP events on the solver's stream."""
        ms, by = L.f64(), L.f64()
        L.check(self._lib.fs_stokes_time_kernel(self.h, kind, precision, reps, C.byref(ms), C.byref(by)),
                "fs_stokes_time_kernel")
        return ms.value, by.value

    def ordering(self, which=0):
        n = self.n_u if which == 0 else self.n_p
        perm = np.zeros(n, np.int32)
        ns = L.i32()
        L.check(self._lib.fs_stokes_get_ordering(self.h, which, perm.ctypes.data_as(L.P(L.i32)),
                                                 C.byref(ns)), "fs_stokes_get_ordering")
        return perm, ns.value


class SparseILU:
    """SparseILU<double> on the GPU."""

    class AdditionalData:
        def __init__(self, use_previous_sparsity=False, level_of_fill=0, block_size=1000):
            self.use_previous_sparsity = use_previous_sparsity
            self.max_level_of_fill = level_of_fill
            self.block_size_per_thread = block_size

    def __init__(self, ordering=L.FS_ORDER_REFERENCE, block_rows=0, block_hint=None):
        self._lib = L.lib()
        self.h = None
        self.ordering, self.block_rows = ordering, block_rows
        self._hint = None if block_hint is None else np.ascontiguousarray(block_hint, np.int32)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.fs_ilu_destroy(self.h)
            self.h = None

    def initialize(self, matrix, additional_data):
        if self.h is None or not additional_data.use_previous_sparsity:
            if self.h is not None:
                self._lib.fs_ilu_destroy(self.h)
            keep = []
            c = _csr_struct(matrix, keep)
            self.h = C.c_void_p()
            hint = None if self._hint is None else self._hint.ctypes.data_as(L.P(L.i32))
            L.check(self._lib.fs_ilu_create(C.byref(c), additional_data.max_level_of_fill,
                                            self.ordering, self.block_rows, hint, C.byref(self.h)),
                    "fs_ilu_create")
            self.n = matrix.shape[0]
        v = np.ascontiguousarray(matrix.data, dtype=np.float64)
        return L.check(self._lib.fs_ilu_factorize(self.h, v.ctypes.data_as(L.P(L.f64))),
                       "fs_ilu_factorize")

    def _apply(self, fn, src, precision):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros(self.n)
        L.check(fn(self.h, src.ctypes.data_as(L.P(L.f64)), dst.ctypes.data_as(L.P(L.f64)), precision),
                "fs_ilu_vmult")
        return dst

    def vmult(self, src, precision=L.FS_PREC_FP64):
        return self._apply(self._lib.fs_ilu_vmult, src, precision)

    def Tvmult(self, src, precision=L.FS_PREC_FP64):
        return self._apply(self._lib.fs_ilu_Tvmult, src, precision)

    def vmult_device_noalloc(self, d_src_ptr, d_dst_ptr, precision=L.FS_PREC_FP64):
        """vmult on FP64 device vectors (raw pointers) with the library's static
        work area: no device allocation."""
        L.check(self._lib.fs_ilu_vmult_noalloc(self.h, C.c_void_p(d_src_ptr), C.c_void_p(d_dst_ptr), precision),
                "fs_ilu_vmult_noalloc")

    def factors(self, values=True):
        """(scipy CSR of the factor in its own row order, perm[new] = old)."""
        import scipy.sparse as sp
        n, nnz = L.i64(), L.i64()
        L.check(self._lib.fs_ilu_get_factors(self.h, C.byref(n), C.byref(nnz), None, None, None,
                                             None), "fs_ilu_get_factors")
        rp = np.zeros(n.value + 1, np.int64)
        ci = np.zeros(nnz.value, np.int32)
        v = np.zeros(nnz.value) if values else None
        perm = np.zeros(n.value, np.int32)
        L.check(self._lib.fs_ilu_get_factors(
            self.h, None, None, rp.ctypes.data_as(L.P(L.i64)), ci.ctypes.data_as(L.P(L.i32)),
            v.ctypes.data_as(L.P(L.f64)) if values else None, perm.ctypes.data_as(L.P(L.i32))),
            "fs_ilu_get_factors")
        data = v if values else np.ones(nnz.value)
        return sp.csr_matrix((data, ci, rp), shape=(n.value, n.value)), perm


def ilu_symbolic(matrix, level_of_fill):
    """Host symbolic ILU(k) pattern (scipy CSR of ones)."""
    import scipy.sparse as sp
    lib = L.lib()
    keep = []
    c = _csr_struct(matrix, keep)
    rp, ci = L.P(L.i64)(), L.P(L.i32)()
    L.check(lib.fs_ilu_symbolic(C.byref(c), level_of_fill, C.byref(rp), C.byref(ci)),
            "fs_ilu_symbolic")
    n = matrix.shape[0]
    R = np.ctypeslib.as_array(rp, (n + 1,)).copy()
    nnz = int(R[-1])
    Ci = np.ctypeslib.as_array(ci, (max(nnz, 1),))[:nnz].copy()
    lib.fs_free(C.cast(rp, C.c_void_p))
    lib.fs_free(C.cast(ci, C.c_void_p))
    return sp.csr_matrix((np.ones(nnz), Ci, R), shape=(n, n))


class SparseMatrix:
    """dealii::SparseMatrix<double>::vmult on the GPU."""

    def __init__(self, matrix):
        self._lib = L.lib()
        keep = []
        c = _csr_struct(matrix, keep)
        self.h = C.c_void_p()
        self.shape = matrix.shape
        L.check(self._lib.fs_matrix_create(C.byref(c), C.byref(self.h)), "fs_matrix_create")

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.fs_matrix_destroy(self.h)
            self.h = None

    def vmult(self, x, precision=L.FS_PREC_FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape[0])
        L.check(self._lib.fs_matrix_vmult(self.h, x.ctypes.data_as(L.P(L.f64)),
                                          y.ctypes.data_as(L.P(L.f64)), precision),
                "fs_matrix_vmult")
        return y


def block_color_ordering(matrix, level_of_fill, block_rows=0, block_hint=None):
    """The GPU ILU's block-colour schedule, computed on the host:
    (perm[new] = old, blk_ptr, stage_ptr)."""
    lib = L.lib()
    keep = []
    c = _csr_struct(matrix, keep)
    n = matrix.shape[0]
    perm = np.zeros(n, np.int32)
    blk = np.zeros(n + 1, np.int32)
    stg = np.zeros(n + 1, np.int32)
    nb, ns = L.i32(), L.i32()
    hint = None
    if block_hint is not None:
        hint_arr = np.ascontiguousarray(block_hint, np.int32)
        keep.append(hint_arr)
        hint = hint_arr.ctypes.data_as(L.P(L.i32))
    L.check(lib.fs_block_color_ordering(C.byref(c), level_of_fill, block_rows, hint,
                                        perm.ctypes.data_as(L.P(L.i32)),
                                        blk.ctypes.data_as(L.P(L.i32)), C.byref(nb),
                                        stg.ctypes.data_as(L.P(L.i32)), C.byref(ns)),
            "fs_block_color_ordering")
    return perm, blk[:nb.value + 1], stg[:ns.value + 1]
