"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product (felspa_amd/) never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"])


class OrCsr(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64),
                ("rp", C.POINTER(C.c_int64)), ("ci", C.POINTER(C.c_int32)),
                ("v", C.POINTER(C.c_double))]


class OrOpts(C.Structure):
    _fields_ = [("ilu_k_A", C.c_int), ("ilu_k_S", C.c_int), ("restart", C.c_int),
                ("rtol", C.c_double), ("reduce", C.c_double), ("inner_tol", C.c_double),
                ("apply_diagonal_scaling", C.c_int), ("method", C.c_int),
                ("threads", C.c_int), ("max_steps", C.c_int), ("inner_a_iters", C.c_int)]


class OrStats(C.Structure):
    _fields_ = [("status", C.c_int), ("outer_iters", C.c_int), ("inner_iters", C.c_longlong),
                ("last_value", C.c_double), ("t_setup_s", C.c_double), ("t_krylov_s", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB)
        P = C.POINTER
        L.or_spmv.argtypes = [P(OrCsr), P(C.c_double), P(C.c_double)]
        L.or_ilu_create.argtypes = [P(OrCsr), C.c_int, C.c_int, P(C.c_void_p)]
        L.or_ilu_refactor.argtypes = [C.c_void_p, P(OrCsr)]
        L.or_ilu_vmult.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.or_ilu_Tvmult.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.or_ilu_get.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int64),
                                 P(C.c_int32), P(C.c_double)]
        L.or_ilu_destroy.argtypes = [C.c_void_p]
        L.or_stokes_solve.argtypes = [P(OrCsr)] * 5 + [P(OrOpts), P(C.c_double), P(C.c_double),
                                                       P(OrStats)]
        L.or_block_schur_vmult.argtypes = [P(OrCsr), P(OrCsr), P(OrCsr), C.c_int, C.c_int,
                                           C.c_double, P(C.c_double), P(C.c_double),
                                           P(C.c_longlong)]
        L.or_scaling_coeffs.argtypes = [P(OrCsr), P(OrCsr), C.c_int64, P(C.c_double),
                                        P(C.c_double)]
        L.or_inner_pcg_a.argtypes = [P(OrCsr), C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class _Csr:
    """Keeps numpy arrays alive behind an OrCsr struct. `m` is a scipy CSR."""

    def __init__(self, m):
        self.rp = np.ascontiguousarray(m.indptr, dtype=np.int64)
        self.ci = np.ascontiguousarray(m.indices, dtype=np.int32)
        self.v = np.ascontiguousarray(m.data, dtype=np.float64)
        self.s = OrCsr(m.shape[0], m.shape[1], _ptr(self.rp, C.c_int64), _ptr(self.ci, C.c_int32),
                       _ptr(self.v, C.c_double))


def spmv(m, x):
    c = _Csr(m)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(m.shape[0])
    lib().or_spmv(C.byref(c.s), _ptr(x, C.c_double), _ptr(y, C.c_double))
    return y


class SparseILU:
    """SparseILU<double>: initialize / vmult / Tvmult / factors (sorted CSR)."""

    def __init__(self, A, k, force_generic=False):
        self._A = _Csr(A)
        self.h = C.c_void_p()
        self.rc = lib().or_ilu_create(C.byref(self._A.s), int(k), int(force_generic),
                                      C.byref(self.h))
        self.n = A.shape[0]

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_ilu_destroy(self.h)
            self.h = None

    def refactor(self, A):
        self._A = _Csr(A)
        return lib().or_ilu_refactor(self.h, C.byref(self._A.s))

    def vmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().or_ilu_vmult(self.h, _ptr(src, C.c_double), _ptr(dst, C.c_double))
        return dst

    def Tvmult(self, src):
        src = np.ascontiguousarray(src, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().or_ilu_Tvmult(self.h, _ptr(src, C.c_double), _ptr(dst, C.c_double))
        return dst

    def factors(self):
        import scipy.sparse as sp
        n = C.c_int64()
        nnz = C.c_int64()
        lib().or_ilu_get(self.h, C.byref(n), C.byref(nnz), None, None, None)
        rp = np.zeros(n.value + 1, np.int64)
        ci = np.zeros(nnz.value, np.int32)
        v = np.zeros(nnz.value)
        lib().or_ilu_get(self.h, None, None, _ptr(rp, C.c_int64), _ptr(ci, C.c_int32),
                         _ptr(v, C.c_double))
        return sp.csr_matrix((v, ci, rp), shape=(n.value, n.value))


def stokes_solve(blocks, rhs, x0=None, kA=1, kS=0, restart=23, rtol=1e-8, reduce=1e-6,
                 inner_tol=1e-12, scaling=False, method=0, threads=0, max_steps=0, inner_a_iters=30):
    """StokesLinearSystem::solve on the CPU. blocks = dict of scipy CSR
    A, Bt, B, K11, Mp. Returns (x, stats dict)."""
    cs = [_Csr(blocks[k]) for k in ("A", "Bt", "B", "K11", "Mp")]
    n = blocks["A"].shape[0] + blocks["Mp"].shape[0]
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    o = OrOpts(kA, kS, restart, rtol, reduce, inner_tol, int(scaling), method, threads, max_steps,
               inner_a_iters)
    st = OrStats()
    lib().or_stokes_solve(*[C.byref(c.s) for c in cs], C.byref(o), _ptr(rhs, C.c_double),
                          _ptr(x, C.c_double), C.byref(st))
    return x, dict(status=st.status, outer_iters=st.outer_iters, inner_iters=st.inner_iters,
                   last_value=st.last_value, t_setup_s=st.t_setup_s, t_krylov_s=st.t_krylov_s)


def block_schur_vmult(blocks, src, kA=1, kS=0, inner_tol=1e-12):
    cs = [_Csr(blocks[k]) for k in ("A", "B", "Mp")]
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    it = C.c_longlong()
    rc = lib().or_block_schur_vmult(*[C.byref(c.s) for c in cs], kA, kS, inner_tol,
                                    _ptr(src, C.c_double), _ptr(dst, C.c_double), C.byref(it))
    return dst, rc, it.value


def scaling_coeffs(A, Bt, n_p):
    ca, cb = _Csr(A), _Csr(Bt)
    su = np.zeros(A.shape[0])
    sp_ = np.zeros(n_p)
    lib().or_scaling_coeffs(C.byref(ca.s), C.byref(cb.s), n_p, _ptr(su, C.c_double),
                            _ptr(sp_, C.c_double))
    return su, sp_


def inner_pcg_a(A, b, kA=1, iters=30):
    """The flexible method's A^{-1}: iters ILU(kA)-PCG steps from zero."""
    ca = _Csr(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().or_inner_pcg_a(C.byref(ca.s), int(kA), int(iters), _ptr(b, C.c_double), _ptr(x, C.c_double))
    return x
