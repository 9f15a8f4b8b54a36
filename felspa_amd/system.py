"""Q2-Q1 Stokes systems from the in-library generator (include/fs_assemble.h):
blocks A, Bt, B, K11, Mp as scipy CSR, rhs, constraints and lattice maps."""
import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

from . import _lib as L

BLOCK_NAMES = ("A", "Bt", "B", "K11", "Mp")


def problem(dim, n_cells, field, bc, eta_matrix=1.0, eta_incl=1e3, rho_matrix=1.0, rho_incl=1.0,
            radius=0.1, amp=0.02, gravity=0.0, shear_velocity=1.0, seed=0, n_ellipses=(1, 8),
            axes=(0.05, 0.2), logc=(0.0, 3.0)):
    return L.FsProblem(dim, n_cells, field, bc, eta_matrix, eta_incl, rho_matrix, rho_incl, radius,
                       amp, gravity, shear_velocity, seed, n_ellipses[0], n_ellipses[1], axes[0],
                       axes[1], logc[0], logc[1])


class StokesSystem:
    """Owns an fs_system; exposes copies of its arrays."""

    def __init__(self, prob):
        self._lib = L.lib()
        self.h = C.c_void_p()
        L.check(self._lib.fs_system_create(C.byref(prob), C.byref(self.h)), "fs_system_create")
        info = L.FsSystemInfo()
        L.check(self._lib.fs_system_get_info(self.h, C.byref(info)), "fs_system_get_info")
        self.info = info
        self.dim, self.n_cells = info.dim, info.n_cells
        self.n_u, self.n_p = info.n_u, info.n_p
        self.n = self.n_u + self.n_p
        shapes = [(self.n_u, self.n_u), (self.n_u, self.n_p), (self.n_p, self.n_u),
                  (self.n_p, self.n_p), (self.n_p, self.n_p)]
        self.blocks = {}
        for w, name in enumerate(BLOCK_NAMES):
            rp, ci, v = L.P(L.i64)(), L.P(L.i32)(), L.P(L.f64)()
            L.check(self._lib.fs_system_csr(self.h, w, C.byref(rp), C.byref(ci), C.byref(v)),
                    "fs_system_csr")
            m, n = shapes[w]
            nnz = rp[m]
            R = np.ctypeslib.as_array(rp, (m + 1,)).copy()
            Ci = np.ctypeslib.as_array(ci, (max(nnz, 1),))[:nnz].copy()
            Vv = np.ctypeslib.as_array(v, (max(nnz, 1),))[:nnz].copy()
            self.blocks[name] = sp.csr_matrix((Vv, Ci, R), shape=(m, n))
        self.rhs = np.ctypeslib.as_array(self._lib.fs_system_rhs(self.h), (self.n,)).copy()
        k, ma, va = L.P(C.c_int8)(), L.P(L.i64)(), L.P(L.f64)()
        self._lib.fs_system_constraints(self.h, C.byref(k), C.byref(ma), C.byref(va))
        self.kind = np.ctypeslib.as_array(k, (self.n,)).copy()
        self.master = np.ctypeslib.as_array(ma, (self.n,)).copy()
        self.cval = np.ctypeslib.as_array(va, (self.n,)).copy()
        vn, vc, pn = L.P(L.i64)(), L.P(C.c_int8)(), L.P(L.i64)()
        self._lib.fs_system_dof_maps(self.h, C.byref(vn), C.byref(vc), C.byref(pn))
        self.vnode = np.ctypeslib.as_array(vn, (self.n_u,)).copy()
        self.vcomp = np.ctypeslib.as_array(vc, (self.n_u,)).copy()
        self.pnode = np.ctypeslib.as_array(pn, (self.n_p,)).copy()
        npn = (self.n_cells + 1) ** self.dim
        self.log_eta = np.ctypeslib.as_array(self._lib.fs_system_log_eta_grid(self.h), (npn,)).copy()

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.fs_system_destroy(self.h)
            self.h = None

    @property
    def constrained(self):
        return self.kind != L.FS_DOF_FREE

    def K(self):
        b = self.blocks
        return sp.bmat([[b["A"], b["Bt"]], [b["B"], b["K11"]]]).tocsr()

    def distribute(self, x):
        """AffineConstraints::distribute: constrained entries from the constraints."""
        x = np.array(x, dtype=np.float64)
        self._lib.fs_system_distribute(self.h, x.ctypes.data_as(L.P(L.f64)))
        return x

    @property
    def n_cells_total(self):
        return self.n_cells ** self.dim

    def quadrature_fields(self):
        """(eta_q, rho_q): the field at the assembly's quadrature points,
        [n_cells^dim, 3^dim] each (cells x fastest)."""
        nq = 3 ** self.dim
        eta = np.zeros((self.n_cells_total, nq))
        rho = np.zeros((self.n_cells_total, nq))
        L.check(self._lib.fs_system_quadrature_fields(self.h, eta.ctypes.data_as(L.P(L.f64)),
                                                      rho.ctypes.data_as(L.P(L.f64))),
                "fs_system_quadrature_fields")
        return eta, rho

    def block_hint(self, nodes_per_block):
        out = np.zeros(self.n_u, np.int32)
        L.check(self._lib.fs_system_block_hint(self.h, int(nodes_per_block),
                                                out.ctypes.data_as(L.P(L.i32))),
                "fs_system_block_hint")
        return out


class DeviceAssembler:
    """fs_assembler: the system's blocks and rhs re-assembled on the GPU on
    the system's patterns (StokesAssembler::local_assembly +
    copy_local_to_global). Buffers are torch CUDA tensors."""

    def __init__(self, system):
        import torch
        self._lib = L.lib()
        self.h = C.c_void_p()
        L.check(self._lib.fs_assembler_create(system.h, C.byref(self.h)), "fs_assembler_create")
        self.dim, self.n_cells, self.n_u, self.n_p = system.dim, system.n_cells, system.n_u, system.n_p
        self.n_q = system.n_cells_total * 3 ** system.dim
        dev = torch.device("cuda", torch.cuda.current_device())
        self.values = {name: torch.zeros(system.blocks[name].nnz, dtype=torch.float64, device=dev)
                       for name in BLOCK_NAMES}
        self.rhs = torch.zeros(system.n, dtype=torch.float64, device=dev)

    def sample_field(self):
        """(eta_q, rho_q) device tensors of the system's own field."""
        import torch
        eta = torch.empty(self.n_q, dtype=torch.float64, device=self.rhs.device)
        rho = torch.empty_like(eta)
        torch.cuda.synchronize()
        L.check(self._lib.fs_assembler_sample_field(self.h, C.c_void_p(eta.data_ptr()), C.c_void_p(rho.data_ptr())),
                "fs_assembler_sample_field")
        return eta, rho

    def assemble(self, eta_q, rho_q):
        """Assembles into self.values / self.rhs from device samples; returns
        the GPU milliseconds of the assembly."""
        import torch
        assert eta_q.is_cuda and eta_q.dtype == torch.float64 and eta_q.numel() == self.n_q
        assert rho_q.is_cuda and rho_q.dtype == torch.float64 and rho_q.numel() == self.n_q
        eta_q, rho_q = eta_q.contiguous(), rho_q.contiguous()
        torch.cuda.synchronize()
        v = self.values
        L.check(self._lib.fs_assembler_assemble(
            self.h, C.c_void_p(eta_q.data_ptr()), C.c_void_p(rho_q.data_ptr()),
            *[C.c_void_p(v[name].data_ptr()) for name in BLOCK_NAMES], C.c_void_p(self.rhs.data_ptr())),
            "fs_assembler_assemble")
        return self._lib.fs_assembler_last_ms(self.h)

    def close(self):
        if getattr(self, "h", None):
            self._lib.fs_assembler_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def save_blocks(path, blocks, rhs=None):
    """fs_blocks_save: the five blocks (scipy CSR, as stored) and optionally
    the rhs in the binary interchange format (include/fs_blocks_io.h)."""
    lib = L.lib()
    keep = []
    cs = []
    for k in BLOCK_NAMES:
        m = blocks[k]
        rp = np.ascontiguousarray(m.indptr, np.int64)
        ci = np.ascontiguousarray(m.indices, np.int32)
        v = np.ascontiguousarray(m.data, np.float64)
        keep += [rp, ci, v]
        cs.append(L.FsCsr(m.shape[0], m.shape[1], rp.ctypes.data_as(L.P(L.i64)), ci.ctypes.data_as(L.P(L.i32)),
                          v.ctypes.data_as(L.P(L.f64)), 0))
    r = None
    if rhs is not None:
        rr = np.ascontiguousarray(rhs, np.float64)
        keep.append(rr)
        r = rr.ctypes.data_as(L.P(L.f64))
    L.check(lib.fs_blocks_save(os.fsencode(path), *[C.byref(c) for c in cs], r), "fs_blocks_save")


def load_blocks(path):
    """fs_blocks_load: (dict of scipy CSR blocks A, Bt, B, K11, Mp; rhs or None)."""
    lib = L.lib()
    h = C.c_void_p()
    L.check(lib.fs_blocks_load(os.fsencode(path), C.byref(h)), "fs_blocks_load")
    try:
        out = {}
        for w, k in enumerate(BLOCK_NAMES):
            c = L.FsCsr()
            L.check(lib.fs_blocks_get(h, w, C.byref(c)), "fs_blocks_get")
            nnz = c.row_ptr[c.n_rows]
            rp = np.ctypeslib.as_array(c.row_ptr, (c.n_rows + 1,)).copy()
            ci = np.ctypeslib.as_array(c.col_idx, (max(nnz, 1),))[:nnz].copy()
            v = np.ctypeslib.as_array(c.val, (max(nnz, 1),))[:nnz].copy()
            out[k] = sp.csr_matrix((v, ci, rp), shape=(c.n_rows, c.n_cols))
        nu, npp = L.i64(), L.i64()
        r = lib.fs_blocks_rhs(h, C.byref(nu), C.byref(npp))
        rhs = np.ctypeslib.as_array(r, (nu.value + npp.value,)).copy() if r else None
        return out, rhs
    finally:
        lib.fs_blocks_free(h)
