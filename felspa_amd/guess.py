"""Learned warm start (include/fs_guess.h): the network object, its training
on solves that the GPU solver generates, and the lattice maps between dof
vectors and the network's node fields.

The reference has no learned part; its only hook is the initial guess of
StokesLinearSystem::solve_gmres (src/pde/stokes_ilu.cc:523-524), which
FS_GUESS_LEARNED fills. Training, inference and every gradient run on the
library's own BF16-MFMA kernels; torch is used here only to hold device
buffers.
"""
import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(HERE, "weights")
DEFAULT_2D = os.path.join(WEIGHTS_DIR, "guess_ellipses_2d.fsgm")
DEFAULT_3D = os.path.join(WEIGHTS_DIR, "guess_ellipsoids_3d.fsgm")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _sync_torch():
    """The model runs on its own HIP stream: torch's pending writes to the
    buffers it reads must be complete first."""
    import torch
    torch.cuda.current_stream().synchronize()


class GuessModel:
    """fs_guess_model: cosine-basis spectral neural operator on the GPU."""

    def __init__(self, dim=2, width=32, modes=12, layers=4, proj=64, seed=0, path=None):
        self._lib = L.lib()
        self.h = C.c_void_p()
        if path is not None:
            L.check(self._lib.fs_guess_model_load(os.fsencode(path), C.byref(self.h)), "fs_guess_model_load")
        else:
            cfg = L.FsGuessConfig(dim, width, modes, layers, proj, seed)
            L.check(self._lib.fs_guess_model_create(C.byref(cfg), C.byref(self.h)), "fs_guess_model_create")
        c = L.FsGuessConfig()
        L.check(self._lib.fs_guess_model_get_config(self.h, C.byref(c)), "fs_guess_model_get_config")
        self.cfg = dict(dim=c.dim, width=c.width, modes=c.modes, layers=c.layers, proj=c.proj)
        self.dim = c.dim
        self.n_params = int(self._lib.fs_guess_model_num_params(self.h))

    @classmethod
    def load(cls, path):
        return cls(path=path)

    def close(self):
        """Release the model's device memory now (also done on garbage collection)."""
        if getattr(self, "h", None):
            self._lib.fs_guess_model_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def save(self, path):
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        L.check(self._lib.fs_guess_model_save(self.h, os.fsencode(path)), "fs_guess_model_save")

    # parameters -------------------------------------------------------------
    def params(self):
        out = np.zeros(self.n_params, np.float32)
        L.check(self._lib.fs_guess_model_get_params(self.h, out.ctypes.data_as(L.P(L.f32))),
                "fs_guess_model_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        L.check(self._lib.fs_guess_model_set_params(self.h, p.ctypes.data_as(L.P(L.f32))),
                "fs_guess_model_set_params")

    def grads(self):
        out = np.zeros(self.n_params, np.float32)
        L.check(self._lib.fs_guess_model_get_grads(self.h, out.ctypes.data_as(L.P(L.f32))),
                "fs_guess_model_get_grads")
        return out

    def set_normalization(self, mean, std):
        m = np.ascontiguousarray(mean, np.float32)
        s = np.ascontiguousarray(std, np.float32)
        L.check(self._lib.fs_guess_model_set_normalization(self.h, m.ctypes.data_as(L.P(L.f32)),
                                                           s.ctypes.data_as(L.P(L.f32))),
                "fs_guess_model_set_normalization")

    def normalization(self):
        m = np.zeros(4, np.float32)
        s = np.zeros(4, np.float32)
        L.check(self._lib.fs_guess_model_get_normalization(self.h, m.ctypes.data_as(L.P(L.f32)),
                                                           s.ctypes.data_as(L.P(L.f32))),
                "fs_guess_model_get_normalization")
        return m[: self.dim + 1], s[: self.dim + 1]

    # device passes (torch CUDA tensors as plain device buffers) -------------
    def prepare_input(self, d_log_eta, n_cells):
        import torch
        B = d_log_eta.shape[0]
        P = (2 * n_cells + 1) ** self.dim
        X = torch.empty((B, self.dim + 1, P), dtype=torch.float32, device=d_log_eta.device)
        d_log_eta = d_log_eta.contiguous()
        _sync_torch()
        L.check(self._lib.fs_guess_prepare_input(self.h, n_cells, B, _ptr(d_log_eta), _ptr(X)),
                "fs_guess_prepare_input")
        return X

    def forward(self, X, n_cells):
        import torch
        Y = torch.empty_like(X[:, : self.dim + 1])
        _sync_torch()
        L.check(self._lib.fs_guess_model_forward(self.h, n_cells, X.shape[0], _ptr(X), _ptr(Y)),
                "fs_guess_model_forward")
        return Y

    def loss_grad(self, X, T, n_cells):
        loss = L.f64()
        _sync_torch()
        L.check(self._lib.fs_guess_model_loss_grad(self.h, n_cells, X.shape[0], _ptr(X), _ptr(T),
                                                   C.byref(loss)), "fs_guess_model_loss_grad")
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        L.check(self._lib.fs_guess_model_adam_step(self.h, lr, beta1, beta2, eps, weight_decay),
                "fs_guess_model_adam_step")

    def time_forward(self, n_cells, batch=1, reps=20):
        ms, fl = L.f64(), L.f64()
        L.check(self._lib.fs_guess_model_time_forward(self.h, n_cells, batch, reps, C.byref(ms), C.byref(fl)),
                "fs_guess_model_time_forward")
        return ms.value, fl.value


# ---------------------------------------------------------------------------
# dof vectors <-> node lattices
def q1_to_q2(field_q1, dim, n_cells):
    """Multilinear interpolation of a (N+1)^dim nodal field (x fastest) to the
    (2N+1)^dim lattice: the Q1 finite-element field evaluated at Q2 nodes."""
    nq, n1 = n_cells + 1, 2 * n_cells + 1
    f = np.asarray(field_q1, dtype=np.float64).reshape((nq,) * dim)
    for ax in range(dim):
        shp = list(f.shape)
        shp[ax] = n1
        g = np.zeros(shp)
        ev = [slice(None)] * dim
        ev[ax] = slice(0, n1, 2)
        g[tuple(ev)] = f
        od = [slice(None)] * dim
        od[ax] = slice(1, n1, 2)
        a = [slice(None)] * dim
        b = [slice(None)] * dim
        a[ax] = slice(0, nq - 1)
        b[ax] = slice(1, nq)
        g[tuple(od)] = 0.5 * (f[tuple(a)] + f[tuple(b)])
        f = g
    return f.reshape(-1)


def lattice_targets(S, x):
    """[dim+1, (2N+1)^dim] node fields of a solution x (caller's dof order):
    velocity components at their nodes, pressure interpolated from its nodes."""
    d, nc = S.dim, S.n_cells
    P = (2 * nc + 1) ** d
    T = np.zeros((d + 1, P), np.float64)
    T[S.vcomp.astype(np.int64), S.vnode] = x[: S.n_u]
    pq = np.zeros((nc + 1) ** d)
    pq[S.pnode] = x[S.n_u:]
    T[d] = q1_to_q2(pq, d, nc)
    return T.astype(np.float32)


def lattice_to_dofs(S, T):
    """Inverse of lattice_targets on the dofs: x from the node fields."""
    d, nc = S.dim, S.n_cells
    x = np.zeros(S.n)
    x[: S.n_u] = T[S.vcomp.astype(np.int64), S.vnode]
    n1, nq = 2 * nc + 1, nc + 1
    idx = np.indices((nq,) * d).reshape(d, -1)  # slowest first
    lat = np.zeros(nq ** d, np.int64)
    stride = 1
    for a in range(d):
        lat += 2 * idx[d - 1 - a] * stride
        stride *= n1
    x[S.n_u:] = T[d][lat[S.pnode]]
    return x


# ---------------------------------------------------------------------------
# data generation and training
def generate_solutions(cfg_name, samples, n_cells=None, restart=30, inner_a=60, rtol=1e-8, log=None,
                       budget_s=0.0, ilu_k=None, prefetch=3):
    """Solve the seeded problems of a configuration with the GPU solver
    (mixed-precision FGMRES to `rtol` from a zero guess; ILU level `ilu_k`,
    default the configuration's: the solution does not depend on it),
    stopping early once `budget_s` seconds have passed (0: no limit); the
    next `prefetch` systems are assembled on host threads while the GPU
    solves. Returns log10(eta) on the pressure lattice [S, (N+1)^d] and the
    node-lattice solutions [S, d+1, (2N+1)^d] (float32), plus per-solve
    statistics."""
    from concurrent.futures import ThreadPoolExecutor
    from .configs import make_problem, solver_settings
    from .stokes import StokesLinearSystem
    from .system import StokesSystem
    st_cfg = solver_settings(cfg_name)
    kA = st_cfg["ilu_k_A"] if ilu_k is None else ilu_k
    lin = S0 = None
    logeta, targets, stats = [], [], []
    t0 = time.time()
    pool = ThreadPoolExecutor(max_workers=max(1, prefetch))
    make = lambda s: StokesSystem(make_problem(cfg_name, n_cells=n_cells, sample=s))
    pending = [pool.submit(make, s) for s in samples[:prefetch]]
    for i, s in enumerate(samples):
        if budget_s and time.time() - t0 > budget_s:
            if log:
                log("data budget reached after %d of %d solves" % (i, len(samples)))
            break
        S = pending.pop(0).result()
        if i + prefetch < len(samples):
            pending.append(pool.submit(make, samples[i + prefetch]))
        if lin is None:
            S0 = S
            lin = StokesLinearSystem(S.blocks, A_lof=kA, S_lof=st_cfg["ilu_k_S"],
                                     precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO,
                                     block_hint=S.block_hint(st_cfg["hint_nodes"]),
                                     constraints=S.constrained, restart=restart, inner_a_iters=inner_a,
                                     rtol=rtol)
        else:
            for k in ("A", "Bt", "B", "K11", "Mp"):
                a, b = S.blocks[k], S0.blocks[k]
                if not (np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)):
                    raise ValueError("sample %d: sparsity differs from sample %d" % (s, samples[0]))
            lin.refactor(S.blocks)
        x = np.zeros(S.n)
        st = lin.solve(x, S.rhs)
        x = S.distribute(x)
        logeta.append(S.log_eta.astype(np.float32))
        targets.append(lattice_targets(S, x))
        stats.append(dict(sample=s, outer=st["outer_iters"], res=st["true_rel_residual"],
                          ms=st["t_total_ms"]))
        if S is not S0:
            del S
        if log and (i % 16 == 0 or i == len(samples) - 1):
            log("solved %d/%d (sample %d: %d outer, %.1f ms) %.1fs" % (i + 1, len(samples), s, st["outer_iters"],
                                                                       st["t_total_ms"], time.time() - t0))
    for f in pending:
        f.cancel()
    pool.shutdown(wait=True)
    lin.close()
    return np.stack(logeta), np.stack(targets), stats


def train(model, d_X, d_T, n_cells, epochs=100, batch=16, lr=2e-3, lr_final=1e-4, weight_decay=1e-5,
          d_Xv=None, d_Tv=None, log=None, seed=0, budget_s=0.0):
    """Adam on the normalised MSE; d_X [S, d+1, P] inputs, d_T the normalised
    targets (device tensors). Cosine learning-rate decay from lr to lr_final.
    Stops after the epoch during which `budget_s` seconds passed (0: no
    limit). Returns the per-epoch (train loss, held-out loss) history."""
    import torch
    S = d_X.shape[0]
    nb = max(1, S // batch)
    steps = epochs * nb
    g = torch.Generator().manual_seed(seed)
    hist = []
    step = 0
    t0 = time.time()
    t_log = t0
    for ep in range(epochs):
        perm = torch.randperm(S, generator=g).to(d_X.device)
        tot = 0.0
        for k in range(nb):
            idx = perm[k * batch:(k + 1) * batch]
            xb = d_X.index_select(0, idx).contiguous()
            tb = d_T.index_select(0, idx).contiguous()
            tot += model.loss_grad(xb, tb, n_cells)
            f = 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps - 1)))
            model.adam_step(lr_final + (lr - lr_final) * f, weight_decay=weight_decay)
            step += 1
        vl = None
        if d_Xv is not None:
            vl = evaluate_loss(model, d_Xv, d_Tv, n_cells, batch)
        hist.append((tot / nb, vl))
        over = budget_s and time.time() - t0 > budget_s
        if log and (ep % 10 == 0 or ep == epochs - 1 or over or time.time() - t_log > 60):
            t_log = time.time()
            log("epoch %d: train %.4e held-out %s (%.1fs)" % (ep, tot / nb, "%.4e" % vl if vl is not None else "-",
                                                              time.time() - t0))
        if over:
            break
    return hist


def evaluate_loss(model, d_X, d_T, n_cells, batch=16):
    """Normalised MSE of the network over a set (forward only)."""
    S = d_X.shape[0]
    tot = 0.0
    for k in range(0, S, batch):
        Y = model.forward(d_X[k:k + batch].contiguous(), n_cells)
        tot += float(((Y - d_T[k:k + batch]) ** 2).sum())
    return tot / float(d_T.numel())
