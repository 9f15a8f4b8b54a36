"""CPU oracle (numpy, FP64) of the learned warm-start network.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it, as the checker; the product path never does.

The network has no counterpart in the reference (FELSPA starts GMRES from the
previous solution or zero, src/pde/stokes_ilu.cc:523-524); it is the
cosine-basis spectral neural operator described in include/fs_guess.h, with
the same parameter layout, restated here in FP64 with exact (unrounded)
cosine matrices:
    H0 = W_lift X + b_lift
    Hl = act(S_l(H_{l-1}) + W_l H_{l-1} + b_l)   (act = GELU, identity on the last layer)
    Y  = W_p2 GELU(W_p1 H_L + b_p1) + b_p2
S_l = synthesis o (per-mode channel mixing R_l) o analysis along every axis.
Forward and the hand-written backward are pinned by central finite
differences (tests/test_cpu_guess_oracle.py); the GPU kernels are checked
against both.
"""
import numpy as np
from scipy.special import erf


def layout(cfg):
    """Parameter offsets/shapes in the order of include/fs_guess.h."""
    d, w, M, L, pw = cfg["dim"], cfg["width"], cfg["modes"], cfg["layers"], cfg["proj"]
    cin = cout = d + 1
    Mm = M ** d
    ent = []
    ent.append(("W_lift", (w, cin)))
    ent.append(("b_lift", (w,)))
    for l in range(L):
        ent.append(("R%d" % l, (Mm, w, w)))
        ent.append(("W%d" % l, (w, w)))
        ent.append(("b%d" % l, (w,)))
    ent += [("W_p1", (pw, w)), ("b_p1", (pw,)), ("W_p2", (cout, pw)), ("b_p2", (cout,))]
    out, o = {}, 0
    for name, shp in ent:
        n = int(np.prod(shp))
        out[name] = (o, shp)
        o += n
    return out, o


def unpack(cfg, flat):
    lay, n = layout(cfg)
    assert flat.size == n
    return {k: flat[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def pack(cfg, params):
    lay, n = layout(cfg)
    flat = np.zeros(n)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = np.asarray(params[k], dtype=np.float64).ravel()
    return flat


def cos_mats(n1, M):
    """Analysis Af[m, i] = w_i cos(pi m x_i) (trapezoid weights on x_i = i/(n1-1)),
    synthesis Sb[i, m] = (1 if m == 0 else 2) cos(pi m x_i)."""
    x = np.arange(n1) / (n1 - 1)
    h = 1.0 / (n1 - 1)
    wq = np.full(n1, h)
    wq[0] = wq[-1] = 0.5 * h
    C = np.cos(np.pi * np.arange(M)[:, None] * x[None, :])
    Af = C * wq[None, :]
    Sb = (C * np.where(np.arange(M) == 0, 1.0, 2.0)[:, None]).T
    return Af, Sb


def prep_input(dim, nc, log_eta_q1):
    """X [B, dim+1, (2nc+1)^dim]: multilinear interpolation of log10(eta) from
    the (nc+1)^dim lattice to the (2nc+1)^dim lattice, divided by 3, then the
    node coordinates."""
    q = np.asarray(log_eta_q1, dtype=np.float64)
    B = q.shape[0]
    nq, n1 = nc + 1, 2 * nc + 1
    f = q.reshape((B,) + (nq,) * dim)  # axes (b, [z,] y, x)
    for ax in range(1, dim + 1):
        shp = list(f.shape)
        shp[ax] = n1
        g = np.zeros(shp)
        sl_even = [slice(None)] * (dim + 1)
        sl_even[ax] = slice(0, n1, 2)
        g[tuple(sl_even)] = f
        sl_odd = [slice(None)] * (dim + 1)
        sl_odd[ax] = slice(1, n1, 2)
        a = [slice(None)] * (dim + 1)
        b = [slice(None)] * (dim + 1)
        a[ax] = slice(0, nq - 1)
        b[ax] = slice(1, nq)
        g[tuple(sl_odd)] = 0.5 * (f[tuple(a)] + f[tuple(b)])
        f = g
    P = n1 ** dim
    X = np.zeros((B, dim + 1, P))
    X[:, 0] = f.reshape(B, P) / 3.0
    idx = np.indices((n1,) * dim).reshape(dim, P)  # idx[0] is the slowest axis
    for a in range(dim):
        X[:, 1 + a] = idx[dim - 1 - a] / (n1 - 1)  # a = 0 is x (fastest)
    return X


def _gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def _gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _apply_axes(T, mat, dim, order):
    """T [B, C, s_{dim-1}, ..., s_0] (x last): contract axis a (0 = x) with
    mat [n_out, n_in] for a in order."""
    for a in order:
        ax = T.ndim - 1 - a
        T = np.moveaxis(np.tensordot(T, mat, axes=([ax], [1])), -1, ax)
    return T


class Net:
    def __init__(self, cfg, flat):
        self.cfg = dict(cfg)
        self.p = unpack(cfg, np.asarray(flat, dtype=np.float64))

    def _spec_mats(self, n1):
        Af, Sb = cos_mats(n1, self.cfg["modes"])
        return Af, Sb

    def forward(self, X, keep=False):
        c = self.cfg
        d, w, L = c["dim"], c["width"], c["layers"]
        B, _, P = X.shape
        n1 = int(round(P ** (1.0 / d)))
        assert n1 ** d == P
        M = c["modes"]
        Af, Sb = self._spec_mats(n1)
        p = self.p
        cache = {"X": X, "H": [], "Z": [None], "T2": [None]}
        H = np.einsum("oc,bcp->bop", p["W_lift"], X) + p["b_lift"][None, :, None]
        cache["H"].append(H)
        for l in range(1, L + 1):
            Hp = H
            Z = np.einsum("oc,bcp->bop", p["W%d" % (l - 1)], Hp) + p["b%d" % (l - 1)][None, :, None]
            T = Hp.reshape((B, w) + (n1,) * d)
            T2 = _apply_axes(T, Af, d, range(d)).reshape(B, w, M ** d)
            O = np.einsum("bcm,mco->bom", T2, p["R%d" % (l - 1)])
            S = _apply_axes(O.reshape((B, w) + (M,) * d), Sb, d, reversed(range(d))).reshape(B, w, P)
            Z = Z + S
            H = _gelu(Z) if l < L else Z
            cache["Z"].append(Z)
            cache["T2"].append(T2)
            cache["H"].append(H)
        P1 = np.einsum("oc,bcp->bop", p["W_p1"], H) + p["b_p1"][None, :, None]
        G = _gelu(P1)
        Y = np.einsum("oc,bcp->bop", p["W_p2"], G) + p["b_p2"][None, :, None]
        cache.update(P1=P1, G=G)
        return (Y, cache) if keep else Y

    def loss_grad(self, X, T):
        """loss = mean((Y - T)^2) and its gradient (flat, parameter order)."""
        c = self.cfg
        d, w, L = c["dim"], c["width"], c["layers"]
        Y, ca = self.forward(X, keep=True)
        B, _, P = X.shape
        n1 = int(round(P ** (1.0 / d)))
        M = c["modes"]
        Af, Sb = self._spec_mats(n1)
        p = self.p
        g = {}
        n = Y.size
        loss = float(np.sum((Y - T) ** 2) / n)
        dY = 2.0 * (Y - T) / n
        g["b_p2"] = dY.sum(axis=(0, 2))
        g["W_p2"] = np.einsum("bop,bcp->oc", dY, ca["G"])
        dP1 = np.einsum("oc,bop->bcp", p["W_p2"], dY) * _gelu_grad(ca["P1"])
        g["b_p1"] = dP1.sum(axis=(0, 2))
        g["W_p1"] = np.einsum("bop,bcp->oc", dP1, ca["H"][L])
        dH = np.einsum("oc,bop->bcp", p["W_p1"], dP1)
        for l in range(L, 0, -1):
            dZ = dH * _gelu_grad(ca["Z"][l]) if l < L else dH
            Hp = ca["H"][l - 1]
            g["b%d" % (l - 1)] = dZ.sum(axis=(0, 2))
            g["W%d" % (l - 1)] = np.einsum("bop,bcp->oc", dZ, Hp)
            dHp = np.einsum("oc,bop->bcp", p["W%d" % (l - 1)], dZ)
            dO = _apply_axes(dZ.reshape((B, w) + (n1,) * d), Sb.T, d, range(d)).reshape(B, w, M ** d)
            g["R%d" % (l - 1)] = np.einsum("bcm,bom->mco", ca["T2"][l], dO)
            dT2 = np.einsum("bom,mco->bcm", dO, p["R%d" % (l - 1)])
            dHp = dHp + _apply_axes(dT2.reshape((B, w) + (M,) * d), Af.T, d,
                                    reversed(range(d))).reshape(B, w, P)
            dH = dHp
        g["b_lift"] = dH.sum(axis=(0, 2))
        g["W_lift"] = np.einsum("bop,bcp->oc", dH, ca["X"])
        return loss, pack(c, g)
