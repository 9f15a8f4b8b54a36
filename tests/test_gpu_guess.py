"""GPU tests of the learned warm start (include/fs_guess.h) through the C-ABI:
the BF16 MFMA GEMM against an FP64 product of the same BF16-rounded
operands, the network's forward pass and gradient against the FP64 oracle
(oracle/guess_oracle.py, itself pinned by finite differences), the input
interpolation, weight files, training, and the guess inside the Stokes solve
(convergence to the same tolerance; fallback to zero for a useless guess)."""
import ctypes as C
import os

import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from felspa_amd import guess as G
from conftest import rel, system

import guess_oracle as go

pytestmark = pytest.mark.gpu


def bf16_round(a):
    """FP32 -> BF16 -> FP32, round to nearest even (the kernels' rounding)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def gemm(M, N, K, batch, A, sA, B, sB, Cm, sC, alpha=1.0, beta=0.0, a_bf16=False, b_bf16=False):
    import torch
    dA, dB, dC = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (A, B, Cm))
    rc = L.lib().fs_gemm_bf16(M, N, K, batch, C.c_void_p(dA.data_ptr()), int(a_bf16), *sA,
                              C.c_void_p(dB.data_ptr()), int(b_bf16), *sB, C.c_void_p(dC.data_ptr()), *sC,
                              alpha, beta)
    L.check(rc, "fs_gemm_bf16")
    return dC.cpu().numpy()


@pytest.mark.parametrize("M,N,K,batch", [(64, 64, 32, 1), (50, 70, 45, 3), (1, 129, 300, 2), (130, 7, 9, 5)])
def test_gemm_row_major_matches_fp64_of_rounded_operands(M, N, K, batch):
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.standard_normal((batch, M, K)).astype(np.float32)
    B = rng.standard_normal((batch, K, N)).astype(np.float32)
    C0 = rng.standard_normal((batch, M, N)).astype(np.float32)
    out = gemm(M, N, K, batch, A, (M * K, K, 1), B, (K * N, N, 1), C0, (M * N, N, 1), alpha=0.5, beta=2.0)
    want = 0.5 * np.einsum("bmk,bkn->bmn", bf16_round(A).astype(np.float64), bf16_round(B).astype(np.float64)) \
        + 2.0 * C0
    e = float(np.max(np.abs(out - want)) / np.max(np.abs(want)))
    assert e < 1e-5, e


def test_gemm_transposed_strides_and_bf16_operands():
    rng = np.random.default_rng(7)
    M, N, K, batch = 33, 65, 70, 4
    At = rng.standard_normal((batch, K, M)).astype(np.float32)  # A(i,k) = At[b,k,i]
    Bt = rng.standard_normal((batch, N, K)).astype(np.float32)  # B(k,j) = Bt[b,j,k]
    Bbits = (bf16_round(Bt).view(np.uint32) >> 16).astype(np.uint16)
    C0 = np.zeros((batch, N, M), np.float32)  # C stored transposed too
    out = gemm(M, N, K, batch, At, (K * M, 1, M), Bbits, (N * K, 1, K), C0, (N * M, 1, M), b_bf16=True)
    want = np.einsum("bki,bjk->bji", bf16_round(At).astype(np.float64), bf16_round(Bt).astype(np.float64))
    e = float(np.max(np.abs(out - want)) / np.max(np.abs(want)))
    assert e < 1e-5, e


CFG = dict(dim=2, width=8, modes=4, layers=2, proj=16)


def _model_and_data(cfg, nc, B, seed):
    import torch
    m = G.GuessModel(seed=seed, **cfg)
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(m.n_params) * 0.3).astype(np.float32)
    m.set_params(p)
    nq = (nc + 1) ** cfg["dim"]
    le = rng.uniform(0, 3, (B, nq)).astype(np.float32)
    X = m.prepare_input(torch.from_numpy(le).cuda(), nc)
    P = (2 * nc + 1) ** cfg["dim"]
    T = rng.standard_normal((B, cfg["dim"] + 1, P)).astype(np.float32)
    return m, p, le, X, T


@pytest.mark.parametrize("cfg,nc", [(CFG, 6), (dict(dim=3, width=8, modes=3, layers=2, proj=8), 3)])
def test_prepare_input_and_forward_match_oracle(cfg, nc):
    m, p, le, X, _ = _model_and_data(cfg, nc, 2, 3)
    Xo = go.prep_input(cfg["dim"], nc, le)
    assert np.max(np.abs(X.cpu().numpy() - Xo)) < 1e-6
    Y = m.forward(X, nc).cpu().numpy()
    net = go.Net(cfg, bf16_round(p).astype(np.float64))
    Yo = net.forward(Xo)
    assert rel(Y, Yo) < 3e-2, rel(Y, Yo)


@pytest.mark.parametrize("cfg,nc", [(CFG, 6), (dict(dim=3, width=8, modes=3, layers=2, proj=8), 3)])
def test_gradient_matches_oracle(cfg, nc):
    import torch
    m, p, le, X, T = _model_and_data(cfg, nc, 2, 5)
    loss = m.loss_grad(X, torch.from_numpy(T).cuda(), nc)
    g = m.grads()
    Xo = go.prep_input(cfg["dim"], nc, le)
    lo, go_ = go.Net(cfg, bf16_round(p).astype(np.float64)).loss_grad(Xo, T.astype(np.float64))
    assert abs(loss - lo) / lo < 2e-2
    cos = float(g @ go_) / (np.linalg.norm(g) * np.linalg.norm(go_))
    assert cos > 0.999, cos
    assert rel(g, go_) < 5e-2, rel(g, go_)
    # every parameter group individually
    lay, _ = go.layout(cfg)
    for name, (o, s) in lay.items():
        n = int(np.prod(s))
        assert rel(g[o:o + n], go_[o:o + n]) < 0.1, name


def test_weight_file_roundtrip_and_hash(tmp_path):
    m, p, _, _, _ = _model_and_data(CFG, 4, 1, 9)
    m.set_normalization([0.1, -0.2, 3.0], [1.0, 2.0, 0.5])
    path = str(tmp_path / "w.fsgm")
    m.save(path)
    m2 = G.GuessModel.load(path)
    assert m2.cfg == m.cfg
    assert np.array_equal(m2.params(), p)
    mean, std = m2.normalization()
    assert np.allclose(mean, [0.1, -0.2, 3.0]) and np.allclose(std, [1.0, 2.0, 0.5])
    raw = bytearray(open(path, "rb").read())
    raw[-3] ^= 0xFF
    open(path, "wb").write(bytes(raw))
    with pytest.raises(L.FsException):
        G.GuessModel.load(path)


def test_training_reduces_loss():
    import torch
    nc = 8
    cfg = dict(dim=2, width=16, modes=6, layers=2, proj=16)
    m = G.GuessModel(seed=1, **cfg)
    rng = np.random.default_rng(2)
    B = 8
    le = rng.uniform(0, 3, (B, (nc + 1) ** 2)).astype(np.float32)
    X = m.prepare_input(torch.from_numpy(le).cuda(), nc)
    # a smooth target that depends on the input: T = [sin(pi x) * eta, y, eta^2] (normalised-ish)
    Xc = X.cpu().numpy()
    T = np.stack([np.sin(np.pi * Xc[:, 1]) * Xc[:, 0], Xc[:, 2] - 0.5, Xc[:, 0] ** 2], axis=1).astype(np.float32)
    dT = torch.from_numpy(T).cuda()
    l0 = m.loss_grad(X, dT, nc)
    for _ in range(150):
        m.loss_grad(X, dT, nc)
        m.adam_step(3e-3)
    l1 = m.loss_grad(X, dT, nc)
    assert l1 < 0.1 * l0, (l0, l1)


def test_lattice_maps_roundtrip():
    S = system("C4", 16, 0)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(S.n)
    T = G.lattice_targets(S, x)
    assert np.allclose(G.lattice_to_dofs(S, T), x.astype(np.float32), atol=1e-6)


def _c4_solver(S, **kw):
    return fa.StokesLinearSystem(S.blocks, A_lof=1, precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO,
                                 block_hint=S.block_hint(5), constraints=S.constrained, restart=30,
                                 inner_a_iters=40, **kw)


def test_learned_guess_is_exact_for_the_solution_and_falls_back():
    """The guess path on a system whose network output is replaced by the
    solution itself (a model whose head outputs zeros, with mean = 0 and the
    solution injected through a second handle) is not available; instead:
    (1) an untrained model gives a guess, the solve still reaches rtol;
    (2) a model with an enormous output scale gives a guess worse than zero,
        and the solve falls back to the zero guess and converges."""
    S = system("C4", 16, 1)
    lin = _c4_solver(S)
    m = G.GuessModel(seed=3, width=8, modes=4, layers=2, proj=8)
    lin.set_guess_field(m, S)
    lin.set_guess_mode(L.FS_GUESS_LEARNED)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    assert st["true_rel_residual"] <= 1e-8
    assert st["guess_used"] in (L.FS_GUESS_LEARNED, L.FS_GUESS_ZERO)
    assert st["t_guess_ms"] > 0
    m.set_normalization([0, 0, 0], [1e12, 1e12, 1e12])
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    assert st["guess_used"] == L.FS_GUESS_ZERO
    assert st["guess_rel_residual"] > 1.0
    assert st["true_rel_residual"] <= 1e-8
    lin.set_guess_mode(L.FS_GUESS_ZERO)
    x0 = np.zeros(S.n)
    st0 = lin.solve(x0, S.rhs)
    assert st0["outer_iters"] == st["outer_iters"]


def test_learned_guess_requires_a_field():
    S = system("C4", 8, 0)
    lin = _c4_solver(S)
    lin.set_guess_mode(L.FS_GUESS_LEARNED)
    with pytest.raises(L.FsException):
        lin.solve(np.zeros(S.n), S.rhs)


@pytest.mark.skipif(not os.path.exists(G.DEFAULT_2D), reason="trained weights not present")
def test_trained_network_gains_on_held_out_fields():
    """The committed network (trained on 512 seeded C4 fields) against the
    zero guess on held-out fields (samples 512+): same tolerance, fewer outer
    iterations on average; a guess worse than zero is replaced by zero."""
    m = G.GuessModel.load(G.DEFAULT_2D)
    outer_l, outer_z, gres = [], [], []
    for s in (512, 513, 514, 515):
        S = system("C4", 128, s)
        lin = _c4_solver(S)
        lin.set_guess_field(m, S)
        x = np.zeros(S.n)
        stz = lin.solve(x, S.rhs)
        lin.set_guess_mode(L.FS_GUESS_LEARNED)
        x = np.zeros(S.n)
        stl = lin.solve(x, S.rhs)
        assert stl["true_rel_residual"] <= 1e-8 and stz["true_rel_residual"] <= 1e-8
        # the guess is used exactly when its residual is below the zero guess's
        assert (stl["guess_used"] == L.FS_GUESS_LEARNED) == (stl["guess_rel_residual"] < 1.0)
        if stl["guess_used"] == L.FS_GUESS_ZERO:
            assert stl["outer_iters"] == stz["outer_iters"]
        outer_l.append(stl["outer_iters"])
        outer_z.append(stz["outer_iters"])
        gres.append(stl["guess_rel_residual"])
    print("outer learned", outer_l, "zero", outer_z, "guess residual", gres)
    assert np.mean(outer_l) < np.mean(outer_z)
    assert np.median(gres) < 0.5


@pytest.mark.skipif(not os.path.exists(G.DEFAULT_3D), reason="trained 3D weights not present")
def test_trained_3d_network_gains_on_held_out_fields():
    """The committed 3D network (trained on 253 seeded 32^3 ellipsoid solves,
    samples 0..252) against the zero guess on fields it never saw (samples
    512+) at the training resolution: the same true-residual tolerance, and a
    guess whose residual is under half the zero guess's."""
    m = G.GuessModel.load(G.DEFAULT_3D)
    assert m.dim == 3
    outer_l, outer_z, gres = [], [], []
    lin = None
    for s in (512, 513, 514):
        S = system("C5", 32, s)
        if lin is None:
            lin = fa.StokesLinearSystem(S.blocks, A_lof=1, precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO,
                                        block_hint=S.block_hint(3), constraints=S.constrained, restart=60,
                                        inner_a_iters=15, max_outer=3000)
        else:
            lin.refactor(S.blocks)
        lin.set_guess_field(m, S)
        lin.set_guess_mode(L.FS_GUESS_ZERO)
        stz = lin.solve(np.zeros(S.n), S.rhs)
        lin.set_guess_mode(L.FS_GUESS_LEARNED)
        stl = lin.solve(np.zeros(S.n), S.rhs)
        assert stl["true_rel_residual"] <= 1e-8 and stz["true_rel_residual"] <= 1e-8
        assert (stl["guess_used"] == L.FS_GUESS_LEARNED) == (stl["guess_rel_residual"] < 1.0)
        outer_l.append(stl["outer_iters"])
        outer_z.append(stz["outer_iters"])
        gres.append(stl["guess_rel_residual"])
    lin.close()
    print("3D outer learned", outer_l, "zero", outer_z, "guess residual", gres)
    # the guess removes most of the initial residual; the outer iterations it
    # saves are few (the smooth error it removes is what this preconditioner
    # removes fastest): not more on average than from zero
    assert np.mean(outer_l) <= np.mean(outer_z)
    assert max(gres) < 0.5
