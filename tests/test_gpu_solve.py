"""GPU parity of the Stokes solve (StokesLinearSystem::solve) against the
CPU oracle on identical matrices.

Three levels:
  * reference mode (same algorithm, same row order, same stopping rule as the
    reference): iteration count and solution agree with the oracle;
  * FGMRES (FP64 and mixed FP32/FP64): the FP64 true residual meets rtol;
  * the same FGMRES method restated on the CPU by the oracle: iteration count
    and solution agree in the caller's row order;
  * a tight tolerance, where the velocity error falls below 1e-6 and the
    mean-free pressure error below 1e-4.
The exact discrete solution comes from a sparse direct solve of K.
At rtol 1e-8 the error is bounded only through the conditioning of K (the
reference's own stopping rule leaves larger errors); it is reported by
bench.py, and asserted here at a tight tolerance.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import rel, system

pytestmark = pytest.mark.gpu

CASES = [("C1", 16, 0), ("C2", 16, 1), ("C3", 4, 1)]
_truth = {}


def truth(S):
    key = id(S)
    if key not in _truth:
        _truth[key] = spla.spsolve(S.K().tocsc(), S.rhs)
    return _truth[key]


def errors(S, x):
    """relative L2 error of u and of p against the exact discrete solution;
    the pressure error is taken modulo a constant (the reference pins a single
    pressure dof, so the constant pressure mode is fixed only through the
    couplings of that dof and is amplified by the system's conditioning)."""
    t = truth(S)
    dp = x[S.n_u:] - t[S.n_u:]
    tp = t[S.n_u:]
    return rel(x[:S.n_u], t[:S.n_u]), np.linalg.norm(dp - dp.mean()) / np.linalg.norm(tp - tp.mean())


def test_block_schur_precond_reference_order(oracle):
    S = system("C1", 8)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, ordering=L.FS_ORDER_REFERENCE,
                                method=L.FS_METHOD_GMRES_REF)
    src = np.random.default_rng(4).standard_normal(S.n)
    dst, st = lin.precondition(src)
    want, rc, it = oracle.block_schur_vmult(S.blocks, src, kA=1, kS=0, inner_tol=1e-12)
    assert rc == 0
    assert rel(dst[:S.n_u], want[:S.n_u]) < 1e-12
    assert rel(dst[S.n_u:], want[S.n_u:]) < 1e-9


@pytest.mark.parametrize("cfg,n", [("C1", 8), ("C2", 8), ("C3", 3)])
def test_reference_mode_matches_oracle(oracle, cfg, n):
    """Left-preconditioned GMRES(23), ReductionControl(1e-8||b||, 1e-6), ILU(1)
    applied once per preconditioner call, inner CG to 1e-12, reference order:
    the reference's method, so iterations and solution agree."""
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, method=L.FS_METHOD_GMRES_REF,
                                ordering=L.FS_ORDER_REFERENCE, constraints=S.constrained)
    x = np.zeros(S.n)
    ctrl = fa.StokesSolverControl()
    st = lin.solve(x, S.rhs, ctrl)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=1, kS=0)
    assert sto["status"] == 0
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    assert ctrl.n_gmres_iter == [st["outer_iters"]]
    du, dp = rel(x[:S.n_u], xo[:S.n_u]), rel(x[S.n_u:], xo[S.n_u:])
    assert du < 1e-7 and dp < 1e-7, (du, dp)


@pytest.mark.parametrize("precision", [L.FS_PREC_FP64, L.FS_PREC_MIXED, L.FS_PREC_MIXED_FP64_VECTORS])
@pytest.mark.parametrize("cfg,n,kA", CASES)
def test_fgmres_meets_true_residual(cfg, n, kA, precision):
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, precision=precision,
                                block_hint=S.block_hint(2 * kA + 5), constraints=S.constrained)
    x = np.zeros(S.n)
    ctrl = fa.StokesSolverControl()
    st = lin.solve(x, S.rhs, ctrl)
    r = np.linalg.norm(S.rhs - S.K() @ x) / np.linalg.norm(S.rhs)
    assert r <= 1e-8 * 1.0001 and st["true_rel_residual"] <= 1e-8 * 1.0001
    assert abs(st["true_rel_residual"] - r) <= 1e-3 * r + 1e-15
    assert ctrl.n_gmres_iter == [st["outer_iters"]] and st["inner_iters"] > 0


@pytest.mark.parametrize("cfg,n,kA", [("C1", 8, 0), ("C2", 8, 1), ("C3", 3, 1)])
def test_fgmres_matches_oracle_port(oracle, cfg, n, kA):
    """FP64 FGMRES in the caller's row order against the oracle's CPU
    restatement of the same method (same ILU, same inner steps, true
    residual restarts; the oracle orthogonalises by classical Gram-Schmidt
    twice, the GPU by the delayed form, equal in exact arithmetic): the
    iteration count to one and the same solution."""
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, ordering=L.FS_ORDER_REFERENCE, restart=30,
                                inner_a_iters=30, constraints=S.constrained)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=kA, kS=0, method=2, restart=30, rtol=1e-8,
                                  inner_a_iters=30)
    assert sto["status"] == 0
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    du = rel(x[:S.n_u], xo[:S.n_u])
    dp = x[S.n_u:] - xo[S.n_u:]
    tp = xo[S.n_u:]
    dpm = np.linalg.norm(dp - dp.mean()) / np.linalg.norm(tp - tp.mean())
    assert du < 1e-6 and dpm < 1e-4, (du, dpm)


@pytest.mark.parametrize("cfg,n,kA,restart,inner,precision", [("C2", 16, 1, 250, 20, L.FS_PREC_MIXED),
                                                               ("C3", 3, 1, 10, 30, L.FS_PREC_FP64),
                                                               ("C1", 16, 0, 20, 5, L.FS_PREC_FP64)])
def test_fgmres_delayed_reorth_matches_cgs2(monkeypatch, cfg, n, kA, restart, inner, precision):
    """The flexible method's Arnoldi step with delayed reorthogonalisation
    (the default) against classical Gram-Schmidt twice per step
    (FS_FGMRES_CGS2=1), in one cycle and over restarts (a cycle's last
    column provisional): both meet the true residual tolerance, the step
    counts agree to one and the solutions to 1e-6. (The oracle's FP64
    restatement needs 41 and 78 steps, four or more cycles, on the two
    restarted cases; they run in FP64: with FP32 vectors FGMRES(10) nearly
    stagnates on C3 and rounding alone moves its restart history.)"""
    S = system(cfg, n)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FS_FGMRES_CGS2", mode)
        lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, precision=precision,
                                    block_hint=S.block_hint(2 * kA + 5), constraints=S.constrained,
                                    restart=restart, inner_a_iters=inner, max_outer=3000)
        x = np.zeros(S.n)
        st = lin.solve(x, S.rhs)
        r = np.linalg.norm(S.rhs - S.K() @ x) / np.linalg.norm(S.rhs)
        assert st["status"] == 0 and r <= 1e-8 * 1.0001, (mode, st, r)
        out[mode] = (x, st)
        lin.close()
    (x1, s1), (x0, s0) = out["1"], out["0"]
    assert abs(s1["outer_iters"] - s0["outer_iters"]) <= 1, (s1["outer_iters"], s0["outer_iters"])
    if restart < 250:
        assert s0["restarts"] >= 3, s0
    assert rel(x0[:S.n_u], x1[:S.n_u]) < 1e-6


def test_large_basis_and_krylov_switch_match_oracle_port(oracle):
    """A Krylov basis past 64 vectors (the Gram-Schmidt multi-dots run in
    chunks of 32): FP64 FGMRES(200) with one inner PCG step on C1 16^2 needs
    about 110 outer iterations without a restart; same count (within two) and
    solution as the oracle's restatement. Then fs_stokes_set_krylov on the
    same handle: a restart beyond the allocated basis is refused, and
    FGMRES(30) with 30 inner steps matches the oracle's run of those
    settings."""
    S = system("C1", 16)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=0, S_lof=0, ordering=L.FS_ORDER_REFERENCE, restart=200,
                                inner_a_iters=1, constraints=S.constrained)
    for restart, ia, min_iters in ((200, 1, 65), (30, 30, 1)):
        lin.set_krylov(restart=restart, inner_a_iters=ia)
        x = np.zeros(S.n)
        st = lin.solve(x, S.rhs)
        xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=0, kS=0, method=2, restart=restart, rtol=1e-8,
                                      inner_a_iters=ia)
        assert sto["status"] == 0 and st["status"] == 0
        assert st["outer_iters"] >= min_iters, st["outer_iters"]
        assert abs(st["outer_iters"] - sto["outer_iters"]) <= 2, (restart, st["outer_iters"], sto["outer_iters"])
        assert rel(x[:S.n_u], xo[:S.n_u]) < 1e-6
    with pytest.raises(L.FsException) as e:
        lin.set_krylov(restart=201)
    assert e.value.code == L.FS_ERR_BAD_INPUT
    lin.close()


@pytest.mark.parametrize("precision", [L.FS_PREC_FP64, L.FS_PREC_MIXED, L.FS_PREC_MIXED_FP64_VECTORS])
@pytest.mark.parametrize("cfg,n,kA", CASES)
def test_fgmres_tight_tolerance(cfg, n, kA, precision):
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, precision=precision, rtol=1e-12,
                                block_hint=S.block_hint(2 * kA + 5), constraints=S.constrained)
    x = np.zeros(S.n)
    lin.solve(x, S.rhs)
    eu, ep = errors(S, x)
    assert eu < 1e-6, eu
    # C1 (contrast-1e3 inclusion, one pinned pressure dof): the pressure error
    # is amplified ~1e6 over the residual also in the oracle's FP64 port of
    # this method (4.5e-6 at rtol 1e-12), so only the velocity is asserted
    if cfg != "C1":
        assert ep < 1e-4, ep


def test_inner_cg_one_step_by_launches(oracle):
    """The inner CG on M_p capped at one step runs as ordinary launches: the
    block preconditioner's pressure part is one ILU(0)-PCG step from zero,
    x = (t.d / d.M d) d with d = ILU_S^{-1} t and t = B ILU_A^{-1} src_u -
    src_p (the oracle's factors, reference order), equal to the cooperative
    kernel's result when that one stops after its first step (cap 2), in
    FP64 and with FP32 vectors; a step short of the tolerance is reported as
    FS_NOT_CONVERGED."""
    S = system("C1", 8)
    nu = S.n_u
    src = np.random.default_rng(9).standard_normal(S.n)
    du = oracle.SparseILU(S.blocks["A"], 1).vmult(src[:nu])
    t = S.blocks["B"] @ du - src[nu:]
    d = oracle.SparseILU(S.blocks["Mp"], 0).vmult(t)
    Md = S.blocks["Mp"] @ d
    alpha = (t @ d) / (d @ Md)
    want_p = alpha * d
    # the premise: one step reduces the residual below 0.95 |t|
    assert np.linalg.norm(alpha * Md - t) < 0.95 * np.linalg.norm(t)
    out = {}
    for prec, tol_p in ((L.FS_PREC_FP64, 1e-12), (L.FS_PREC_MIXED, 1e-5)):
        for cap in (1, 2):
            lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, ordering=L.FS_ORDER_REFERENCE,
                                        method=L.FS_METHOD_FGMRES, precision=prec, inner_a_iters=0,
                                        inner_cg_tol=0.95, inner_cg_max=cap, inner_cg_tol_mixed=0.95,
                                        inner_cg_max_mixed=cap)
            dst, st = lin.precondition(src, precision=prec)
            assert st["status"] == L.FS_OK and st["inner_iters"] == 1
            assert rel(dst[:nu], du) < tol_p
            out[prec, cap] = dst[nu:]
            lin.close()
        assert rel(out[prec, 1], want_p) < tol_p
        assert rel(out[prec, 1], out[prec, 2]) < tol_p
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, ordering=L.FS_ORDER_REFERENCE,
                                method=L.FS_METHOD_FGMRES, inner_a_iters=0, inner_cg_tol=1e-8, inner_cg_max=1)
    dst, st = lin.precondition(src)
    assert st["status"] == L.FS_NOT_CONVERGED and st["inner_iters"] == 1
    assert rel(dst[nu:], want_p) < 1e-12
    lin.close()


def test_refactor_and_warm_start():
    """New values on the same pattern; a converged guess needs no iteration."""
    S = system("C2", 8)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, block_hint=S.block_hint(7))
    x = np.zeros(S.n)
    lin.solve(x, S.rhs)
    blocks2 = {k: v.copy() for k, v in S.blocks.items()}
    blocks2["A"].data *= 2.0
    blocks2["Mp"].data *= 0.5
    lin.refactor(blocks2)
    K2 = __import__("scipy.sparse", fromlist=["bmat"]).bmat(
        [[blocks2["A"], blocks2["Bt"]], [blocks2["B"], blocks2["K11"]]]).tocsr()
    x2 = np.zeros(S.n)
    lin.solve(x2, S.rhs)
    assert np.linalg.norm(S.rhs - K2 @ x2) <= 1e-8 * np.linalg.norm(S.rhs) * 1.0001
    st = lin.solve(x2, S.rhs)
    assert st["outer_iters"] == 0


def test_vmult_matches_block_operator():
    S = system("C3", 3)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, block_hint=S.block_hint(7))
    v = np.random.default_rng(5).standard_normal(S.n)
    assert rel(lin.vmult(v), S.K() @ v) < 1e-14
    assert rel(lin.vmult(v, precision=L.FS_PREC_MIXED), S.K() @ v) < 1e-5


def test_error_paths():
    S = system("C1", 6)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, max_outer=2, block_hint=S.block_hint(7))
    x = np.zeros(S.n)
    with pytest.raises(fa.NoConvergence):
        lin.solve(x, S.rhs)
    bad = S.rhs.copy()
    bad[3] = np.nan
    with pytest.raises(L.FsException) as e:
        lin.solve(np.zeros(S.n), bad)
    assert e.value.code == L.FS_ERR_BAD_INPUT


_refined = {}


def _refined_truth(S, steps=2):
    """Sparse direct solve of K x = b with `steps` rounds of iterative
    refinement (FP64 residuals); one per system."""
    if id(S) not in _refined:
        K = S.K().tocsc()
        lu = spla.splu(K)
        x = lu.solve(S.rhs)
        for _ in range(steps):
            x = x + lu.solve(S.rhs - K @ x)
        _refined[id(S)] = x
    return _refined[id(S)]


@pytest.mark.parametrize("cfg,n,kA,hint,p_bar", [("C1", 64, 0, 5, 1e-4), ("C4", 128, 1, 5, 2e-6)])
@pytest.mark.parametrize("precision", [L.FS_PREC_FP64, L.FS_PREC_MIXED])
def test_full_size_parity_against_direct_solve(cfg, n, kA, hint, p_bar, precision):
    """The product method (block-colour ILU, FGMRES with inner ILU-PCG on A;
    FP64, and FP32 preconditioner inside FP64 FGMRES) at the configurations'
    full sizes, C1 64^2 and C4 128^2, against a refined sparse direct solve:
    velocity within the 1e-7 relative L2 of the accuracy target, at rtol
    1e-14 (at 1e-12 the oracle's FP64 CPU restatement of this method is still
    3.7e-6 from the direct solution in velocity on C4).
    The pressure is ill-conditioned: one refinement step moves the direct
    solution's pressure by 3e-7 (C4) and 5e-7 (C1) relative. The FP64 CPU
    oracle of this method reaches 7e-7 on C4's mean-free pressure at rtol
    1e-14, held here to 2e-6. On C1 (contrast-1e3 inclusion, periodic in x)
    the same oracle stays 2.5e-5 from the direct solution at rtol 1e-12 and at
    1e-14 alike, with both solutions' residuals at round-off, and whether or
    not the pinned pressure dof is counted: FP64 fixes the pressure there
    only to that level (this GPU: 1.9e-5 FP64, 3.6e-5 mixed), held to 1e-4."""
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, precision=precision, rtol=1e-14, restart=60,
                                inner_a_iters=60, max_outer=3000, block_hint=S.block_hint(hint),
                                constraints=S.constrained, inner_cg_tol=1e-2, inner_cg_max=50)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    assert st["true_rel_residual"] <= 1e-14
    t = _refined_truth(S)
    eu = rel(x[:S.n_u], t[:S.n_u])
    dp = x[S.n_u:] - t[S.n_u:]
    tp = t[S.n_u:]
    ep = np.linalg.norm(dp - dp.mean()) / np.linalg.norm(tp - tp.mean())
    assert eu < 1e-7, eu
    assert ep < p_bar, ep
    lin.close()


def test_user_guess_matches_oracle(oracle):
    """The initial-guess slot (stokes_ilu.cc:523-524: GMRES starts from the
    caller's soln with its constrained entries zeroed): reference-mode GMRES
    from a random nonzero guess, in the reference's row order, against the
    oracle's solve from the same guess: iteration count within one, solution
    within 1e-7."""
    S = system("C2", 8)
    rng = np.random.default_rng(11)
    x0 = rng.standard_normal(S.n)
    x0[S.constrained] = 0.0
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, ordering=L.FS_ORDER_REFERENCE,
                                method=L.FS_METHOD_GMRES_REF, constraints=S.constrained)
    x = x0.copy()
    st = lin.solve(x, S.rhs)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, x0=x0, kA=1, kS=0)
    assert sto["status"] == 0 and st["status"] == 0
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    assert rel(x, xo) < 1e-7
    # a different guess changes the iterates (the slot is used)
    x1 = np.zeros(S.n)
    st1 = lin.solve(x1, S.rhs)
    xo1, sto1 = oracle.stokes_solve(S.blocks, S.rhs, kA=1, kS=0)
    assert abs(st1["outer_iters"] - sto1["outer_iters"]) <= 1
    lin.close()


def test_system_from_block_file_solves_identically(tmp_path):
    """A system written and read back through the block-CSR interchange
    (include/fs_blocks_io.h) gives the same solve bit for bit."""
    S = system("C2", 8)
    path = str(tmp_path / "c2.fsbk")
    fa.save_blocks(path, S.blocks, S.rhs)
    blocks, rhs = fa.load_blocks(path)
    xs = []
    for bl, b in ((S.blocks, S.rhs), (blocks, rhs)):
        lin = fa.StokesLinearSystem(bl, A_lof=1, precision=L.FS_PREC_MIXED, block_hint=S.block_hint(5),
                                    constraints=S.constrained, restart=60, inner_a_iters=30)
        x = np.zeros(S.n)
        st = lin.solve(x, b)
        assert st["true_rel_residual"] <= 1e-8
        xs.append(x)
        lin.close()
    assert np.array_equal(xs[0], xs[1])
