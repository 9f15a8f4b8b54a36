"""GPU parity of the diagonal scaling (StokesSolverControl::apply_diagonal_scaling,
StokesScalingOperator, include/felspa/pde/src/stokes_common.implement.h:13-140,
pre/post_solve_scaling src/pde/stokes_ilu.cc:418-438) against the CPU
oracle, which scales the matrices explicitly; the GPU applies the same
scaling to the Krylov operators (K~ = D^-1 K D^-1, P~^-1 = D P^-1 D), which
is the same iteration in exact arithmetic."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import rel, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n,kA", [("C1", 8, 1), ("C2", 8, 1), ("C3", 3, 1)])
def test_scaled_reference_gmres_matches_oracle(oracle, cfg, n, kA):
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, method=L.FS_METHOD_GMRES_REF, ordering=L.FS_ORDER_REFERENCE,
                                constraints=S.constrained, apply_diagonal_scaling=True)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=kA, kS=0, method=0, scaling=True)
    assert sto["status"] == 0
    # the stopping rule sits on a slowly falling residual, so rounding moves
    # the count by a few percent
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= max(1, round(0.05 * sto["outer_iters"])), (
        st["outer_iters"], sto["outer_iters"])
    # both stop at a 1e-6 reduction of the scaled preconditioned residual, so
    # they agree to about that level (the GPU scales the operators, the oracle
    # the matrices: the same iteration with different rounding)
    assert rel(x[:S.n_u], xo[:S.n_u]) < 1e-4


def test_scaled_fgmres_converges_to_the_unscaled_solution():
    S = system("C2", 16, 1)
    common = dict(A_lof=1, precision=L.FS_PREC_MIXED, block_hint=S.block_hint(5), constraints=S.constrained,
                  restart=30, inner_a_iters=40, rtol=1e-10)
    x0 = np.zeros(S.n)
    fa.StokesLinearSystem(S.blocks, **common).solve(x0, S.rhs)
    x1 = np.zeros(S.n)
    st = fa.StokesLinearSystem(S.blocks, apply_diagonal_scaling=True, **common).solve(x1, S.rhs)
    assert st["status"] == L.FS_OK
    # the scaled solve stops on ||D^-1 (b - K x)|| <= rtol ||D^-1 b||
    assert st["true_rel_residual"] < 1e-8
    truth = spla.spsolve(S.K().tocsc(), S.rhs)
    e0, e1 = rel(x0[:S.n_u], truth[:S.n_u]), rel(x1[:S.n_u], truth[:S.n_u])
    assert e1 < max(10 * e0, 1e-6), (e0, e1)


@pytest.mark.parametrize("cfg,n,kA", [("C1", 8, 0), ("C2", 8, 1), ("C3", 3, 0)])
def test_scaled_scr_matches_oracle(oracle, cfg, n, kA):
    # solve_cg with pre_solve_scaling (src/pde/stokes_ilu.cc:584-586,673-678):
    # the oracle scales the matrices; the GPU runs the Schur CG and its inner
    # A^-1 CG in the unscaled variables and measures ||D^-1 r||
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, method=L.FS_METHOD_SCR, ordering=L.FS_ORDER_REFERENCE,
                                constraints=S.constrained, apply_diagonal_scaling=True)
    x = np.zeros(S.n)
    ctrl = fa.StokesSolverControl()
    rhs0 = S.rhs.copy()
    st = lin.solve(x, S.rhs, ctrl)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=kA, kS=0, method=1, scaling=True)
    assert sto["status"] == 0 and st["status"] == L.FS_OK
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    # the final scaled residual norms agree in magnitude (both just under the
    # same tolerance on the same scaled right-hand side)
    assert 0.1 < st["last_value"] / sto["last_value"] < 10, (st["last_value"], sto["last_value"])
    assert rel(x[:S.n_u], xo[:S.n_u]) < 1e-6
    dp, dpo = x[S.n_u:] - x[S.n_u:].mean(), xo[S.n_u:] - xo[S.n_u:].mean()
    assert rel(dp, dpo) < 1e-5
    # the caller's rhs is left as it was
    assert np.array_equal(S.rhs, rhs0)
