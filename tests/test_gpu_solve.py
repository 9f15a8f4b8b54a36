"""GPU parity of the Stokes solve (StokesLinearSystem::solve) against the
CPU oracle on identical matrices."""
import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import rel, system

pytestmark = pytest.mark.gpu


def _tight(oracle, S, kA, kS):
    """The exact discrete solution (sparse direct solve), far below the test
    tolerance."""
    import scipy.sparse.linalg as spla
    return spla.spsolve(S.K().tocsc(), S.rhs)


def _errors(S, x, truth):
    nu = S.n_u
    return rel(x[:nu], truth[:nu]), rel(x[nu:], truth[nu:])


def test_block_schur_precond_reference_order(oracle):
    S = system("C1", 8)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, ordering=L.FS_ORDER_REFERENCE)
    src = np.random.default_rng(4).standard_normal(S.n)
    dst, st = lin.precondition(src)
    want, rc, it = oracle.block_schur_vmult(S.blocks, src, kA=1, kS=0, inner_tol=1e-12)
    assert rc == 0
    assert rel(dst[:S.n_u], want[:S.n_u]) < 1e-12
    assert rel(dst[S.n_u:], want[S.n_u:]) < 1e-9


@pytest.mark.parametrize("cfg,n,kA", [("C1", 16, 0), ("C2", 16, 1), ("C3", 4, 1)])
def test_solve_fgmres_fp64(oracle, cfg, n, kA):
    S = system(cfg, n)
    truth = _tight(oracle, S, kA, 0)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, block_hint=S.block_hint(2 * kA + 3),
                                constraints=S.constrained)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs, fa.StokesSolverControl())
    K = S.K()
    assert np.linalg.norm(S.rhs - K @ x) <= 1e-8 * np.linalg.norm(S.rhs) * 1.0001
    assert st["true_rel_residual"] <= 1e-8
    eu, ep = _errors(S, x, truth)
    assert eu < 1e-6 and ep < 1e-5, (eu, ep)


@pytest.mark.parametrize("cfg,n,kA", [("C1", 16, 0), ("C2", 16, 1), ("C3", 4, 1)])
def test_solve_mixed_precision(oracle, cfg, n, kA):
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=0, precision=L.FS_PREC_MIXED,
                                block_hint=S.block_hint(2 * kA + 3), constraints=S.constrained)
    x = np.zeros(S.n)
    st = lin.solve(x, S.rhs)
    K = S.K()
    assert np.linalg.norm(S.rhs - K @ x) <= 1e-8 * np.linalg.norm(S.rhs) * 1.0001
    truth = _tight(oracle, S, kA, 0)
    eu, ep = _errors(S, x, truth)
    assert eu < 1e-6 and ep < 1e-5, (eu, ep)


def test_gmres_reference_mode_matches_oracle_iterations(oracle):
    """Reference order + left-preconditioned GMRES(23) + ReductionControl:
    the same method as the reference, so the iteration count agrees."""
    S = system("C1", 8)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, S_lof=0, method=L.FS_METHOD_GMRES_REF,
                                ordering=L.FS_ORDER_REFERENCE, constraints=S.constrained)
    x = np.zeros(S.n)
    ctrl = fa.StokesSolverControl()
    st = lin.solve(x, S.rhs, ctrl)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=1, kS=0)
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    assert rel(x, xo) < 1e-6
