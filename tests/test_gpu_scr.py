"""GPU parity of the Schur-complement reduction (StokesSolutionMethod::SCR,
src/pde/stokes_ilu.cc:564-710) and of CompareTest (:471-505) against the CPU
oracle's restatement on identical matrices, in the reference row order (so
the ILU factors are the reference's)."""
import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from conftest import rel, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n,kA", [("C1", 8, 0), ("C2", 8, 1), ("C3", 3, 0)])
def test_scr_matches_oracle(oracle, cfg, n, kA):
    S = system(cfg, n)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, method=L.FS_METHOD_SCR, ordering=L.FS_ORDER_REFERENCE,
                                constraints=S.constrained)
    x = np.zeros(S.n)
    ctrl = fa.StokesSolverControl()
    st = lin.solve(x, S.rhs, ctrl)
    xo, sto = oracle.stokes_solve(S.blocks, S.rhs, kA=kA, kS=0, method=1)
    assert sto["status"] == 0 and st["status"] == L.FS_OK
    assert abs(st["outer_iters"] - sto["outer_iters"]) <= 1, (st["outer_iters"], sto["outer_iters"])
    assert rel(x[:S.n_u], xo[:S.n_u]) < 1e-6
    dp, dpo = x[S.n_u:] - x[S.n_u:].mean(), xo[S.n_u:] - xo[S.n_u:].mean()
    assert rel(dp, dpo) < 1e-5
    assert ctrl.log_cg and ctrl.n_cg_outer_iter[-1] == st["outer_iters"]


def test_compare_test_records_difference():
    S = system("C1", 8)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, method=L.FS_METHOD_GMRES_REF, ordering=L.FS_ORDER_REFERENCE,
                                constraints=S.constrained)
    ctrl = fa.StokesSolverControl()
    x = np.zeros(S.n)
    out = lin.solve(x, S.rhs, ctrl, solution_method=fa.CompareTest)
    assert ctrl.log_cg and ctrl.log_gmres
    assert len(ctrl.soln_diff_l2) == 1 and ctrl.soln_diff_l2[0] == out["soln_diff_l2"]
    # both methods stop at a 1e-6 reduction of their own residuals
    assert out["soln_diff_l2"] < 1e-3
    # the handle keeps its FC method afterwards
    assert lin.opts.method == L.FS_METHOD_GMRES_REF


def test_scr_rejects_mixed_precision():
    S = system("C1", 4)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=0, method=L.FS_METHOD_SCR, precision=L.FS_PREC_MIXED)
    with pytest.raises(L.FsException) as e:
        lin.solve(np.zeros(S.n), S.rhs)
    assert e.value.code == L.FS_ERR_UNSUPPORTED
