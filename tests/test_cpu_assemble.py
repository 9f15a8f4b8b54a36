"""The Q2-Q1 generator: sizes, symmetry, and closed-form solutions that
Q2-Q1 reproduces exactly (linear velocity, linear pressure)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import felspa_amd as fa
from felspa_amd import _lib as L

from conftest import rel


def _solve(S):
    return spla.spsolve(S.K().tocsc(), S.rhs)


def _vel_coords(S):
    n1 = 2 * S.n_cells + 1
    node = S.vnode
    c = np.stack([(node // n1 ** k) % n1 for k in range(S.dim)], axis=1)
    return c / (n1 - 1)


def _p_coords(S):
    n1 = S.n_cells + 1
    node = S.pnode
    c = np.stack([(node // n1 ** k) % n1 for k in range(S.dim)], axis=1)
    return c / (n1 - 1)


@pytest.mark.parametrize("dim,n", [(2, 4), (2, 9), (3, 3)])
def test_sizes_and_symmetry(dim, n):
    S = fa.StokesSystem(fa.problem(dim, n, L.FS_FIELD_INCLUSION, L.FS_BC_LINEAR_SHEAR))
    assert S.n_u == dim * (2 * n + 1) ** dim and S.n_p == (n + 1) ** dim
    b = S.blocks
    assert abs(b["A"] - b["A"].T).max() < 1e-12 * abs(b["A"]).max()
    assert abs(b["Mp"] - b["Mp"].T).max() < 1e-12 * abs(b["Mp"]).max()
    assert abs(b["B"] - b["Bt"].T).max() == 0
    assert b["A"].nnz <= dim * dim * (8 * n + 1) ** dim
    assert b["B"].nnz <= dim * (5 * n + 1) ** dim
    assert b["K11"].nnz == S.n_p


@pytest.mark.parametrize("dim,n", [(2, 6), (3, 3)])
def test_linear_shear_is_exact(dim, n):
    """Constant viscosity + LinearSimpleShear BCs: u_0 = (x_{d-1}-0.5)/0.5,
    other components 0, p = 0 is in the Q2-Q1 space and is the solution."""
    S = fa.StokesSystem(fa.problem(dim, n, L.FS_FIELD_CONSTANT, L.FS_BC_LINEAR_SHEAR))
    x = _solve(S)
    X = _vel_coords(S)
    want = np.where(S.vcomp == 0, (X[np.arange(S.n_u), dim - 1] - 0.5) / 0.5, 0.0)
    assert np.abs(x[:S.n_u] - want).max() < 1e-10
    assert np.abs(x[S.n_u:]).max() < 1e-9


def test_periodic_topbottom_shear_is_exact():
    S = fa.StokesSystem(fa.problem(2, 8, L.FS_FIELD_CONSTANT, L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC))
    x = S.distribute(_solve(S))
    X = _vel_coords(S)
    want = np.where(S.vcomp == 0, 2 * X[:, 1] - 1, 0.0)
    assert np.abs(x[:S.n_u] - want).max() < 1e-10
    assert np.abs(x[S.n_u:]).max() < 1e-9


def test_hydrostatic_pressure_is_exact():
    """Free slip, constant density, gravity g: u = 0, p = -rho g y."""
    S = fa.StokesSystem(fa.problem(2, 6, L.FS_FIELD_CONSTANT, L.FS_BC_FREE_SLIP, gravity=2.0,
                                   rho_matrix=1.5))
    x = _solve(S)
    assert np.abs(x[:S.n_u]).max() < 1e-10
    Y = _p_coords(S)[:, 1]
    assert np.abs(x[S.n_u:] - (-1.5 * 2.0 * Y)).max() < 1e-9


def test_constraints_and_distribute():
    S = fa.StokesSystem(fa.make_problem("C1", n_cells=6))
    x = S.distribute(np.zeros(S.n))
    dir_ = S.kind == L.FS_DOF_DIRICHLET
    assert np.all(x[dir_] == S.cval[dir_])
    assert (S.kind == L.FS_DOF_PINNED).sum() == 1 and S.kind[S.n_u] == L.FS_DOF_PINNED
    per = np.where(S.kind == L.FS_DOF_PERIODIC)[0]
    assert len(per) > 0
    y = np.arange(S.n, dtype=float)
    y = S.distribute(y)
    assert np.all(y[per] == y[S.master[per]])


def test_ellipse_fields_are_seeded():
    a = fa.StokesSystem(fa.make_problem("C4", n_cells=16, sample=3))
    b = fa.StokesSystem(fa.make_problem("C4", n_cells=16, sample=3))
    c = fa.StokesSystem(fa.make_problem("C4", n_cells=16, sample=4))
    assert np.array_equal(a.log_eta, b.log_eta)
    assert not np.array_equal(a.log_eta, c.log_eta)
    assert a.log_eta.min() >= 0 and a.log_eta.max() <= 3.0 + 1e-6


def test_inclusion_viscosity_rule():
    """SphericalInclusion: inside -> 1e3, |x| <= r outside the sphere -> mean."""
    S = fa.StokesSystem(fa.make_problem("C1", n_cells=20))
    g = S.log_eta.reshape(21, 21)
    assert g[10, 10] == pytest.approx(3.0)
    assert g[0, 0] == pytest.approx(np.log10(0.5 * (1 + 1e3)), rel=1e-6)
    assert g[0, 20] == pytest.approx(0.0)


def test_bad_problem_rejected():
    with pytest.raises(L.FsException):
        fa.StokesSystem(fa.problem(4, 4, 0, 0))
