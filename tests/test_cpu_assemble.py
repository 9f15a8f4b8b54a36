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


def test_quadrature_fields_inclusion_and_layers():
    """The samples the assembly uses: the field at (cell + Gauss point) * h,
    with Gauss points 0.5 -+ sqrt(3/5)/2 and 0.5 (3-point Gauss-Legendre on
    [0, 1]), restated with numpy; inclusion and layered fields."""
    g = np.sqrt(3.0 / 5.0) / 2.0
    gp = np.array([0.5 - g, 0.5, 0.5 + g])
    for dim, n in ((2, 7), (3, 4)):
        S = fa.StokesSystem(fa.problem(dim, n, L.FS_FIELD_INCLUSION, L.FS_BC_FREE_SLIP, eta_matrix=2.0,
                                 eta_incl=50.0, rho_matrix=1.0, rho_incl=3.0, radius=0.31))
        eta, rho = S.quadrature_fields()
        assert eta.shape == (n ** dim, 3 ** dim)
        cells = np.indices((n,) * dim).reshape(dim, -1)[::-1]          # axis k of cell c, x fastest
        qs = np.indices((3,) * dim).reshape(dim, -1)[::-1]             # Gauss index k of point q
        x = (cells[:, :, None] + gp[qs][:, None, :]) / n                # [dim, cell, q]
        dist = np.sqrt(((x - 0.5) ** 2).sum(0))
        nrm = np.sqrt((x ** 2).sum(0))
        want = np.where(dist < 0.31, 50.0, np.where(nrm > 0.31, 2.0, 26.0))
        assert np.array_equal(eta, want)
        assert np.array_equal(rho, np.where(dist < 0.31, 3.0, np.where(nrm > 0.31, 1.0, 2.0)))
        S = fa.StokesSystem(fa.problem(dim, n, L.FS_FIELD_LAYERS, L.FS_BC_FREE_SLIP, eta_matrix=1.0,
                                 eta_incl=10.0, amp=0.1))
        eta, _ = S.quadrature_fields()
        y = 0.5 + 0.1 * np.cos(2 * np.pi * x[0])
        upper = x[dim - 1] > y
        near = np.abs(x[dim - 1] - y) < 1e-12
        assert np.array_equal(eta[~near], np.where(upper, 10.0, 1.0)[~near])


def test_pinned_pressure_diagonal_is_mean_free_diagonal():
    """Constrained pressure rows carry K11 = mean |A_ii| over the free
    velocity rows (summed by row chunks; equal to the plain mean to rounding)."""
    S = fa.StokesSystem(fa.problem(2, 12, L.FS_FIELD_INCLUSION, L.FS_BC_TOPBOTTOM_SHEAR_PERIODIC))
    A, K11 = S.blocks["A"], S.blocks["K11"]
    free = ~S.constrained[:S.n_u]
    want = np.abs(A.diagonal()[free]).mean()
    pc = S.constrained[S.n_u:]
    assert pc.sum() >= 2  # the pinned dof and the periodic pressure slaves
    assert np.allclose(K11.diagonal()[pc], want, rtol=1e-14, atol=0)
    assert np.all(K11.diagonal()[~pc] == 0.0)
