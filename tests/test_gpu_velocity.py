"""Velocity hand-off (VelocityExtractor::extract, stokes_common.implement.h:
582-591) on the GPU: the solution's velocity on the Q2 node lattice and the
Q2 field evaluated at points, against a host scatter, against quadratic
fields the Q2 basis reproduces exactly, and against the analytic simple-shear
solution of a constant-viscosity solve."""
import numpy as np
import pytest

import felspa_amd as fa
from felspa_amd import _lib as L
from felspa_amd.system import StokesSystem, problem
from conftest import system

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _handle(S, **kw):
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, constraints=S.constrained, block_hint=S.block_hint(3 if S.dim == 3 else 5),
                                **kw)
    lin.set_dof_lattice(S)
    return lin


@pytest.mark.parametrize("cfg,n", [("C4", 16), ("C3", 4)])
def test_velocity_lattice_matches_host_scatter(cfg, n):
    import torch
    S = system(cfg, n)
    lin = _handle(S)
    x = np.random.default_rng(3).standard_normal(S.n)
    P = (2 * n + 1) ** S.dim
    d_u = torch.full((S.dim * P,), float("nan"), dtype=torch.float64, device="cuda")
    d_x = _dev(x)
    torch.cuda.synchronize()
    lin.velocity_lattice(d_x.data_ptr(), d_u.data_ptr())
    want = np.zeros((S.dim, P))
    want[S.vcomp.astype(np.int64), S.vnode] = x[:S.n_u]
    assert np.array_equal(d_u.cpu().numpy(), want.ravel())
    lin.close()


@pytest.mark.parametrize("cfg,n", [("C4", 6), ("C3", 3)])
def test_points_reproduce_quadratic_fields(cfg, n):
    """Every quadratic polynomial (per coordinate) lies in the Q2 space: set
    on the nodes (node j of axis a at j / 2N), it is reproduced exactly at
    arbitrary points, including points on cell faces and outside the cube
    (clamped)."""
    import torch
    S = system(cfg, n)
    d = S.dim
    lin = _handle(S)
    n1 = 2 * n + 1
    idx = np.indices((n1,) * d).reshape(d, -1)[::-1] / (2.0 * n)  # coordinate a of node k: axis a x fastest
    coef = np.random.default_rng(5).standard_normal((d, 10))

    def f(c, X):  # a quadratic in each coordinate, with cross terms
        r = coef[c, 0] + sum(coef[c, 1 + a] * X[a] + coef[c, 4 + a] * X[a] ** 2 for a in range(d))
        return r + coef[c, 7] * X[0] * X[1] * (X[2] if d == 3 else 1.0)
    U = np.stack([f(c, idx) for c in range(d)])
    pts = np.random.default_rng(6).uniform(0, 1, (400, d))
    pts[:20] = np.round(pts[:20] * n) / n  # on cell faces and nodes
    d_u, d_p = _dev(U.ravel()), _dev(pts.ravel())
    d_v = torch.zeros(400 * d, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    lin.velocity_at_points(d_u.data_ptr(), 400, d_p.data_ptr(), d_v.data_ptr())
    got = d_v.cpu().numpy().reshape(400, d)
    want = np.stack([f(c, pts.T) for c in range(d)], axis=1)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    lin.close()


def test_simple_shear_velocity_at_points():
    """Constant viscosity, linear simple shear on every face (u_0 = (x_{d-1} -
    0.5) / 0.5, other components 0): the discrete solution is the linear
    profile itself, so the handed-off velocity at arbitrary points is."""
    import torch
    S = StokesSystem(problem(dim=3, n_cells=4, field=L.FS_FIELD_CONSTANT, bc=L.FS_BC_LINEAR_SHEAR,
                             eta_matrix=1.0, shear_velocity=1.0))
    lin = _handle(S, rtol=1e-12)
    x = np.zeros(S.n)
    lin.solve(x, S.rhs)
    x = S.distribute(x)
    P = 9 ** 3
    d_x = _dev(x)
    d_u = torch.zeros(3 * P, dtype=torch.float64, device="cuda")
    pts = np.random.default_rng(7).uniform(0, 1, (300, 3))
    d_p = _dev(pts.ravel())
    d_v = torch.zeros(900, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    lin.velocity_lattice(d_x.data_ptr(), d_u.data_ptr())
    lin.velocity_at_points(d_u.data_ptr(), 300, d_p.data_ptr(), d_v.data_ptr())
    v = d_v.cpu().numpy().reshape(300, 3)
    assert np.abs(v[:, 0] - (pts[:, 2] - 0.5) / 0.5).max() < 1e-9
    assert np.abs(v[:, 1:]).max() < 1e-9
    lin.close()


def test_velocity_handoff_needs_the_lattice():
    import torch
    S = system("C4", 6)
    lin = fa.StokesLinearSystem(S.blocks, A_lof=1, block_hint=S.block_hint(5))
    d = torch.zeros(S.n, dtype=torch.float64, device="cuda")
    with pytest.raises(L.FsException):
        lin.velocity_lattice(d.data_ptr(), d.data_ptr())
    lin.close()
