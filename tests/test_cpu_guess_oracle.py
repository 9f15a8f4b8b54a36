"""Pins of the learned-guess network oracle (oracle/guess_oracle.py) on the
CPU: the cosine analysis/synthesis pair against the DCT-I orthogonality that
fixes it exactly, the input interpolation against a field it must reproduce
exactly, and the hand-written backward pass against central finite
differences of the forward pass (2D and 3D)."""
import numpy as np
import pytest

import guess_oracle as go


def test_cosine_pair_reproduces_cosine_series():
    n1, M = 17, 5
    Af, Sb = go.cos_mats(n1, M)
    x = np.arange(n1) / (n1 - 1)
    rng = np.random.default_rng(0)
    a = rng.standard_normal(M)
    f = sum(a[k] * np.cos(np.pi * k * x) for k in range(M))
    # discrete orthogonality of the DCT-I: analysis returns a_0, a_k / 2
    c = Af @ f
    assert np.allclose(c, a * np.where(np.arange(M) == 0, 1.0, 0.5), atol=1e-13)
    assert np.allclose(Sb @ c, f, atol=1e-13)
    # a mode above the cut-off is annihilated
    assert np.allclose(Af @ np.cos(np.pi * M * x), 0.0, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_prep_input_interpolates_multilinear_fields_exactly(dim):
    nc = 4
    nq, n1 = nc + 1, 2 * nc + 1
    coef = np.array([0.3, -1.1, 0.7, 0.25])[: dim + 1]
    def field(n):
        idx = np.indices((n,) * dim).reshape(dim, -1)  # slowest axis first
        xs = [idx[dim - 1 - a] / (n - 1) for a in range(dim)]  # a = 0 is x
        return coef[0] + sum(coef[1 + a] * xs[a] for a in range(dim)), xs
    q1, _ = field(nq)
    X = go.prep_input(dim, nc, q1[None, :])
    want, xs = field(n1)
    assert np.allclose(X[0, 0] * 3.0, want, atol=1e-14)
    for a in range(dim):
        assert np.allclose(X[0, 1 + a], xs[a], atol=1e-15)


def _fd_check(cfg, n1, B, seed):
    rng = np.random.default_rng(seed)
    _, npar = go.layout(cfg)
    flat = rng.standard_normal(npar) * 0.5
    P = n1 ** cfg["dim"]
    X = rng.standard_normal((B, cfg["dim"] + 1, P))
    T = rng.standard_normal((B, cfg["dim"] + 1, P))
    loss, g = go.Net(cfg, flat).loss_grad(X, T)
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        lp = np.mean((go.Net(cfg, flat + e).forward(X) - T) ** 2)
        lm = np.mean((go.Net(cfg, flat - e).forward(X) - T) ** 2)
        fd[i] = (lp - lm) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert err < 1e-6, err
    # every parameter group is exercised
    lay, _ = go.layout(cfg)
    for name, (o, s) in lay.items():
        assert np.linalg.norm(fd[o:o + int(np.prod(s))]) > 0, name


def test_backward_matches_finite_differences_2d():
    _fd_check(dict(dim=2, width=3, modes=3, layers=2, proj=4), n1=7, B=2, seed=1)


def test_backward_matches_finite_differences_3d():
    _fd_check(dict(dim=3, width=2, modes=2, layers=2, proj=3), n1=5, B=1, seed=2)


def test_layout_counts():
    cfg = dict(dim=2, width=32, modes=12, layers=4, proj=64)
    lay, n = go.layout(cfg)
    w, cin = 32, 3
    want = w * cin + w + 4 * (144 * w * w + w * w + w) + 64 * w + 64 + 3 * 64 + 3
    assert n == want
    assert lay["W_lift"][0] == 0 and lay["b_p2"][0] + 3 == n
