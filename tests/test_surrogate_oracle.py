"""Pins of the surrogate-network oracle (oracle/mlp_oracle.py), no GPU: its analytic
gradient against central finite differences, and its projection against a general
constrained least-squares solve (scipy SLSQP) and the properties that define it."""
import numpy as np
import pytest
from scipy.optimize import minimize

from gdtk_amd.chem import element_matrix
from oracle import mlp_oracle as M


def _small_net(seed=0, n_in=8, hidden=64, n_hidden=2, n_out=5):
    lay, npar = M.layout(n_in, hidden, n_hidden, n_out)
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.3, npar)
    return lay, p


def test_layout_counts():
    lay, n = M.layout(8, 256, 3, 5)
    assert n == 8 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 5 + 5 == 135173
    assert lay[-1][3] + 5 == n


def test_gradient_matches_finite_differences():
    lay, p = _small_net()
    rng = np.random.default_rng(1)
    X = rng.normal(size=(7, 8))
    Y = rng.normal(size=(7, 5))
    loss, g = M.loss_grad(p, lay, X, Y)
    idx = rng.choice(p.size, 60, replace=False)
    # include every bias block and the first/last weight of every layer
    for (i, o, w, b) in lay:
        idx = np.concatenate([idx, [w, w + o * i - 1, b, b + o - 1]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_grad(pp, lay, X, Y)[0] - M.loss_grad(pm, lay, X, Y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gelu_closed_forms():
    import math
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    # GELU(0) = 0, GELU(z) - GELU(-z) = z, derivative by finite differences, and the tanh
    # form stays within 1e-3 of the erf definition z Phi(z)
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert np.allclose(M.gelu(z) - M.gelu(-z), z, atol=1e-15)
    h = 1e-6
    assert np.allclose((M.gelu(z + h) - M.gelu(z - h)) / (2 * h), M.gelu_grad(z), atol=1e-8)
    zz = np.linspace(-6, 6, 241)
    exact = np.array([0.5 * v * (1.0 + math.erf(v / math.sqrt(2.0))) for v in zz])
    assert np.max(np.abs(M.gelu(zz) - exact)) < 1e-3


def _air_E():
    Msp = np.array([0.0280134, 0.0319988, 0.0300061, 0.0140067, 0.0159994])
    atoms = element_matrix(["N2", "O2", "NO", "N", "O"])
    Mel = np.array([0.0140067, 0.0159994])
    return atoms * Mel[:, None] / Msp[None, :]


@pytest.mark.parametrize("seed", range(6))
def test_projection_against_constrained_least_squares(seed):
    E = _air_E()
    assert np.allclose(E.sum(axis=0), 1.0, atol=1e-6)  # each species' mass is all assigned
    rng = np.random.default_rng(seed)
    Y0 = rng.dirichlet(np.ones(5) * 0.7)
    Y0[rng.integers(5)] *= 1e-3  # a nearly absent species
    Y0 /= Y0.sum()
    d0 = rng.normal(0.0, 0.05, 5)
    d = M.project(Y0, d0, E)
    assert d is not None
    assert np.max(np.abs(E @ d)) < 1e-15
    assert np.min(Y0 + d) >= 0.0
    res = minimize(lambda v: 0.5 * np.sum((v - d0) ** 2), np.zeros(5), jac=lambda v: v - d0,
                   constraints=[{"type": "eq", "fun": lambda v: E @ v, "jac": lambda v: E},
                                {"type": "ineq", "fun": lambda v: Y0 + v, "jac": lambda v: np.eye(5)}],
                   method="SLSQP", options={"ftol": 1e-15, "maxiter": 500})
    assert res.success
    # the active-set answer is the constrained minimiser (it is no farther from d0)
    assert np.sum((d - d0) ** 2) <= np.sum((res.x - d0) ** 2) + 1e-10
    assert np.allclose(d, res.x, atol=1e-6)


def test_projection_identity_on_conserving_change():
    E = _air_E()
    Y0 = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    # O2 -> 2 O at fixed mass: conserves elements, so the projection leaves it alone
    d0 = np.array([0.0, -0.01, 0.0, 0.0, 0.01])
    assert np.allclose(M.project(Y0, d0, E), d0, atol=1e-15)
