"""CPU float64 restatement of the chemistry surrogate network and its output projection.

TEST INFRASTRUCTURE ONLY (the oracle): imported by tests/ and never by the product path.
The surrogate has no counterpart in the reference; this module states what the GPU
kernels compute so that they can be checked:
  - forward: h = GELU(W_l h + b_l) over the hidden layers, y = W_out h + b_out, with GELU
    in its tanh form, GELU(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)));
  - mean-squared-error loss over the batch and outputs, and its gradient by the chain rule;
  - the element-conservation projection with an active set (every element's mass kept,
    no negative mass fraction).
Pins (tests/test_surrogate_oracle.py): gradients against central finite differences, the
projection against a general constrained least-squares solve and its defining properties.
"""
from __future__ import annotations

import math

import numpy as np

def layout(n_in, hidden, n_hidden, n_out):
    """[(in, out, w_offset, b_offset)] per layer and the parameter count, in the library's
    order: W [out][in] row-major, then b [out], layer by layer."""
    dims = [n_in] + [hidden] * n_hidden + [n_out]
    out, off = [], 0
    for l in range(len(dims) - 1):
        i, o = dims[l], dims[l + 1]
        out.append((i, o, off, off + o * i))
        off += o * i + o
    return out, off


def unpack(params, lay):
    p = np.asarray(params, dtype=np.float64)
    return [(p[w:w + o * i].reshape(o, i), p[b:b + o]) for (i, o, w, b) in lay]


_K = math.sqrt(2.0 / math.pi)
_C = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_K * (z + _C * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_K * (z + _C * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _K * (1.0 + 3.0 * _C * z * z)


def forward(params, lay, X):
    """X [batch][n_in] -> Y [batch][n_out]; also the pre-activations and activations."""
    Ws = unpack(params, lay)
    h = np.asarray(X, dtype=np.float64)
    Zs, As = [], [h]
    for l, (W, b) in enumerate(Ws):
        z = h @ W.T + b
        Zs.append(z)
        h = gelu(z) if l < len(Ws) - 1 else z
        As.append(h)
    return h, Zs, As


def loss_grad(params, lay, X, Y):
    """Mean over batch and outputs of the squared error, and its gradient (flat, layout order)."""
    Yp, Zs, As = forward(params, lay, X)
    Y = np.asarray(Y, dtype=np.float64)
    d = Yp - Y
    loss = float(np.mean(d * d))
    dZ = 2.0 * d / d.size
    Ws = unpack(params, lay)
    grad = np.zeros(len(np.asarray(params)))
    for l in range(len(Ws) - 1, -1, -1):
        i, o, w, b = lay[l]
        grad[w:w + o * i] = (dZ.T @ As[l]).reshape(-1)
        grad[b:b + o] = dZ.sum(axis=0)
        if l == 0:
            break
        dA = dZ @ Ws[l][0]
        dZ = dA * gelu_grad(Zs[l - 1])
    return loss, grad


def project(Y0, d0, E):
    """Least-squares change d nearest d0 with E d = 0 and Y0 + d >= 0 (active set: species
    driven negative are fixed at d = -Y0 and the rest re-projected).  Returns None when the
    free species cannot carry the element balance."""
    Y0 = np.asarray(Y0, dtype=np.float64)
    d0 = np.asarray(d0, dtype=np.float64)
    E = np.asarray(E, dtype=np.float64)
    nsp = Y0.size
    fixed = np.zeros(nsp, dtype=bool)
    for _ in range(nsp + 1):
        U = ~fixed
        rhs = E @ np.where(fixed, -Y0, d0)
        G = E[:, U] @ E[:, U].T
        try:
            L = np.linalg.cholesky(G)
        except np.linalg.LinAlgError:
            return None
        if np.min(np.diag(L) ** 2) <= 1e-14:
            return None
        lam = np.linalg.solve(G, rhs)
        d = np.where(fixed, -Y0, d0 - E.T @ lam)
        neg = (~fixed) & (Y0 + d < 0.0)
        if not neg.any():
            return d
        fixed |= neg
    return None
