"""Test infrastructure: numpy restatement of the field-surrogate MLP of
traceon_amd/csrc/mlp.hip (2-128-128-128-128-5, SiLU), forward and backward.

``rounding=True`` emulates the kernels' BF16 rounding points: the activations entering the
hidden MFMA layers, the hidden weights, and the hidden-layer gradients entering the MFMA.
Everything else is FP32 in the kernels and float64 here.  ``rounding=False`` is the plain
float64 network; its gradient is pinned against central finite differences
(tests/test_oracle.py::test_mlp_oracle_gradient_vs_finite_differences).
"""
from __future__ import annotations

import numpy as np

W, OUT = 128, 5
OFF = {}
_o = 0
for _name, _size in (("W1", W * 2), ("b1", W), ("W2", W * W), ("b2", W), ("W3", W * W), ("b3", W), ("W4", W * W), ("b4", W),
                     ("W5", OUT * W), ("b5", OUT)):
    OFF[_name] = (_o, _size)
    _o += _size
NPARAM = _o


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64 values."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    out = {}
    for k, (o, n) in OFF.items():
        v = P[o:o + n]
        if k == "W1":
            v = v.reshape(W, 2)
        elif k in ("W2", "W3", "W4"):
            v = v.reshape(W, W)
        elif k == "W5":
            v = v.reshape(OUT, W)
        out[k] = v
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, X, rounding=True):
    """Y (B,5) and the cache of pre-activations / MFMA inputs."""
    p = unpack(P)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ p["W1"].T + p["b1"]
    H = [r(silu(Z1))]
    Z = [Z1]
    for l in ("2", "3", "4"):
        Zl = H[-1] @ r(p["W" + l]).T + p["b" + l]
        Z.append(Zl)
        H.append(r(silu(Zl)))
    Y = silu(Z[3]) @ p["W5"].T + p["b5"]
    return Y, (X, Z, H, p, r)


def grad(P, X, T, rounding=True):
    """Gradient of mean((Y - T)^2) over batch and outputs (flat, NPARAM) and the sum of squares."""
    Y, (X, Z, H, p, r) = forward(P, X, rounding)
    B = X.shape[0]
    D = Y - np.asarray(T, dtype=np.float64)
    dY = 2.0 * D / (OUT * B)
    G = {}
    h4 = silu(Z[3])
    G["W5"] = dY.T @ h4
    G["b5"] = dY.sum(axis=0)
    dZ = (dY @ p["W5"]) * dsilu(Z[3])
    for l, (zi, hi) in zip(("4", "3", "2"), ((2, 2), (1, 1), (0, 0))):
        dZr = r(dZ)
        G["W" + l] = dZr.T @ H[hi]
        G["b" + l] = dZ.sum(axis=0)
        dH = dZr @ r(p["W" + l])
        dZ = dH * dsilu(Z[zi])
    G["W1"] = dZ.T @ X
    G["b1"] = dZ.sum(axis=0)
    flat = np.zeros(NPARAM)
    for k, (o, n) in OFF.items():
        flat[o:o + n] = G[k].ravel()
    return flat, float(np.sum(D * D))


def init_params(seed=0, scale=1.0):
    """He-style initialisation (the surrogate's own init is in traceon_amd/surrogate.py)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM)
    for k, (o, n) in OFF.items():
        if k.startswith("W"):
            fan_in = 2 if k == "W1" else W
            P[o:o + n] = rng.standard_normal(n) * scale * np.sqrt(2.0 / fan_in)
        else:
            P[o:o + n] = rng.standard_normal(n) * 0.1 * scale
    return P
