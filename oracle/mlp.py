"""CPU restatement of the learned limiter's network -- TEST INFRASTRUCTURE ONLY.

numpy float64 forward and backward pass of the 10-64-64-1 network (ReLU,
ReLU, sigmoid) with the bf16 roundings placed exactly where the device
kernels round (inputs, weights, hidden activations, and the back-propagated
pre-activation gradients that enter a matrix product), so the device's BF16
matrix-core results can be checked to FP32-accumulation accuracy. The loss
is scale * sum((sigmoid(z) - y)^2) over rows with y >= 0.

Parameter layout (flattened FP32): W1[10][64], b1[64], W2[64][64], b2[64],
w3[64], b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 10, 64
OFF = {}
_o = 0
for _k, _n in (("W1", N_IN * N_HID), ("b1", N_HID), ("W2", N_HID * N_HID), ("b2", N_HID), ("w3", N_HID),
               ("b3", 1)):
    OFF[_k] = (_o, _o + _n)
    _o += _n
N_PARAMS = _o


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = lambda k: p[OFF[k][0]:OFF[k][1]].astype(np.float64)  # noqa: E731
    return (g("W1").reshape(N_IN, N_HID), g("b1"), g("W2").reshape(N_HID, N_HID), g("b2"), g("w3"), g("b3")[0])


def exact(x):
    """No rounding (float64 throughout): the finite-difference check uses it."""
    return np.asarray(x, dtype=np.float64)


def forward(p, X, rnd=bf16):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xb = rnd(X)
    z1 = xb @ rnd(W1) + b1
    a1 = rnd(np.maximum(z1, 0.0))
    z2 = a1 @ rnd(W2) + b2
    a2 = rnd(np.maximum(z2, 0.0))
    z3 = a2 @ rnd(w3) + b3
    return 1.0 / (1.0 + np.exp(-z3)), dict(xb=xb, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3)


def gradient(p, X, y, scale, rnd=bf16):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    pr, c = forward(p, X, rnd)
    y = np.asarray(y, dtype=np.float64)
    valid = y >= 0
    d = np.where(valid, pr - y, 0.0)
    loss = float(np.sum(d * d))
    dz3 = scale * 2.0 * d * pr * (1.0 - pr)
    g = np.zeros(N_PARAMS)
    g[OFF["b3"][0]] = dz3.sum()
    g[OFF["w3"][0]:OFF["w3"][1]] = c["a2"].T @ dz3
    dz2 = np.where(c["z2"] > 0, rnd(w3)[None, :] * dz3[:, None], 0.0)
    g[OFF["b2"][0]:OFF["b2"][1]] = dz2.sum(axis=0)
    dz2b = rnd(dz2)
    g[OFF["W2"][0]:OFF["W2"][1]] = (c["a1"].T @ dz2b).ravel()
    da1 = dz2b @ rnd(W2).T
    dz1 = np.where(c["z1"] > 0, da1, 0.0)
    g[OFF["b1"][0]:OFF["b1"][1]] = dz1.sum(axis=0)
    g[OFF["W1"][0]:OFF["W1"][1]] = (c["xb"].T @ rnd(dz1)).ravel()
    return g, loss


def glorot(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    for k, (a, b) in OFF.items():
        if k == "W1":
            lim = np.sqrt(6.0 / (N_IN + N_HID))
        elif k == "W2":
            lim = np.sqrt(6.0 / (2 * N_HID))
        elif k == "w3":
            lim = np.sqrt(6.0 / (N_HID + 1))
        else:
            continue
        p[a:b] = rng.uniform(-lim, lim, b - a)
    return p
