"""Pins of the CPU restatement of the learned limiter's network (oracle/mlp.py).

The backward pass is checked against central finite differences of the loss
in float64 with the bf16 roundings switched off (a dropped term, a missing
ReLU mask or a transposed weight fails it), and the bf16 rounding against
hand-computed bit patterns.
"""
import numpy as np
import pytest

from oracle import mlp


def _loss(p, X, y, scale):
    pr, _ = mlp.forward(p.astype(np.float64), X, mlp.exact)
    d = np.where(y >= 0, pr - y, 0.0)
    return scale * float(np.sum(d * d))


def test_bf16_rounding_bits():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.140625, 65504.0, 1e-40], dtype=np.float32)
    r = mlp.bf16(x)
    # 1 + 2^-8 is the tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert r[3] == -3.140625
    assert r[4] == 65536.0
    assert abs(r[5]) < 1e-39


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(5).standard_normal(4096).astype(np.float32) * np.float32(1e3)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_forward_closed_form():
    """A network wired by hand: hidden unit 0 carries relu(2*x0 - 1), unit 1 carries relu(-x3);
    layer 2 passes them through with gains 1 and 3; the output is sigmoid(w.a2 + b3).
    All weights are bf16-exact, so the rounded and the exact forward agree with the formula."""
    p = np.zeros(mlp.N_PARAMS, dtype=np.float32)
    W1 = np.zeros((10, 64), np.float32)
    W1[0, 0], W1[3, 1] = 2.0, -1.0
    b1 = np.zeros(64, np.float32)
    b1[0] = -1.0
    W2 = np.zeros((64, 64), np.float32)
    W2[0, 0], W2[1, 1] = 1.0, 3.0
    w3 = np.zeros(64, np.float32)
    w3[0], w3[1] = 0.5, -0.25
    for k, v in (("W1", W1), ("b1", b1), ("W2", W2), ("w3", w3)):
        p[mlp.OFF[k][0]:mlp.OFF[k][1]] = v.ravel()
    p[mlp.OFF["b3"][0]] = 0.125
    X = np.zeros((5, 10))
    X[:, 0] = [0.0, 0.5, 1.0, 2.0, -1.0]
    X[:, 3] = [1.0, -2.0, -0.5, 0.0, -4.0]
    z = 0.5 * np.maximum(2 * X[:, 0] - 1, 0) - 0.25 * 3 * np.maximum(-X[:, 3], 0) + 0.125
    want = 1.0 / (1.0 + np.exp(-z))
    for rnd in (mlp.bf16, mlp.exact):
        got, _ = mlp.forward(p, X, rnd)
        assert np.allclose(got, want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = mlp.glorot(seed).astype(np.float64)
    p[mlp.OFF["b1"][0]:mlp.OFF["b1"][1]] = rng.uniform(-0.1, 0.1, 64)
    p[mlp.OFF["b2"][0]:mlp.OFF["b2"][1]] = rng.uniform(-0.1, 0.1, 64)
    X = rng.uniform(-2, 2, (40, 10))
    y = rng.uniform(0, 1, 40)
    y[::7] = -1.0
    g, _ = mlp.gradient(p, X, y, 0.5, mlp.exact)
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
                          for a, b in mlp.OFF.values()])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (_loss(pp, X, y, 0.5) - _loss(pm, X, y, 0.5)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), f"param {k}: fd {fd:.6e} analytic {g[k]:.6e}"


def test_forward_shapes_and_range():
    p = mlp.glorot(3)
    X = np.random.default_rng(3).uniform(-8, 8, (100, 10))
    pr, c = mlp.forward(p, X)
    assert pr.shape == (100,) and np.all((pr > 0) & (pr < 1))
    assert c["a1"].shape == (100, 64) and np.all(c["a1"] >= 0)
