# SPDX-License-Identifier: 0BSD
"""The surrogate's CPU oracle (oracle/mlp_ref.py) pinned without a GPU: its analytic gradient
equals central finite differences of its own loss in float64, bfloat16 rounding matches the
format's definition, and Adam's first step moves every parameter by lr in the gradient's sign."""
import numpy as np

import mlp_ref


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = mlp_ref.to_bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # a tie rounds to even
    assert r[2] == 1.0 + 2 ** -7             # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8


def test_gradient_matches_finite_differences():
    inputs, hidden, layers = 5, 8, 3
    rng = np.random.default_rng(0)
    p = mlp_ref.init(inputs, hidden, layers, seed=1).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)  # nonzero biases
    x = rng.uniform(0, 1, (16, inputs))
    t = rng.normal(0, 1, 16)
    loss, g = mlp_ref.backward(p, x, t, inputs, hidden, layers, bf16=False)
    h = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        lp = mlp_ref.backward(p + e, x, t, inputs, hidden, layers, bf16=False)[0]
        lm = mlp_ref.backward(p - e, x, t, inputs, hidden, layers, bf16=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    assert mlp_ref.param_count(inputs, hidden, layers) == p.size


def test_adam_first_step():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    q, m, v = mlp_ref.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(q, p - 0.01 * np.sign(g), rtol=0, atol=1e-6)
