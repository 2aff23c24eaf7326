# SPDX-License-Identifier: 0BSD
"""CPU oracle of the surrogate MLP (numpy). TEST INFRASTRUCTURE ONLY (tests/ and the benchmark's
checker import it; the product never does).

The network: x -> [W_l, b_l, SiLU] x layers -> w_out . h + b_out, loss mean((y - t)^2).
`forward`/`backward` follow the GPU kernels' rounding: inputs, weights, hidden activations and
the pre-activation gradients dZ are rounded to bfloat16 (round to nearest even) where the GPU
feeds them to the matrix cores; everything else is float64 here (float32 on the GPU). With
`bf16=False` the same code is plain float64, which the finite-difference test pins."""
import numpy as np


def to_bf16(a):
    """round float32 values to the nearest bfloat16 (ties to even), returned as float32"""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _r(a, bf16):
    return to_bf16(a).astype(np.float64) if bf16 else np.asarray(a, dtype=np.float64)


def unpack(params, inputs, hidden, layers):
    p = np.asarray(params, dtype=np.float64)
    out, at = [], 0
    for l in range(layers):
        fan = inputs if l == 0 else hidden
        W = p[at:at + hidden * fan].reshape(hidden, fan)
        at += hidden * fan
        b = p[at:at + hidden]
        at += hidden
        out.append((W, b))
    wout = p[at:at + hidden]
    bout = p[at + hidden]
    assert at + hidden + 1 == len(p)
    return out, wout, bout


def param_count(inputs, hidden, layers):
    return hidden * inputs + hidden + (layers - 1) * (hidden * hidden + hidden) + hidden + 1


def init(inputs, hidden, layers, seed=0):
    """He-style initialisation (float32)"""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(layers):
        fan = inputs if l == 0 else hidden
        parts.append(rng.normal(0, np.sqrt(2.0 / fan), hidden * fan))
        parts.append(np.zeros(hidden))
    parts.append(rng.normal(0, np.sqrt(1.0 / hidden), hidden))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(params, x, inputs, hidden, layers, bf16=True):
    Ws, wout, bout = unpack(params, inputs, hidden, layers)
    h = _r(x, bf16)
    zs, hs = [], [h]
    for W, b in Ws:
        z = h @ _r(W, bf16).T + b
        h = _r(z * _sig(z), bf16)
        zs.append(z)
        hs.append(h)
    y = h @ wout + bout
    return y, zs, hs


def backward(params, x, t, inputs, hidden, layers, bf16=True):
    """(loss, gradient in the parameter layout)"""
    Ws, wout, bout = unpack(params, inputs, hidden, layers)
    y, zs, hs = forward(params, x, inputs, hidden, layers, bf16)
    B = len(t)
    d = y - t
    loss = np.mean(d * d)
    delta = 2.0 * d / B
    grads_w, grads_b = [None] * layers, [None] * layers
    g_wout = hs[-1].T @ delta
    g_bout = delta.sum()
    s = _sig(zs[-1])
    dz = _r(delta[:, None] * wout[None, :] * (s * (1 + zs[-1] * (1 - s))), bf16)
    for l in range(layers - 1, -1, -1):
        grads_w[l] = dz.T @ hs[l]
        grads_b[l] = dz.sum(axis=0)
        if l > 0:
            dh = dz @ _r(Ws[l][0], bf16)
            s = _sig(zs[l - 1])
            dz = _r(dh * (s * (1 + zs[l - 1] * (1 - s))), bf16)
    parts = []
    for l in range(layers):
        parts += [grads_w[l].reshape(-1), grads_b[l]]
    parts += [g_wout, np.array([g_bout])]
    return loss, np.concatenate(parts)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
