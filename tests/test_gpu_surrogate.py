# SPDX-License-Identifier: 0BSD
"""The surrogate's GPU kernels (BF16 MFMA forward, backward, Adam) against the CPU oracle
oracle/mlp_ref.py, which rounds to bfloat16 where the kernels do. Tolerances: outputs 2e-3 of
their scale; gradients 2e-2 of the gradient's norm per layer block (fp32 sums in another order
can move a bf16-rounded dZ by one unit in the last place); one Adam step 1e-5 absolute."""
import numpy as np
import pytest

from pacuda_amd import surrogate
from pacuda_amd.runtime import Device
import mlp_ref

pytestmark = pytest.mark.gpu

SHAPES = [(14, 256, 4, 256), (14, 64, 2, 128), (5, 128, 3, 64)]


@pytest.mark.parametrize("inputs,hidden,layers,batch", SHAPES)
def test_forward_backward_adam(inputs, hidden, layers, batch):
    print("mlp", inputs, hidden, layers, batch, flush=True)
    rng = np.random.default_rng(inputs + hidden + layers)
    dev = Device()
    net = surrogate.MLP(inputs, hidden, layers, max_batch=batch, lib=dev.lib)
    p = mlp_ref.init(inputs, hidden, layers, seed=3)
    p = (p + rng.normal(0, 0.02, p.size)).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    x = rng.uniform(0, 1, (batch, inputs)).astype(np.float32)
    t = rng.normal(0, 1, batch).astype(np.float32)
    d_x, d_t, d_y = dev.array(x), dev.array(t), dev.empty(batch, np.float32)
    net.forward(d_x, batch, d_y, train=True)
    y = d_y.download(batch, np.float32)
    want_y, _, _ = mlp_ref.forward(p, x, inputs, hidden, layers)
    assert np.max(np.abs(y - want_y)) <= 2e-3 * max(1.0, np.max(np.abs(want_y)))
    loss = net.backward(d_y, d_t, batch)
    want_loss, want_g = mlp_ref.backward(p, x, t, inputs, hidden, layers)
    assert abs(loss - want_loss) <= 1e-3 * want_loss
    g = net.grads()
    at = 0
    for l in range(layers):
        fan = inputs if l == 0 else hidden
        for size in (hidden * fan, hidden):
            a, b = g[at:at + size], want_g[at:at + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, (l, size)
            at += size
    a, b = g[at:], want_g[at:]
    assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7
    # one Adam step from zero moments on the GPU's own gradient
    net.adam(1e-3)
    q = net.params()
    want_q, _, _ = mlp_ref.adam(p.astype(np.float64), g.astype(np.float64), 0, 0, 1, lr=1e-3)
    assert np.max(np.abs(q - want_q)) <= 1e-5
    for d in (d_x, d_t, d_y):
        d.free()
    net.destroy()


def test_training_reduces_loss():
    """a smooth target learned by the GPU's own forward/backward/Adam"""
    rng = np.random.default_rng(9)
    dev = Device()
    inputs, batch = 14, 1024
    net = surrogate.MLP(inputs, 256, 4, max_batch=batch, lib=dev.lib)
    net.set_params(mlp_ref.init(inputs, 256, 4, seed=4))
    x = rng.uniform(0, 1, (8 * batch, inputs)).astype(np.float32)
    t = (np.sin(3 * x[:, 0]) + x[:, 1] * x[:, 2] - np.log(0.1 + x[:, 3])).astype(np.float32)
    t = (t - t.mean()) / t.std()
    d_x, d_t, d_y = dev.array(x), dev.array(t), dev.empty(len(t), np.float32)
    from pacuda_amd.runtime import DeviceArray
    losses = []
    for epoch in range(30):
        for b in range(8):
            xb = DeviceArray.view(d_x, b * batch * inputs * 4)
            tb = DeviceArray.view(d_t, b * batch * 4)
            yb = DeviceArray.view(d_y, b * batch * 4)
            net.forward(xb, batch, yb, train=True)
            losses.append(net.backward(yb, tb, batch))
            net.adam(1e-3)
    assert np.mean(losses[-8:]) < 0.1 * np.mean(losses[:8]), (losses[:8], losses[-8:])
    net.destroy()


@pytest.mark.parametrize("inputs,hidden,layers,batch", [(14, 256, 4, 4096), (5, 128, 3, 192)])
def test_native_training_loop_equals_the_calls(inputs, hidden, layers, batch):
    """pacuda_mlp_train_steps gives the parameters of forward(train) + backward + adam per batch,
    with the cosine learning rate, bit for bit"""
    import math
    rng = np.random.default_rng(hidden + layers)
    dev = Device()
    steps, total = 5, 12
    x = rng.uniform(0, 1, (steps * batch, inputs)).astype(np.float32)
    t = rng.normal(0, 1, steps * batch).astype(np.float32)
    d_x, d_t, d_y = dev.array(x), dev.array(t), dev.empty(batch, np.float32)
    p0 = mlp_ref.init(inputs, hidden, layers, seed=6)
    nets = [surrogate.MLP(inputs, hidden, layers, max_batch=batch, lib=dev.lib) for _ in range(2)]
    from pacuda_amd.runtime import DeviceArray
    lr = 2e-3
    first = 3
    for net in nets:
        net.set_params(p0)
    a, b = nets
    a.step = first
    for s in range(steps):
        xb = DeviceArray.view(d_x, s * batch * inputs * 4)
        tb = DeviceArray.view(d_t, s * batch * 4)
        a.forward(xb, batch, d_y, train=True)
        loss_a = a.backward(d_y, tb, batch)
        a.adam(lr * 0.5 * (1 + math.cos(math.pi * (first + s) / total)))
    loss_b = b.train_steps(d_x, d_t, batch, steps, lr, first, total)
    assert np.array_equal(a.params(), b.params())
    assert np.array_equal(a.grads(), b.grads())
    assert loss_a == loss_b
    for d in (d_x, d_t, d_y):
        d.free()
    for net in nets:
        net.destroy()


@pytest.mark.parametrize("layers,batch", [(1, 64), (2, 192), (4, 4096 + 64), (2, 32768)])
def test_inference_kernel(layers, batch):
    """the inference kernel (hidden 256, forward without train) against the oracle and against
    the training forward (16- and 64-point workgroups), on whole and ragged 128-point tiles"""
    inputs, hidden = 14, 256
    print("mlp inference", layers, batch, flush=True)
    rng = np.random.default_rng(layers * 7 + batch)
    dev = Device()
    net = surrogate.MLP(inputs, hidden, layers, max_batch=batch, lib=dev.lib)
    p = mlp_ref.init(inputs, hidden, layers, seed=5)
    p = (p + rng.normal(0, 0.02, p.size)).astype(np.float32)
    net.set_params(p)
    x = rng.uniform(0, 1, (batch, inputs)).astype(np.float32)
    d_x, d_y, d_y2 = dev.array(x), dev.empty(batch, np.float32), dev.empty(batch, np.float32)
    net.forward(d_x, batch, d_y, train=False)
    net.forward(d_x, batch, d_y2, train=True)
    y, y2 = d_y.download(batch, np.float32), d_y2.download(batch, np.float32)
    want_y, _, _ = mlp_ref.forward(p, x, inputs, hidden, layers)
    scale = max(1.0, np.max(np.abs(want_y)))
    assert np.max(np.abs(y - want_y)) <= 2e-3 * scale
    assert np.max(np.abs(y - y2)) <= 2e-3 * scale
    for d in (d_x, d_y, d_y2):
        d.free()
    net.destroy()
