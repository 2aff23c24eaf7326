# SPDX-License-Identifier: 0BSD
"""The surrogate network on the GPU (include/pacuda_amd.h, pacuda_mlp_*): a Python handle over the
library's BF16-MFMA forward, backward and Adam kernels, and the phase-space features it reads."""
import ctypes as C
import math

import numpy as np

from . import abi
from .runtime import DeviceArray

F32P = C.POINTER(C.c_float)


def he_init(inputs, hidden, layers, seed=0):
    """He-normal weights, zero biases, in the library's parameter layout (float32)"""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(layers):
        fan = inputs if l == 0 else hidden
        parts.append(rng.normal(0, np.sqrt(2.0 / fan), hidden * fan))
        parts.append(np.zeros(hidden))
    parts.append(rng.normal(0, np.sqrt(1.0 / hidden), hidden))
    parts.append(np.zeros(1))
    return np.concatenate(parts).astype(np.float32)


class MLP:
    def __init__(self, inputs, hidden=256, layers=4, max_batch=4096, lib=None):
        self.lib = lib or abi.load()
        self.inputs, self.hidden, self.layers, self.max_batch = inputs, hidden, layers, max_batch
        h = C.c_void_p()
        st = self.lib.pacuda_mlp_create(inputs, hidden, layers, max_batch, C.byref(h))
        if st != 0:
            raise RuntimeError(f"pacuda_mlp_create: status {st}")
        self.handle = h
        self.count = self.lib.pacuda_mlp_param_count(h)
        self.step = 0

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.count
        if self.lib.pacuda_mlp_set_params(self.handle, p.ctypes.data_as(F32P), self.count) != 0:
            raise RuntimeError("pacuda_mlp_set_params failed")
        self.step = 0

    def params(self):
        out = np.empty(self.count, dtype=np.float32)
        if self.lib.pacuda_mlp_get_params(self.handle, out.ctypes.data_as(F32P), self.count) != 0:
            raise RuntimeError("pacuda_mlp_get_params failed")
        return out

    def grads(self):
        out = np.empty(self.count, dtype=np.float32)
        if self.lib.pacuda_mlp_get_grads(self.handle, out.ctypes.data_as(F32P), self.count) != 0:
            raise RuntimeError("pacuda_mlp_get_grads failed")
        return out

    def forward(self, d_x, batch, d_y, train=False):
        if self.lib.pacuda_mlp_forward(self.handle, d_x.ptr, batch, d_y.ptr, 1 if train else 0) != 0:
            raise RuntimeError("pacuda_mlp_forward failed")

    def backward(self, d_y, d_t, batch, want_loss=True):
        """gradients of the mean squared error; returns the batch loss (a device synchronisation)
        when want_loss, else None (the launches stay queued)"""
        loss = C.c_float()
        if self.lib.pacuda_mlp_backward(self.handle, d_y.ptr, d_t.ptr, batch, C.byref(loss) if want_loss else None) != 0:
            raise RuntimeError("pacuda_mlp_backward failed")
        return loss.value if want_loss else None

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        if self.lib.pacuda_mlp_adam(self.handle, lr, beta1, beta2, eps, self.step) != 0:
            raise RuntimeError("pacuda_mlp_adam failed")

    def train_steps(self, d_x, d_t, batch, steps, lr, first_step, total_steps, beta1=0.9, beta2=0.999, eps=1e-8,
                    want_loss=True):
        """`steps` forward/backward/Adam steps on consecutive batches of d_x, d_t in one native
        call (the learning rate of step s: lr 0.5 (1 + cos(pi s / total_steps))); returns the last
        batch's loss when want_loss (a device synchronisation), else None"""
        loss = C.c_float()
        st = self.lib.pacuda_mlp_train_steps(self.handle, d_x.ptr, d_t.ptr, batch, steps, lr, first_step, total_steps,
                                             beta1, beta2, eps, C.byref(loss) if want_loss else None)
        if st != 0:
            raise RuntimeError(f"pacuda_mlp_train_steps: status {st}")
        self.step = first_step + steps
        return loss.value if want_loss else None

    def last_loss(self):
        """the loss of the last backward or training step (waits for the queued work)"""
        loss = C.c_float()
        if self.lib.pacuda_mlp_last_loss(self.handle, C.byref(loss)) != 0:
            raise RuntimeError("pacuda_mlp_last_loss failed")
        return loss.value

    def destroy(self):
        if self.handle:
            self.lib.pacuda_mlp_destroy(self.handle)
            self.handle = None
