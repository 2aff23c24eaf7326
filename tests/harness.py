# SPDX-License-Identifier: 0BSD
"""Test harness: drive a runner library (the HIP one, or the oracle's v1 emulation) through the C
ABI the way umami/umami_pacuda.c:644-747 does, and the fused entry point."""
import ctypes as C
import os
import subprocess

import numpy as np

from pacuda_amd import abi
from pacuda_amd.abi import ptr, u32p, f64p, u64p


def plan_counts(plan):
    return dict(current_planes=plan.current_planes, amplitude_planes=plan.amplitude_planes,
                form_count=len(plan.forms), external_count=plan.legs)


class V1:
    """A library with the v1 ABI (prefix pacman_recurrence_v1_ or refcpu_v1_)."""

    def __init__(self, lib, prefix="pacman_recurrence_v1_"):
        self.lib = lib
        self.prefix = prefix
        for name, res, args in abi.V1_SYMBOLS:
            short = name[len("pacman_recurrence_v1_"):]
            f = getattr(lib, prefix + short)
            f.restype = res
            f.argtypes = args
            setattr(self, short, f)

    def configure_kernel(self, kernel, block=0, no_sort=0):
        o = abi.Options(abi.ABI, C.sizeof(abi.Options), kernel, block, no_sort, 0)
        return self.configure(C.byref(o))

    def run_plan_keep_only(self, plan, ext, points, stride=None):
        """the v1 run of run_plan with KEEP and no reduction: returns with the amplitudes kept
        (and the runner's lock held by this thread)"""
        self.run_plan(plan, ext, points, stride, keep=True, reduce=False)

    def run_plan(self, plan, ext, points, stride=None, keep=True, reduce=True, words_key=0):
        """v1 run of a generated plan with the forms on the device, then reduce_resolved from
        zero: [points][outputs] (each output before the group factor). A nonzero words_key
        keys the call (the HIP library runs such calls on its fused kernels)."""
        a = plan.arrays()
        stride = stride or points
        n_amp = plan.amplitude_planes
        keepalive = []
        cre = np.zeros(plan.current_planes * stride)
        cim = np.zeros(plan.current_planes * stride)
        are = np.zeros(n_amp * stride)
        aim = np.zeros(n_amp * stride)
        amp_runs = np.array([0, n_amp], dtype=np.uint32)
        ext = np.ascontiguousarray(ext[:, :points])
        keepalive += [cre, cim, are, aim, amp_runs, ext, a]
        c = abi.Call()
        c.abi_version = abi.ABI
        c.struct_size = C.sizeof(abi.Call)
        c.words = ptr(a["words"], u64p)
        c.words_key = words_key
        c.word_count = len(a["words"])
        for k in range(8):
            c.width[k] = int(a["width"][k])
        c.current_re, c.current_im = ptr(cre), ptr(cim)
        c.amplitude_re, c.amplitude_im = ptr(are), ptr(aim)
        c.scale_re, c.scale_im = ptr(a["scale_re"]), ptr(a["scale_im"])
        c.factor_re, c.factor_im = ptr(a["factor_re"]), ptr(a["factor_im"])
        c.download_amplitude = ptr(amp_runs, u32p)
        c.current_planes = plan.current_planes
        c.amplitude_planes = n_amp
        c.momentum_planes = 4 * len(plan.forms)
        c.scale_count = len(a["scale_re"])
        c.factor_count = len(a["factor_re"])
        c.download_amplitude_runs = 1
        c.stride = stride
        c.points = points
        c.zero_amplitude = ptr(amp_runs, u32p)
        c.zero_amplitude_runs = 1
        c.flags = (abi.KEEP_AMPLITUDES if keep else 0) | abi.FORMS_ON_DEVICE
        c.external_momenta = ptr(ext)
        c.external_count = plan.legs
        c.term_count = len(a["term_slots"])
        c.form_terms = ptr(a["form_terms"], u32p)
        c.term_slots = ptr(a["term_slots"], u32p)
        c.term_coefficients = ptr(a["term_coefficients"])
        st = self.run(C.byref(c))
        if st != 0:
            raise RuntimeError(f"run: status {st}")
        if not keep:
            return are.reshape(n_amp, stride)[:, :points], aim.reshape(n_amp, stride)[:, :points]
        if not reduce:
            return None
        outputs = len(a["offsets"]) - 1
        out = np.full(points * outputs, np.nan)
        r = abi.ReduceResolved(abi.ABI, C.sizeof(abi.ReduceResolved), ptr(a["planes"], u32p), ptr(a["weights"]),
                               ptr(a["offsets"], u32p), ptr(out), outputs, points, 1, outputs, abi.OUTPUT_ZERO)
        st = self.reduce_resolved(C.byref(r))
        if st != 0:
            raise RuntimeError(f"reduce_resolved: status {st}")
        return out.reshape(points, outputs)


def helicity_sum(resolved, numerator=1.0, denominator=1.0):
    """umami's combination: v = v * num / den when they differ, summed from 0 in output order"""
    t = np.zeros(resolved.shape[0])
    for o in range(resolved.shape[1]):
        v = resolved[:, o]
        if numerator != denominator:
            v = (v * numerator) / denominator
        t = t + v
    return t


from pacuda_amd.runtime import DeviceArray, FusedPlan  # noqa: E402,F401
from pacuda_amd.runtime import evaluate as fused_evaluate  # noqa: E402,F401


def rectest(runner_library, *args, timeout=120):
    """runs the reference's runner/rectest.c (built into oracle/_ref) against a runner library;
    (exit code, stdout)"""
    from refcpu import RECTEST
    if not os.path.exists(RECTEST):
        raise RuntimeError(f"{RECTEST} is missing: build it with `make -C oracle ref` where the reference is")
    p = subprocess.run([RECTEST, runner_library, *args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
