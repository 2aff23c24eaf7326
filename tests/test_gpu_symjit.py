# SPDX-License-Identifier: 0BSD
"""Symjit rows (op 13) on gfx950 through the C ABI, bit for bit against the oracle.

- the probe plan of tests/symjit_cases.py (every binding code kind, accumulate, direct, the
  fermion-pair kernels) through pacman_recurrence_v1_run: f64 equals the numpy restatement of
  runner/recurrence_runner.cu:631-684 and the oracle; df64, mixed and mixed64 (one configuration
  per process) equal the oracle's emulation of the same kernel;
- the same plan through the fused path in every kernel equals the oracle's whole plan run."""
import os
import subprocess
import sys

import numpy as np
import pytest

from pacuda_amd import abi
import harness
from harness import V1, FusedPlan, DeviceArray, fused_evaluate, bits_equal
import refcpu
import symjit_cases
from test_symjit_cpu import probe_inputs, sources_of

pytestmark = pytest.mark.gpu


def _oracle_currents(plan, ext, kernel):
    lib = refcpu.load()
    lib.refcpu_reset()
    v1 = V1(lib, prefix="refcpu_v1_")
    assert v1.configure_kernel(kernel) == 0
    cur = symjit_cases.run_currents(v1, plan, ext, ext.shape[1])
    lib.refcpu_reset()
    return cur


def test_v1_f64_symjit_rows_bitwise():
    plan, ext = probe_inputs(points=1000 + 7)
    v1 = V1(abi.load())
    assert v1.configure_kernel(abi.KERNEL_F64) == 0
    print("symjit v1 f64", flush=True)
    got = symjit_cases.run_currents(v1, plan, ext, ext.shape[1])
    want = _oracle_currents(plan, ext, abi.KERNEL_F64)
    assert bits_equal(got.real, want.real) and bits_equal(got.imag, want.imag)
    forms = refcpu.fill_forms(plan.arrays(), ext, ext.shape[1])
    for name, value in symjit_cases.expected(sources_of(plan, ext), forms).items():
        base = plan.layout[name]
        g = got[base:base + value.shape[0]]
        assert np.array_equal(g.view(np.uint64), value.view(np.uint64)), name


def test_v1_other_kernels_symjit_rows_bitwise(tmp_path):
    plan, ext = probe_inputs(points=513)
    np.save(tmp_path / "ext.npy", ext)
    for kernel in ("df64", "mixed", "mixed64"):
        print("symjit v1 driver", kernel, flush=True)
        out = tmp_path / f"{kernel}.npy"
        p = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "symjit_driver.py"), kernel,
                            str(tmp_path / "ext.npy"), str(out)], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        got = np.load(out)
        want = _oracle_currents(plan, ext, abi.KERNELS[kernel])
        assert bits_equal(got.real, want.real) and bits_equal(got.imag, want.imag), kernel


@pytest.mark.parametrize("kernel", ["f64", "df64", "mixed", "mixed64"])
def test_fused_symjit_rows_bitwise(kernel):
    lib = abi.load()
    plan, ext = probe_inputs(points=3000 + 5)
    points = ext.shape[1]
    k = abi.KERNELS[kernel]
    fp = FusedPlan(lib, plan)
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    outputs = len(plan.outputs)
    d_res = DeviceArray(lib, points * outputs * 8)
    d_tot = DeviceArray(lib, points * 8)
    try:
        for shape in ((0, 0), (4, 256), (16, 64)):
            print("symjit fused", kernel, shape, flush=True)
            fused_evaluate(lib, [(fp, d_ext, points, points, d_res, d_tot)], k, waves_per_cu=shape[0], block=shape[1])
            want_res, want_tot = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, points, kernel=k)
            assert bits_equal(d_res.download((points, outputs)), want_res), (kernel, shape)
            assert bits_equal(d_tot.download(points), want_tot), (kernel, shape)
    finally:
        for d in (d_ext, d_res, d_tot):
            d.free()
        fp.destroy()
