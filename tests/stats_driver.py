# SPDX-License-Identifier: 0BSD
"""Configures the HIP runner with statistics on (f64), runs the 4-gluon plan twice (KEEP +
reduce_resolved) on 100 and 28 points, and exits: the runner prints its statistics line at exit
(runner/recurrence_runner.cu:1263-1288)."""
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from pacuda_amd import abi, plans, phase_space as ps  # noqa: E402
from harness import V1  # noqa: E402

v1 = V1(abi.load())
o = abi.Options(abi.ABI, C.sizeof(abi.Options), abi.KERNEL_F64, 0, 0, 1)
assert v1.configure(C.byref(o)) == 0
plan = plans.gluon_plan(4)
for points, seed in ((100, 5), (28, 6)):
    mom, _ = ps.generate(4, points, seed)
    v1.run_plan(plan, ps.external_planes(mom, plan.crossing), points)
print("driver done", flush=True)
