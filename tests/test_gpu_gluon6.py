# SPDX-License-Identifier: 0BSD
"""GPU parity of the longest plan, g g > g g g g (6 gluons, 64 helicity outputs), through the
fused path in every precision mode (the reference's four and the two compensated mixed
policies) and at every launch shape the library and the autotuner use.

The 6-gluon program is the one that exercises the on-chip allocator under pressure (eviction,
write-back of evicted currents to the slot, gluon sources made again), so each (kernel, waves per
CU, block) combination below picks a different register tier and LDS variant. Every result is
compared bit for bit with the oracle's whole plan run (oracle/refcpu.cpp, the same precision
policy emulated on the CPU), resolved outputs and totals, on a ragged point count that makes
wavefronts take several 64-point units each at low occupancy."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
from pacuda_amd.runtime import TUNE_SHAPES
import harness
from harness import FusedPlan, DeviceArray, fused_evaluate, bits_equal
import refcpu

pytestmark = pytest.mark.gpu

POINTS = 20000 + 11  # 313 units, the last one ragged
# 1 wave per CU: 256 wavefronts, each takes one or two units (slot reuse); the rest are the
# launch-shape list of the 5-gluon tests and the autotuner's candidates
SHAPES = sorted({(1, 64), (4, 256), (6, 64), (8, 256), (12, 256), (16, 256), (16, 64), (24, 64), (32, 256),
                 *TUNE_SHAPES})


@pytest.fixture(scope="module")
def gg_gggg():
    plan = plans.gluon_plan(6)
    mom, _ = ps.generate(6, POINTS, seed=21)
    ext = ps.external_planes(mom, plan.crossing)
    lib = abi.load()
    fp = FusedPlan(lib, plan)
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    yield lib, plan, ext, fp, d_ext
    d_ext.free()
    fp.destroy()


_oracle = {}


def _want(plan, ext, kernel):
    if kernel not in _oracle:
        _oracle[kernel] = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, POINTS, kernel=kernel,
                                          threads=8)
    return _oracle[kernel]


@pytest.mark.parametrize("kernel", ["f64", "df64", "mixed", "mixed64", "compensated", "compensated64"])
def test_gluon6_every_mode_and_shape_bitwise(gg_gggg, kernel):
    lib, plan, ext, fp, d_ext = gg_gggg
    k = abi.FUSED_KERNELS[kernel]
    outputs = plan.amplitude_planes
    assert outputs == 64
    want_res, want_tot = _want(plan, ext, k)
    d_res = DeviceArray(lib, POINTS * outputs * 8)
    d_tot = DeviceArray(lib, POINTS * 8)
    try:
        for wpc, block in SHAPES:
            print("gg>gggg", kernel, "waves per CU", wpc, "block", block, flush=True)
            fused_evaluate(lib, [(fp, d_ext, POINTS, ext.shape[1], d_res, d_tot)], k, waves_per_cu=wpc, block=block)
            res = d_res.download((POINTS, outputs))
            tot = d_tot.download(POINTS)
            assert bits_equal(res, want_res), \
                f"{kernel} {wpc}x{block}: resolved differs at {np.count_nonzero(res != want_res)} values"
            assert bits_equal(tot, want_tot), f"{kernel} {wpc}x{block}: totals differ"
    finally:
        d_res.free()
        d_tot.free()


def test_gluon6_default_and_autotuned_shape_bitwise(gg_gggg):
    """the default launch (waves_per_cu = block = 0) and the shape the autotuner picks, per mode"""
    from pacuda_amd.runtime import autotune
    lib, plan, ext, fp, d_ext = gg_gggg
    d_tot = DeviceArray(lib, POINTS * 8)
    try:
        for kernel in ("f64", "df64", "mixed", "mixed64", "compensated", "compensated64"):
            k = abi.FUSED_KERNELS[kernel]
            _, want_tot = _want(plan, ext, k)
            jobs = [(fp, d_ext, POINTS, ext.shape[1], None, d_tot)]
            (wpc, block), _ = autotune(lib, jobs, k, repeats=1)
            for shape in ((0, 0), (wpc, block)):
                print("gg>gggg", kernel, "shape", shape, flush=True)
                fused_evaluate(lib, jobs, k, waves_per_cu=shape[0], block=shape[1])
                assert bits_equal(d_tot.download(POINTS), want_tot), f"{kernel} {shape}"
    finally:
        d_tot.free()
