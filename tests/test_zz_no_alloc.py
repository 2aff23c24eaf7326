# SPDX-License-Identifier: 0BSD
"""pacuda_evaluate_static (the fused evaluation with its counter, item table and slots in the
library's fixed device array instead of allocated buffers) against pacuda_evaluate through the C
ABI: bit for bit on two calls in a row, the second with a smaller input, in the f64 and mixed
kernels; and its limits are refused with status 2 before anything runs."""
import ctypes as C

import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
from harness import FusedPlan, DeviceArray, fused_evaluate, bits_equal

pytestmark = pytest.mark.gpu


def _items(jobs):
    items = (abi.Item * len(jobs))()
    for k, (fp, ext, points, stride, res, tot) in enumerate(jobs):
        items[k].plan = fp.handle
        items[k].external = ext.ptr
        items[k].points = points
        items[k].ext_stride = stride
        items[k].resolved = res.ptr if res is not None else None
        items[k].total = tot.ptr if tot is not None else None
    return items


def _static(lib, jobs, kernel, waves_per_cu=0, block=0, kernel_ms=None):
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), kernel, waves_per_cu, block, 0, None, kernel_ms)
    return lib.pacuda_evaluate_static(_items(jobs), len(jobs), C.byref(l))


@pytest.fixture(scope="module")
def setup():
    lib = abi.load()
    data = []
    for n, points, seed in [(5, 5003, 41), (6, 3007, 42)]:
        plan = plans.gluon_plan(n)
        mom, _ = ps.generate(n, points, seed)
        ext = ps.external_planes(mom, plan.crossing)
        fp = FusedPlan(lib, plan)
        d_ext = DeviceArray(lib, ext.nbytes)
        d_ext.upload(ext)
        data.append((plan, fp, d_ext, points, ext.shape[1]))
    yield lib, data
    for _, fp, d_ext, _, _ in data:
        d_ext.free()
        fp.destroy()


def _run_both(lib, data, sizes, kernel, shape=(0, 0)):
    """[(item index, points)]: pacuda_evaluate and pacuda_evaluate_static on the same items;
    returns the downloads of both"""
    outs = []
    for use_static in (False, True):
        jobs, bufs = [], []
        for k, points in sizes:
            plan, fp, d_ext, _, stride = data[k]
            outputs = plan.amplitude_planes
            res = DeviceArray(lib, points * outputs * 8)
            tot = DeviceArray(lib, points * 8)
            res.upload(np.full(points * outputs, np.nan))
            tot.upload(np.full(points, np.nan))
            jobs.append((fp, d_ext, points, stride, res, tot))
            bufs.append((res, tot, points, outputs))
        if use_static:
            print("static", kernel, sizes, shape, flush=True)
            assert _static(lib, jobs, kernel, *shape) == 0
            assert lib.pacuda_synchronize() == 0
        else:
            fused_evaluate(lib, jobs, kernel, waves_per_cu=shape[0], block=shape[1])
        got = [(r.download((p, o)), t.download(p)) for r, t, p, o in bufs]
        for r, t, _, _ in bufs:
            r.free()
            t.free()
        outs.append(got)
    return outs


@pytest.mark.parametrize("kernel", ["f64", "mixed"])
def test_static_equals_allocating_entry_point(setup, kernel):
    lib, data = setup
    k = abi.KERNELS[kernel]
    # two calls in a row, the second smaller (one item, fewer points, another plan first)
    for sizes in ([(0, 5003), (1, 3007)], [(1, 700)]):
        want, got = _run_both(lib, data, sizes, k)
        for (wr, wt), (gr, gt), (item, points) in zip(want, got, sizes):
            assert not np.isnan(gt).any(), f"{kernel} {sizes}: outputs not written"
            assert bits_equal(gr, wr), f"{kernel} item {item} {points}: resolved differs"
            assert bits_equal(gt, wt), f"{kernel} item {item} {points}: totals differ"


def test_static_at_low_and_high_occupancy(setup):
    """the slots of a 16-per-CU launch of g g > g g g g do not all fit: the call runs fewer
    resident wavefronts and still matches"""
    lib, data = setup
    for shape in ((16, 256), (2, 64)):
        want, got = _run_both(lib, data, [(1, 3007), (0, 5003)], abi.KERNEL_F64, shape)
        for (wr, wt), (gr, gt) in zip(want, got):
            assert bits_equal(gr, wr) and bits_equal(gt, wt), shape


def test_static_limits_are_refused(setup):
    lib, data = setup
    plan, fp, d_ext, points, stride = data[0]
    tot = DeviceArray(lib, points * 8)
    try:
        job = (fp, d_ext, points, stride, None, tot)
        ms = C.c_float(0)
        assert _static(lib, [job], abi.KERNEL_F64, kernel_ms=C.pointer(ms)) == abi.INVALID
        assert _static(lib, [job] * 4097, abi.KERNEL_F64) == abi.INVALID
        assert _static(lib, [job], abi.KERNEL_F64, block=96) == abi.INVALID
        assert _static(lib, [], abi.KERNEL_F64) == 0
    finally:
        tot.free()
