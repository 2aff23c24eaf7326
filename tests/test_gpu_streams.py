# SPDX-License-Identifier: 0BSD
"""pacuda_evaluate on caller streams: two evaluations issued back to back on two different
streams, without a synchronisation in between, share the library's item table, work counter and
slots; the second call must wait for the first launch before it rewrites them. Both results
equal the oracle's bit for bit."""
import ctypes as C

import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
import harness
from harness import FusedPlan, DeviceArray, bits_equal
import refcpu

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return hip


def _launch(lib, jobs, kernel, stream):
    items = (abi.Item * len(jobs))()
    for k, (fp, ext, points, stride, tot) in enumerate(jobs):
        items[k].plan = fp.handle
        items[k].external = ext.ptr
        items[k].points = points
        items[k].ext_stride = stride
        items[k].resolved = None
        items[k].total = tot.ptr
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), kernel, 0, 0, 0, stream, None)
    return lib.pacuda_evaluate(items, len(jobs), C.byref(l))


def test_back_to_back_evaluations_on_two_streams():
    lib = abi.load()
    hip = _hip()
    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        assert hip.hipStreamCreate(C.byref(s)) == 0
    cases = []
    try:
        # a long 6-gluon batch first, then a short 5-gluon one: the second call's table would
        # replace the first's while the first launch still reads it
        for n, points, seed in [(6, 60000, 51), (5, 3001, 52)]:
            plan = plans.gluon_plan(n)
            mom, _ = ps.generate(n, points, seed)
            ext = ps.external_planes(mom, plan.crossing)
            fp = FusedPlan(lib, plan)
            d_ext = DeviceArray(lib, ext.nbytes)
            d_ext.upload(ext)
            tot = DeviceArray(lib, points * 8)
            tot.upload(np.full(points, np.nan))
            cases.append((plan, ext, fp, d_ext, tot, points))
        for rep in range(3):
            print("two streams, repetition", rep, flush=True)
            for (plan, ext, fp, d_ext, tot, points), s in zip(cases, streams):
                assert _launch(lib, [(fp, d_ext, points, ext.shape[1], tot)], abi.KERNEL_F64, s) == 0
            for s in streams:
                assert hip.hipStreamSynchronize(s) == 0
            for plan, ext, fp, d_ext, tot, points in cases:
                _, want = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, points, kernel=1,
                                          resolved=False, threads=8)
                assert bits_equal(tot.download(points), want), f"{plan.name} repetition {rep}"
                tot.upload(np.full(points, np.nan))
    finally:
        for plan, ext, fp, d_ext, tot, points in cases:
            d_ext.free()
            tot.free()
            fp.destroy()
        for s in streams:
            if s.value:
                hip.hipStreamDestroy(s)
