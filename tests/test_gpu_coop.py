# SPDX-License-Identifier: 0BSD
"""GPU parity of cooperative launches (pacuda_launch.unit_waves): the W wavefronts of a workgroup
share each 64-point unit, running the plan's cooperative program phase by phase between
workgroup barriers with the unit's currents, forms and amplitudes in the workgroup's LDS and slot.
Every shape is compared bit for bit with the oracle's plan run (oracle/refcpu.cpp), resolved
outputs and totals, on ragged point counts: g g > g g g, g g > g g g g (whose live currents
overflow LDS into the slot), a quark-line plan, the one-quark-line p p > j j j subprocesses in
one launch with uneven points, a graph-captured sweep and the allocation-free entry point."""
import ctypes as C

import numpy as np
import pytest

from pacuda_amd import abi, plans, processes, phase_space as ps
from pacuda_amd.runtime import COOP_SHAPES, Sweep
import harness
from harness import FusedPlan, DeviceArray, fused_evaluate, bits_equal
import refcpu

pytestmark = pytest.mark.gpu

F64 = abi.FUSED_KERNELS["f64"]


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def _run_plan(lib, plan, points, seed, shapes=COOP_SHAPES, kernel="f64"):
    mom, _ = ps.generate(plan.legs, points, seed=seed)
    ext = ps.external_planes(mom, plan.crossing)
    k = abi.FUSED_KERNELS[kernel]
    want_res, want_tot = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, points, kernel=k, threads=8)
    outputs = want_res.shape[1]
    fp = FusedPlan(lib, plan)
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    d_res = DeviceArray(lib, points * outputs * 8)
    d_tot = DeviceArray(lib, points * 8)
    try:
        for unit_waves, per_cu in shapes:
            print(plan.name, kernel, points, "points: cooperative", unit_waves, "wavefronts per unit,", per_cu, "per CU",
                  flush=True)
            fused_evaluate(lib, [(fp, d_ext, points, ext.shape[1], d_res, d_tot)], k, waves_per_cu=per_cu,
                           unit_waves=unit_waves)
            res = d_res.download((points, outputs))
            tot = d_tot.download(points)
            assert bits_equal(res, want_res), \
                f"{plan.name} {kernel} {unit_waves}/{per_cu}: {np.count_nonzero(res != want_res)} values differ"
            assert bits_equal(tot, want_tot), f"{plan.name} {kernel} {unit_waves}/{per_cu}: totals differ"
    finally:
        d_ext.free()
        d_res.free()
        d_tot.free()
        fp.destroy()


def test_coop_gluon5_bitwise(lib):
    _run_plan(lib, plans.gluon_plan(5), 40000 + 37, seed=61)


def test_coop_gluon6_bitwise(lib):
    _run_plan(lib, plans.gluon_plan(6), 12000 + 5, seed=62)


def test_coop_quark_line_bitwise(lib):
    _run_plan(lib, plans.quark_line_plan(5), 9000 + 63, seed=63)


# the shapes the kernels other than f64 run
ANY_SHAPES = ((16, 16), (8, 16))


@pytest.mark.parametrize("kernel", ["df64", "mixed", "mixed64", "compensated", "compensated64"])
def test_coop_other_policies_bitwise(lib, kernel):
    """every precision policy's cooperative kernel (the sources in f64, then the policy's forms and
    constants) equals the oracle's emulation of the policy bit for bit: g g > g g g and, in the
    f32-current policies, g g > g g g g, whose currents overflow LDS into the slot"""
    _run_plan(lib, plans.gluon_plan(5), 6000 + 17, seed=65, shapes=ANY_SHAPES, kernel=kernel)
    if kernel in ("mixed", "mixed64", "compensated64"):
        _run_plan(lib, plans.gluon_plan(6), 3000 + 41, seed=66, shapes=ANY_SHAPES, kernel=kernel)


def test_coop_one_launch_of_many_plans(lib):
    """the one-quark-line p p > j j j subprocesses and g g > g g g in one cooperative launch with
    uneven point counts (a unit count of 1 included), in every f64 shape"""
    subs = [s for s in processes.subprocesses() if s.plan.name in ("5g", "qqbar+3g")]
    assert len(subs) == 29
    _many_plans(lib, subs, COOP_SHAPES, ("f64",))


def test_coop_pp_jjj_with_symjit_rows(lib):
    """every p p > j j j subprocess, the 152 four-quark ones with their symjit rows among them, in
    one cooperative launch, in the shapes and policies that run symjit rows cooperatively"""
    subs = processes.subprocesses()
    assert len(subs) == 181
    _many_plans(lib, subs, ANY_SHAPES, ("f64", "mixed64", "compensated64"))


def _many_plans(lib, subs, shapes, kernels):
    jobs, devs, fps, exts = [], [], [], []
    for k, sp in enumerate(subs):
        points = 64 * (k % 7) + 29 * k + 1
        mom, _ = ps.generate(sp.plan.legs, points, 500 + k)
        ext = ps.external_planes(mom, sp.plan.crossing)
        fp = FusedPlan(lib, sp.plan)
        d_ext = DeviceArray(lib, ext.nbytes)
        d_ext.upload(ext)
        d_tot = DeviceArray(lib, points * 8)
        jobs.append((fp, d_ext, points, ext.shape[1], None, d_tot))
        devs += [d_ext, d_tot]
        fps.append(fp)
        exts.append(ext)
    try:
        for kernel in kernels:
            k = abi.FUSED_KERNELS[kernel]
            wants = [refcpu.evaluate(sp.plan.arrays(), harness.plan_counts(sp.plan), ext, job[2], kernel=k,
                                     resolved=False)[1] for sp, job, ext in zip(subs, jobs, exts)]
            for unit_waves, per_cu in shapes:
                print(len(subs), "p p > j j j subprocesses:", kernel, "cooperative", unit_waves, per_cu, flush=True)
                fused_evaluate(lib, jobs, k, waves_per_cu=per_cu, unit_waves=unit_waves)
                for sp, (fp, _, points, _, _, d_tot), want in zip(subs, jobs, wants):
                    assert bits_equal(d_tot.download(points), want), f"{kernel} {unit_waves}/{per_cu} {sp.name}"
    finally:
        for d in devs:
            d.free()
        for fp in fps:
            fp.destroy()


def test_coop_sweep_replays_and_static_entry_point(lib):
    """a cooperative launch captured in a sweep's graph gives the eager result on every replay
    (with eager launches between), and pacuda_evaluate_static gives it too"""
    plan = plans.gluon_plan(5)
    points = 5000 + 3
    mom, _ = ps.generate(5, points, seed=64)
    ext = ps.external_planes(mom, plan.crossing)
    _, want = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, points, kernel=1, resolved=False)
    fp = FusedPlan(lib, plan)
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    d_tot = DeviceArray(lib, points * 8)
    job = [(fp, d_ext, points, ext.shape[1], None, d_tot)]
    sw = Sweep(lib, job, F64, waves_per_cu=16, unit_waves=16)
    try:
        for r in range(3):
            print("cooperative sweep replay", r, flush=True)
            lib.pacuda_copy_to_device(d_tot.ptr, np.zeros(points).ctypes.data, points * 8)
            sw.run()
            assert lib.pacuda_synchronize() == 0
            assert bits_equal(d_tot.download(points), want), f"replay {r}"
            fused_evaluate(lib, job, F64, waves_per_cu=8, unit_waves=8)
            assert bits_equal(d_tot.download(points), want)
        print("cooperative static evaluation", flush=True)
        lib.pacuda_copy_to_device(d_tot.ptr, np.zeros(points).ctypes.data, points * 8)
        items = (abi.Item * 1)()
        items[0].plan, items[0].external, items[0].points = fp.handle, d_ext.ptr, points
        items[0].ext_stride, items[0].resolved, items[0].total = ext.shape[1], None, d_tot.ptr
        l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), F64, 16, 0, 16, None, None)
        assert lib.pacuda_evaluate_static(items, 1, C.byref(l)) == 0
        assert lib.pacuda_synchronize() == 0
        assert bits_equal(d_tot.download(points), want)
    finally:
        sw.destroy()
        d_ext.free()
        d_tot.free()
        fp.destroy()


def test_coop_refusals(lib):
    """host-side refusals (nothing is launched): a plan with symjit rows in a shape without a
    symjit instance, a shape only the f64 kernel runs asked of another, a shape that does not
    exist"""
    sub = processes.four_quark_subprocesses()[0]
    fp = FusedPlan(lib, sub.plan)  # (symjit rows: shapes 0 and 2 only)
    good = FusedPlan(lib, plans.gluon_plan(5))
    d = DeviceArray(lib, 5 * 4 * 64 * 8)
    d_tot = DeviceArray(lib, 64 * 8)
    try:
        def status(p, kernel, per_cu, unit_waves):
            items = (abi.Item * 1)()
            items[0].plan, items[0].external, items[0].points = p.handle, d.ptr, 64
            items[0].ext_stride, items[0].resolved, items[0].total = 64, None, d_tot.ptr
            l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), kernel, per_cu, 0, unit_waves, None, None)
            return lib.pacuda_evaluate(items, 1, C.byref(l))
        assert status(fp, F64, 8, 8) == 2
        assert status(good, abi.FUSED_KERNELS["mixed"], 8, 8) == 2  # (a shape only f64 runs)
        assert status(good, F64, 12, 16) == 2
        assert status(good, F64, 16, 2) == 2
    finally:
        d.free()
        d_tot.free()
        fp.destroy()
        good.destroy()
