# SPDX-License-Identifier: 0BSD
"""The lock contract of the v1 ABI (runner/recurrence_runner.cu:1725-1755), as scenarios run
against any library with the v1 entry points (the HIP runner, the oracle's emulation).

- pairs: N threads each issue M run(KEEP) + reduce_resolved pairs on their own points; every
  result equals the single-thread result for the same points bit for bit (a kept run holds the
  lock until its thread's reduction, so no other call overwrites the kept amplitudes in between);
- blocking: while one thread holds a kept run, another thread's run waits until the holder's
  reduction, and capabilities / symjit_kernel (which take no lock) return at once;
- rerun: a thread that runs again without reducing (an error between run and reduce) keeps the
  lock and its second kept run is reduced as usual."""
import threading
import time

import numpy as np

from pacuda_amd import plans, phase_space as ps


def _jobs(threads, per_thread):
    plan = plans.gluon_plan(4)
    jobs = []
    for t in range(threads):
        for m in range(per_thread):
            points = 40 + 17 * t + 5 * m
            mom, _ = ps.generate(4, points, seed=1000 + 37 * t + m)
            jobs.append((t, m, ps.external_planes(mom, plan.crossing)))
    return plan, jobs


def pairs(v1, threads=4, per_thread=5, key=0):
    """(single-thread results, concurrent results) keyed by (thread, repetition); `key` is the
    calls' words_key"""
    plan, jobs = _jobs(threads, per_thread)
    single = {(t, m): v1.run_plan(plan, ext, ext.shape[1], words_key=key) for t, m, ext in jobs}
    concurrent, errors = {}, []
    start = threading.Barrier(threads)

    def work(t):
        try:
            start.wait()
            for tt, m, ext in jobs:
                if tt == t:
                    concurrent[(t, m)] = v1.run_plan(plan, ext, ext.shape[1], words_key=key)
        except Exception as e:  # noqa: BLE001 (reported by the caller)
            errors.append(f"thread {t}: {e}")

    ts = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for x in ts:
        x.start()
    for x in ts:
        x.join(timeout=120)
    assert not any(x.is_alive() for x in ts), "a thread did not finish"
    assert not errors, errors
    return single, concurrent


def blocking(v1, hold_seconds=0.4):
    """timestamps: (holder's reduce started, waiter's run returned, capability calls' longest
    time) for one thread holding a kept run while another runs"""
    import ctypes as C
    from pacuda_amd import abi
    from pacuda_amd.abi import ptr, u32p
    plan = plans.gluon_plan(4)
    mom, _ = ps.generate(4, 64, seed=77)
    ext = ps.external_planes(mom, plan.crossing)
    a = plan.arrays()
    held = threading.Event()
    times = {}
    errors = []

    def holder():
        try:
            v1.run_plan_keep_only(plan, ext, 64)  # returns with the lock held
            held.set()
            time.sleep(hold_seconds)
            outputs = len(a["offsets"]) - 1
            out = np.zeros(64 * outputs)
            r = abi.ReduceResolved(abi.ABI, C.sizeof(abi.ReduceResolved), ptr(a["planes"], u32p), ptr(a["weights"]),
                                   ptr(a["offsets"], u32p), ptr(out), outputs, 64, 1, outputs, abi.OUTPUT_ZERO)
            times["reduce"] = time.perf_counter()
            assert v1.reduce_resolved(C.byref(r)) == 0
        except Exception as e:  # noqa: BLE001
            errors.append(f"holder: {e}")

    def waiter():
        try:
            held.wait(30)
            t = time.perf_counter()
            v1.capabilities()
            v1.symjit_kernel(b"0" * 64)
            times["unlocked"] = time.perf_counter() - t
            v1.run_plan(plan, ext, 64)
            times["run"] = time.perf_counter()
        except Exception as e:  # noqa: BLE001
            errors.append(f"waiter: {e}")

    ts = [threading.Thread(target=holder), threading.Thread(target=waiter)]
    for x in ts:
        x.start()
    for x in ts:
        x.join(timeout=120)
    assert not any(x.is_alive() for x in ts), "a thread did not finish"
    assert not errors, errors
    return times["reduce"], times["run"], times["unlocked"]


def rerun(v1):
    """a kept run, another kept run by the same thread without a reduction, then the reduction:
    the reduction sees the second run's points"""
    plan = plans.gluon_plan(4)
    mom, _ = ps.generate(4, 50, seed=78)
    ext = ps.external_planes(mom, plan.crossing)
    mom2, _ = ps.generate(4, 30, seed=79)
    ext2 = ps.external_planes(mom2, plan.crossing)
    v1.run_plan_keep_only(plan, ext, 50)
    return v1.run_plan(plan, ext2, 30), v1.run_plan(plan, ext2, 30)
