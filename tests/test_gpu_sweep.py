# SPDX-License-Identifier: 0BSD
"""HIP-graph sweeps (pacuda_sweep_*): a captured batch of several plans, replayed three times with
the program's own eager launches and synchronisations in between; every replay equals the eager
result bit for bit, also after the momenta are rewritten in place between replays."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
from pacuda_amd.runtime import Device, FusedPlan, Sweep, evaluate

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager():
    dev = Device()
    lib = dev.lib
    jobs, exts, holds = [], [], []
    for legs, n, seed in [(5, 3000, 21), (4, 130, 22), (6, 700, 23)]:
        plan = plans.gluon_plan(legs)
        ext = ps.external_planes(ps.generate(legs, n, seed)[0], plan.crossing)
        fp, d_ext, d_tot = FusedPlan(lib, plan), dev.array(ext), dev.empty(n)
        jobs.append((fp, d_ext, n, n, None, d_tot))
        exts.append((plan, d_ext, legs, n))
        holds += [fp, d_ext, d_tot]
    sweep = Sweep(lib, jobs, abi.KERNEL_F64, waves_per_cu=2)
    evaluate(lib, jobs, abi.KERNEL_F64)
    eager = [j[5].download(j[2]) for j in jobs]
    for rep in range(3):
        for j in jobs:  # clobber the outputs, then eager work on another stream-ordered launch
            j[5].upload(np.full(j[2], np.nan))
        evaluate(lib, jobs[1:2], abi.KERNEL_MIXED)
        dev.synchronize()
        print("replay", rep, flush=True)
        sweep.run(timed=True)
        evaluate(lib, jobs[1:2], abi.KERNEL_F64)
        dev.synchronize()
        for j, want in zip(jobs, eager):
            got = j[5].download(j[2])
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), f"replay {rep}"
    # new momenta in place: the replay follows them
    plan, d_ext, legs, n = exts[0]
    ext2 = ps.external_planes(ps.generate(legs, n, 99)[0], plan.crossing)
    d_ext.upload(ext2)
    sweep.run()
    dev.synchronize()
    got = jobs[0][5].download(n)
    evaluate(lib, jobs[:1], abi.KERNEL_F64)
    want = jobs[0][5].download(n)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert not np.array_equal(got, eager[0])
    sweep.destroy()
    for h in holds:
        if hasattr(h, "destroy"):
            h.destroy()
        else:
            h.free()
