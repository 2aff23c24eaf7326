# SPDX-License-Identifier: 0BSD
"""Cooperative programs on the CPU (pacuda_plan_coop_program; no GPU).

A cooperative program spreads one 64-point unit's rows over the wavefronts of a workgroup in
phases separated by barriers, with the unit's currents, forms and amplitudes placed in the
workgroup's LDS or slot. Its serial order (phase by phase, wavefront by wavefront) run through the
oracle's fused-program interpreter (oracle/refcpu.cpp, LDS planes mapped after the slot's) must
equal the oracle's run of the plan bit for bit: that pins the dependences the phases honour and
the placement (a clobbered current changes values). The streams' structure (the same barriers in
every wavefront, one fetch of the next unit, the same rows as the serial order) is what the kernel
relies on to neither deadlock nor skip work."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, processes, phase_space as ps
from pacuda_amd.runtime import coop_program, plan_bindings, COOP_SHAPES
import harness
import refcpu

K_BARRIER, K_FETCH = 8, 9
LDS_CAP = {0: 159, 1: 159, 2: 79, 3: 39}


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def _plans():
    subs = processes.subprocesses()
    four = processes.four_quark_subprocesses()
    return [("gg>ggg", plans.gluon_plan(5)), ("gg>gggg", plans.gluon_plan(6)), ("qline5", plans.quark_line_plan(5)),
            (subs[7].name, subs[7].plan), (subs[25].name, subs[25].plan), (four[0].name, four[0].plan),
            (four[57].name, four[57].plan)]


# (class, shape, policy): the 16-byte currents' programs in every shape under f64 and df64, the
# f32 currents' in the shapes the other kernels run under mixed and compensated64
CASES = [(False, k, "f64") for k in range(len(COOP_SHAPES))] + [(False, 2, "df64"), (True, 0, "mixed"),
                                                                (True, 2, "mixed"), (True, 2, "compensated64")]


@pytest.mark.parametrize("f32,shape,policy", CASES)
@pytest.mark.parametrize("which", range(7))
def test_serial_order_equals_the_plan_run(lib, f32, shape, policy, which):
    name, plan = _plans()[which]
    got = coop_program(lib, plan, shape, serial=True, f32=f32)
    assert got is not None, name
    ser, info, _ = got
    assert info["waves"] == COOP_SHAPES[shape][0]
    assert info["lds_planes"] <= LDS_CAP[shape] * (2 if f32 else 1)
    n = 4 * 64 + 13
    mom, _ = ps.generate(plan.legs, n, seed=40 + which)
    ext = ps.external_planes(mom, plan.crossing)
    counts = harness.plan_counts(plan)
    k = abi.FUSED_KERNELS[policy]
    # the first phase is the sources (in f64, before the policy's forms and constants), which
    # the oracle runs as its sources phase
    kinds = ser[:, 0] & 0xff
    lead = int(np.argmax(kinds != 6))
    assert lead > 0 and not np.any(kinds[lead:] == 6)
    want_res, want_tot = refcpu.evaluate(plan.arrays(), counts, ext, n, kernel=k)
    res, tot = refcpu.evaluate_program(ser, lead, info["slot_planes"], info["lds_planes"], plan.arrays(), counts, ext, n,
                                       kernel=k, numerator=plan.numerator, denominator=plan.denominator,
                                       bindings=plan_bindings(lib, plan))
    assert np.array_equal(res.view(np.uint64), want_res.view(np.uint64)), name
    assert np.array_equal(tot.view(np.uint64), want_tot.view(np.uint64)), name


@pytest.mark.parametrize("f32,shape", [(False, k) for k in range(len(COOP_SHAPES))] + [(True, 0), (True, 2)])
def test_streams_meet_the_same_barriers(lib, f32, shape):
    for name, plan in _plans():
        dev, info, starts = coop_program(lib, plan, shape, f32=f32)
        ser, _, _ = coop_program(lib, plan, shape, serial=True, f32=f32)
        W, P = info["waves"], info["phases"]
        assert len(starts) == W + 1 and starts[0] == 0 and starts[-1] == len(dev)
        assert all(a <= b for a, b in zip(starts, starts[1:]))
        kinds = dev[:, 0] & 0xff
        for w in range(W):
            assert np.count_nonzero(kinds[starts[w]:starts[w + 1]] == K_BARRIER) == P - 1, (name, w)
        # one wavefront takes the next unit off the queue, once
        assert np.count_nonzero(kinds == K_FETCH) == 1, name
        # every row of the plan's program once (the serial order holds them in phase order)
        assert len(dev) - W * (P - 1) - 1 == len(ser), name
        rows = (kinds != K_BARRIER) & (kinds != K_FETCH)
        assert sorted(dev[rows, 0].tolist()) == sorted(ser[:, 0].tolist()), name


def test_more_lds_leaves_fewer_slot_accesses(lib):
    plan = plans.gluon_plan(5)
    small = coop_program(lib, plan, 2)[1]
    large = coop_program(lib, plan, 0)[1]
    assert large["slot_accesses"] < small["slot_accesses"]
    assert large["slot_accesses"] + large["lds_accesses"] == small["slot_accesses"] + small["lds_accesses"]
    # g g > g g g: most current accesses stay in the unit's LDS at 159 planes
    assert large["lds_accesses"] > 2 * large["slot_accesses"]


def test_symjit_rows_in_cooperative_programs(lib):
    """a four-quark plan's symjit rows (fermion-pair currents, kernel 1 of the built-in table) are
    modelled from their bindings: each reads both two-component parents and creates a
    four-component destination, which the program then propagates in place"""
    sub = processes.four_quark_subprocesses()[0]
    dev, info, starts = coop_program(lib, sub.plan, 0)
    kinds = dev[:, 0] & 0xff
    assert np.count_nonzero(kinds == 7) == 16
    assert info["slot_accesses"] < info["lds_accesses"]
    bind = plan_bindings(lib, sub.plan)
    for w in dev[kinds == 7]:
        at = int(w[6])
        inputs, pairs = int(bind[at + 1]), int(bind[at + 2])
        assert list(bind[at + 4 + inputs:at + 4 + inputs + pairs]) == [0, 1, 2, 3]


def test_f32_currents_fit_twice_the_planes(lib):
    """the f32 currents' programs place planes of 512 bytes: twice as many in the same LDS, so
    fewer slot accesses"""
    plan = plans.gluon_plan(5)
    wide, narrow = coop_program(lib, plan, 2)[1], coop_program(lib, plan, 2, f32=True)[1]
    assert narrow["lds_planes"] > wide["lds_planes"] and narrow["slot_accesses"] < wide["slot_accesses"]


def test_the_oracle_sees_a_broken_order(lib):
    """the parity check has teeth: the serial rows with two dependent phases swapped (the
    closures before the currents they read) give other values"""
    plan = plans.gluon_plan(5)
    ser, info, _ = coop_program(lib, plan, 0, serial=True)
    kinds = ser[:, 0] & 0xff
    closures = ser[kinds == 5]
    rest = ser[kinds != 5]
    broken = np.concatenate([closures, rest])
    n = 64
    mom, _ = ps.generate(5, n, seed=3)
    ext = ps.external_planes(mom, plan.crossing)
    counts = harness.plan_counts(plan)
    _, want = refcpu.evaluate(plan.arrays(), counts, ext, n, kernel=1)
    _, got = refcpu.evaluate_program(broken, 0, info["slot_planes"], info["lds_planes"], plan.arrays(), counts, ext, n,
                                     kernel=1)
    assert not np.array_equal(got.view(np.uint64), want.view(np.uint64))
