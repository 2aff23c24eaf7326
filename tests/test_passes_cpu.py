# SPDX-License-Identifier: 0BSD
"""The fused path's program passes, checked on the CPU (no GPU): the library compiles a plan
host-side (pacuda_plan_program) into its programs — arena compaction only, with forwarding
through the accumulators, and with each LDS-residency variant, alone and with the register tier
(the ones the kernels run) — and the oracle runs each instruction stream
(refcpu_evaluate_program, on-chip slots mapped to extra planes) beside its own plan run from the
words. Every value must be equal bit for bit, in
the f64, df64 and mixed policies (the f64 kernel's own programs, which run source rows again
instead of loading sources, in f64), for 4–7 gluon plans, quark-line plans and every p p > j j j
subprocess plan; and the passes must have done their work (fewer planes, forwarded rows, LDS
slots)."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, processes, phase_space as ps
from pacuda_amd.runtime import plan_program
import harness
import refcpu

LDS_PLANE = 1 << 24
REG_PLANES = 16
FORM_REGS = 6
F_HOLD, F_FROMACC, F_ZERO = 8, 16, 32


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def _plans():
    out = [(f"{n} gluons", plans.gluon_plan(n)) for n in (4, 5, 6, 7)]
    out += [(f"quark line, {n} legs", plans.quark_line_plan(n)) for n in (4, 5, 6)]
    # (plans with symjit rows are run as built: no pass rewrites them, see below)
    out += [(sp.name, sp.plan) for sp in processes.subprocesses(four_quark=False)]
    return out


def _check(lib, name, plan, points, seed, kernels):
    mom, _ = ps.generate(plan.legs, points, seed)
    ext = ps.external_planes(mom, plan.crossing)
    counts = harness.plan_counts(plan)
    arrays = plan.arrays()
    for kernel in kernels:
        want_res, want_tot = refcpu.evaluate(arrays, counts, ext, points, kernel=kernel)
        # (12-21: the f64 kernel's programs, which run source rows again)
        for variant in range(22) if kernel == abi.KERNEL_F64 else range(12):
            instr, sources, planes, lds = plan_program(lib, plan, variant)
            if variant == 0:
                assert not np.any((instr[:, 0] >> 16) & (F_HOLD | F_FROMACC)), name
            if variant < 2:
                assert lds == 0
            res, tot = refcpu.evaluate_program(instr, sources, planes, lds, arrays, counts, ext, points,
                                               kernel=kernel)
            assert harness.bits_equal(res, want_res), f"{name}: variant {variant}, kernel {kernel}"
            assert harness.bits_equal(tot, want_tot), f"{name}: variant {variant}, kernel {kernel}"


@pytest.mark.parametrize("name,plan", _plans()[:7], ids=[n for n, _ in _plans()[:7]])
def test_passes_keep_every_value(lib, name, plan):
    kernels = (abi.KERNEL_F64, abi.KERNEL_MIXED) + ((abi.KERNEL_DF64,) if plan.legs < 7 else ())
    _check(lib, name, plan, 96 if plan.legs < 7 else 40, 21, kernels)


def test_passes_keep_every_value_pp_jjj(lib):
    for k, (name, plan) in enumerate(_plans()[7:]):
        _check(lib, name, plan, 40, 40 + k, (abi.KERNEL_F64,))


def test_symjit_plans_run_as_built(lib):
    """a plan with symjit rows (four quarks and a gluon) gets its program as built in every
    variant: no compaction, forwarding, on-chip placement or zero masks (their accesses are the
    components the bindings name, which the passes do not model)"""
    plan = plans.four_quark_plan()
    plain, sources, planes, _ = plan_program(lib, plan, 0)
    assert planes == plan.current_planes
    for variant in range(1, 2 + 4 * 5):
        prog, src, _, _ = plan_program(lib, plan, variant)
        assert src == sources and np.array_equal(prog, plain), variant
    assert not np.any(plain[:, 0] >> 16 & (F_HOLD | F_FROMACC | F_ZERO))


def test_passes_do_their_work(lib):
    plan = plans.gluon_plan(5)
    plain, _, planes, _ = plan_program(lib, plan, 0)
    assert planes < plan.current_planes  # compaction: 384 -> 278
    fwd, _, _, _ = plan_program(lib, plan, 1)
    assert len(fwd) == len(plain)
    moved = np.count_nonzero((fwd[:, 0] >> 16) & F_FROMACC)
    assert moved == np.count_nonzero((fwd[:, 0] >> 16) & F_HOLD) > 0
    # structural zeros: the gluon sources' time components and the tensor components built from
    # them are masked (bits 22-31 of a contribution's first word), and the oracle run above,
    # which reads +0 for every masked component and skips masked stores, kept every bit
    contrib = fwd[(fwd[:, 0] & 0xff) == 2]
    assert np.count_nonzero(contrib[:, 0] >> 22) > 0
    sources = fwd[(fwd[:, 0] & 0xff) == 6]
    assert np.all(sources[:, 6] & 1)  # component 0 of every massless vector source
    # the first closure of each amplitude plane starts from zero (no load of the zeroed plane)
    closures = fwd[(fwd[:, 0] & 0xff) == 5]
    zero = closures[((closures[:, 0] >> 16) & F_ZERO) != 0]
    assert sorted(zero[:, 3]) == sorted(set(closures[:, 3]))
    for variant, cap in zip((2, 3, 4, 5, 6), (10, 13, 20, 26, 40)):
        prog, _, _, lds = plan_program(lib, plan, variant)
        assert 0 < lds <= cap
        assert np.any(prog[:, 1:5] >= LDS_PLANE)
        # the register-tier variant: LDS slots, plus whole four-plane currents after them
        reg, _, _, onchip = plan_program(lib, plan, variant + 5)
        assert cap < onchip <= cap + REG_PLANES
        tier = (reg[:, 1:5] >= LDS_PLANE + cap) & (reg[:, 1:5] < LDS_PLANE + cap + REG_PLANES)
        assert np.any(tier) and np.all((reg[:, 1:5][tier] - LDS_PLANE - cap) % 4 == 0)
        # ... and the most-read forms marked as register-resident (slot + 1 in bits 24 and up)
        kind, op = reg[:, 0] & 0xff, (reg[:, 0] >> 8) & 0xff
        fields = np.concatenate([reg[(kind == 2) & (op == 6)][:, 5:7].ravel(), reg[kind == 3][:, 2]])
        slots = fields >> 24
        assert set(np.unique(slots)) == set(range(FORM_REGS + 1))


SHAPES = {2: (2, 4, 2), 3: (2, 4, 2), 4: (6, 4, 4), 5: (4, 4, 6), 6: (4, 4, 4)}
DUAL_FILL, DUAL_THROUGH = 1 << 31, 1 << 30


def _slot_accesses(prog):
    """current components a fused program loads from or stores to the slot per point (on-chip
    fields take none, dual fields a load when bit 31 is set on a read and a store when bit 30 is
    set on a write; known-zero components none)"""
    n = 0

    def count(f, length, zero, load):
        nonlocal n
        f = int(f)
        if f >= LDS_PLANE and not (f & (DUAL_FILL if load else DUAL_THROUGH)):
            return
        n += sum(1 for j in range(length) if not (int(zero) >> j) & 1)

    for w in prog.astype(np.int64):
        kind, op, fl = w[0] & 0xff, (w[0] >> 8) & 0xff, (w[0] >> 16) & 0x3f
        if kind == 2:
            c0, c1, cd = SHAPES[op]
            count(w[2], c0, (w[0] >> 22) & 0x3f, True)
            count(w[3], c1, (w[0] >> 28) & 0xf, True)
            if (fl & 2) and not (fl & 1):
                count(w[4], cd, (w[1] >> 16) & 0x3f, True)
            if (fl & 4) and not (fl & F_HOLD):
                count(w[4], cd, (w[7] >> 16) & 0x3f, False)
        elif kind in (3, 4):
            m = (4 if op == 9 else 2) if kind == 3 else w[2]
            zin, zout = (w[4] & 0xff, (w[4] >> 8) & 0xff) if kind == 3 else (w[4] & 0xffff, w[4] >> 16)
            if not fl & F_FROMACC:
                count(w[1], m, zin, True)
            if not fl & F_HOLD:
                count(w[1], m, zout, False)
        elif kind == 5:
            count(w[1], w[6], w[7] & 0xff, True)
            count(w[2], w[6], (w[7] >> 8) & 0xff, True)
        elif kind == 1:
            count(w[1], w[2], 0, False)
        elif kind == 6:
            fam = w[5] & 7
            count(w[1], 1 if fam == 0 else 2 if fam == 1 else 4, w[6], False)
    return n


def test_on_chip_allocation(lib):
    """the allocator's programs: currents move on chip for parts of their lives (dual fields:
    a read that keeps an LDS copy, a write that also stores to the slot), and the slot carries
    far fewer accesses than with the whole program in it (g g > g g g: 2,210 after forwarding
    and the zero masks; 1,414 with the whole-lifetime placement of 20 LDS and 16 register
    planes, which the allocator and the row orders bring below 1,000)"""
    plan = plans.gluon_plan(5)
    fwd, _, _, _ = plan_program(lib, plan, 1)
    reg, _, _, _ = plan_program(lib, plan, 9)  # 20 LDS planes and the register tier
    fields = reg[:, 1:5].astype(np.int64)
    kind = reg[:, 0] & 0xff
    dual = (fields >= LDS_PLANE) & ((fields & (DUAL_FILL | DUAL_THROUGH)) != 0) & (kind[:, None] != 6)
    assert np.count_nonzero(dual & ((fields & DUAL_FILL) != 0)) > 0
    assert np.count_nonzero(dual & ((fields & DUAL_THROUGH) != 0)) > 0
    # a copy kept by a read is in LDS (slots below the variant's 20 planes)
    fill = fields[dual & ((fields & DUAL_FILL) != 0)]
    assert np.all((((fill & 0xffffff) + (fill >> 24)) & 63) < 20)
    assert _slot_accesses(fwd) == 2210
    assert _slot_accesses(reg) < 1000


def test_sources_made_again(lib):
    """the f64 kernel's programs run a source row again where a read would have filled an on-chip
    copy from the slot, and leave fewer slot accesses than the programs the other kernels run"""
    plan = plans.gluon_plan(5)
    for variant in (3, 4, 8, 9):
        prog, sources, _, _ = plan_program(lib, plan, variant)
        again, _, _, _ = plan_program(lib, plan, variant + 10)
        kind = again[:, 0] & 0xff
        assert np.all(kind[:sources] == 6) and np.count_nonzero(kind[sources:] == 6) > 0
        assert not np.any((prog[sources:, 0] & 0xff) == 6)
        # each one writes an on-chip slot (plain on-chip field) and stores nothing to the slot
        made = again[sources:][kind[sources:] == 6]
        assert np.all((made[:, 1] >= LDS_PLANE) & (made[:, 1] < LDS_PLANE + 64))
        assert _slot_accesses(again) < _slot_accesses(prog)
