# SPDX-License-Identifier: 0BSD
"""The C-ABI library without a GPU: it loads, exports every entry point include/*.h declares,
and rejects invalid calls before touching the device (status 2, as
runner/recurrence_runner.cu:1293-1305, 1587-1596, 1757-1776 do); the same error contract holds
for the oracle's emulation. No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pacuda_amd import abi
from pacuda_amd.abi import ptr, u32p, u64p
import harness
import refcpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("pacman_recurrence.h", "pacuda_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(pac(?:man_recurrence_v1|uda)_\w+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert "pacman_recurrence_v1_run" in names and "pacuda_evaluate" in names
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in abi.V1_SYMBOLS + abi.EXTENSION_SYMBOLS}
    assert set(names) <= bound, set(names) - bound


def test_struct_layouts_match_the_reference_sizes():
    # runner/recurrence_runner.cu:1293-1298 accepts 248, 208 and 192
    assert C.sizeof(abi.Call) == 248
    assert abi.Call.zero_amplitude.offset == 192
    assert abi.Call.external_momenta.offset == 208
    assert abi.Call.words.offset == 8 and abi.Call.width.offset == 28 and abi.Call.stride.offset == 180
    assert C.sizeof(abi.Reduce) == 40
    assert C.sizeof(abi.ReduceResolved) == 64
    assert C.sizeof(abi.ReduceSample) == 96
    assert abi.ReduceSample.random.offset == 48 and abi.ReduceSample.chosen.offset == 88
    assert C.sizeof(abi.Options) == 24


def test_capabilities_without_device(lib):
    import hashlib
    from pacuda_amd import build
    caps = lib.pacman_recurrence_v1_capabilities()
    assert caps == abi.CAP_SOURCES | abi.CAP_FORMS | abi.CAP_SYMJIT | abi.CAP_FUSED
    assert lib.pacman_recurrence_v1_symjit_kernel(None) == -1
    assert lib.pacman_recurrence_v1_symjit_kernel(b"0" * 64) == -1
    # the generated kernels are named by the SHA-256 of their plane-kernel files, either case
    # (runner/recurrence_runner.cu:1792-1797)
    for k, path in enumerate(build.kernel_files()):
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        assert lib.pacman_recurrence_v1_symjit_kernel(digest.encode()) == k
        assert lib.pacman_recurrence_v1_symjit_kernel(digest.upper().encode()) == k


def _invalid_calls():
    words = np.array([1 | 3 << 4 | 2 << 16], dtype=np.uint64)  # a clear row
    base = abi.Call()
    base.abi_version = abi.ABI
    base.struct_size = C.sizeof(abi.Call)
    base.words = ptr(words, u64p)
    base.word_count = 1
    base.width[0] = 12
    base.current_planes = 8
    base.stride = 64
    base.points = 64
    cases = []
    for field, value in [("abi_version", 2), ("struct_size", 200), ("struct_size", 0), ("points", 0), ("points", 65)]:
        c = abi.Call.from_buffer_copy(base)
        setattr(c, field, value)
        cases.append((f"{field}={value}", c))
    c = abi.Call.from_buffer_copy(base)
    c.zero_amplitude_runs = 1
    cases.append(("zero runs without a list", c))
    c = abi.Call.from_buffer_copy(base)
    c.words = None
    cases.append(("no words", c))
    return words, cases


@pytest.mark.parametrize("which", ["hip", "oracle"])
def test_invalid_runs(lib, which):
    v1 = harness.V1(lib) if which == "hip" else harness.V1(refcpu.load(), prefix="refcpu_v1_")
    words, cases = _invalid_calls()
    for name, c in cases:
        assert v1.run(C.byref(c)) == abi.INVALID, name


@pytest.mark.parametrize("which", ["hip", "oracle"])
def test_reductions_without_a_kept_run(lib, which):
    v1 = harness.V1(lib) if which == "hip" else harness.V1(refcpu.load(), prefix="refcpu_v1_")
    out = np.zeros(4)
    planes = np.zeros(1, dtype=np.uint32)
    weights = np.ones(1)
    offsets = np.array([0, 1], dtype=np.uint32)
    r = abi.Reduce(abi.ABI, C.sizeof(abi.Reduce), ptr(planes, u32p), ptr(weights), ptr(out), 1, 4)
    assert v1.reduce(C.byref(r)) == abi.INVALID
    r.struct_size = 8
    assert v1.reduce(C.byref(r)) == abi.INVALID
    rr = abi.ReduceResolved(abi.ABI, C.sizeof(abi.ReduceResolved), ptr(planes, u32p), ptr(weights),
                            ptr(offsets, u32p), ptr(out), 1, 4, 1, 1, 0)
    assert v1.reduce_resolved(C.byref(rr)) == abi.INVALID
    rr.outputs = 70000
    assert v1.reduce_resolved(C.byref(rr)) == abi.INVALID
    rnd = np.zeros(4)
    tot = np.zeros(4)
    chosen = np.zeros(4, dtype=np.uint32)
    hel = np.zeros(1, dtype=np.uint32)
    rs = abi.ReduceSample(abi.ABI, C.sizeof(abi.ReduceSample), ptr(planes, u32p), ptr(weights), ptr(offsets, u32p),
                          ptr(hel, u32p), 1, 4, ptr(rnd), None, 1.0, 1.0, ptr(tot), ptr(chosen, u32p))
    assert v1.reduce_sample(C.byref(rs)) == abi.INVALID
    rs.random = None
    assert v1.reduce_sample(C.byref(rs)) == abi.INVALID


@pytest.mark.parametrize("which", ["hip", "oracle"])
def test_configure_rules(lib, which):
    v1 = harness.V1(lib) if which == "hip" else harness.V1(refcpu.load(), prefix="refcpu_v1_")
    o = abi.Options(abi.ABI, C.sizeof(abi.Options), 4, 0, 0, 0)
    assert v1.configure(C.byref(o)) == abi.INVALID  # no such kernel
    o = abi.Options(abi.ABI, C.sizeof(abi.Options), abi.KERNEL_F64, 16, 0, 0)
    assert v1.configure(C.byref(o)) == abi.INVALID  # block below 32
    o = abi.Options(abi.ABI, C.sizeof(abi.Options), abi.KERNEL_F64, 2048, 0, 0)
    assert v1.configure(C.byref(o)) == abi.INVALID
    o = abi.Options(abi.ABI, 20, abi.KERNEL_F64, 0, 0, 0)
    assert v1.configure(C.byref(o)) == abi.INVALID  # struct too short
    o = abi.Options(2, C.sizeof(abi.Options), abi.KERNEL_F64, 0, 0, 0)
    assert v1.configure(C.byref(o)) == abi.INVALID


def test_fused_rejects_invalid_plans_and_launches(lib):
    from pacuda_amd import plans
    plan = plans.gluon_plan(4)
    a = plan.arrays()
    d = abi.PlanDesc()
    d.abi_version = abi.ABI
    d.struct_size = C.sizeof(abi.PlanDesc) - 8
    h = C.c_void_p()
    assert lib.pacuda_plan_create(C.byref(d), C.byref(h)) == abi.INVALID
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), 7, 0, 0, 0, None, None)
    items = (abi.Item * 1)()
    assert lib.pacuda_evaluate(items, 1, C.byref(l)) == abi.INVALID  # kernel 7
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), abi.KERNEL_F64, 0, 96 + 1, 0, None, None)
    assert lib.pacuda_evaluate(items, 1, C.byref(l)) == abi.INVALID  # block not a multiple of 64
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), abi.KERNEL_F64, 0, 0, 0, None, None)
    assert lib.pacuda_evaluate(items, 1, C.byref(l)) == abi.INVALID  # no plan
    assert lib.pacuda_evaluate(items, 0, C.byref(l)) == abi.OK       # nothing to do
    del a


def test_plan_compilation_rejects_bad_descriptions(lib):
    """pacuda_plan_program (host only) refuses what pacuda_plan_create refuses, without a GPU:
    out-of-range tables, decreasing offsets, reads of unwritten planes, symjit rows, a block of
    1024 (the fused kernel's limit is 256), and a too-small output buffer"""
    from pacuda_amd import plans
    from pacuda_amd.runtime import plan_desc
    plan = plans.gluon_plan(4)
    n = C.c_uint32()

    def compile_with(**changes):
        a = {k: np.array(v, copy=True) for k, v in plan.arrays().items()}
        for k, v in changes.items():
            if k in a:
                a[k] = v
        d = plan_desc(plan, a)
        for k, v in changes.items():
            if k not in a:
                setattr(d, k, v)
        return lib.pacuda_plan_program(C.byref(d), 1, None, 0, C.byref(n), None, None, None), a

    st, _ = compile_with()
    assert st == abi.OK and n.value > 0
    assert compile_with(current_planes=plan.current_planes - 1)[0] == abi.INVALID  # indexes past the planes
    a = plan.arrays()
    bad = np.array(a["planes"], copy=True)
    bad[0] = plan.amplitude_planes
    assert compile_with(planes=bad)[0] == abi.INVALID  # an output plane past the amplitudes
    off = np.array(a["offsets"], copy=True)
    off[1], off[2] = off[2], off[1]
    assert compile_with(offsets=off)[0] == abi.INVALID  # decreasing offsets
    slots = np.array(a["term_slots"], copy=True)
    slots[0] = plan.legs
    assert compile_with(term_slots=slots)[0] == abi.INVALID  # a form term past the legs
    # a program that reads a plane before anything writes it: drop the source rows
    words = np.array(a["words"], copy=True)
    ops = words & np.uint64(15)
    assert compile_with(words=words[ops != 12])[0] == abi.INVALID
    # a malformed binding (an op-14 word made from another row's bits) is refused
    sym = np.concatenate([words[:1] & ~np.uint64(15) | np.uint64(14), words])
    assert compile_with(words=sym)[0] == abi.INVALID
    # variant out of range, missing count, too small a buffer
    d = plan_desc(plan, a)
    assert lib.pacuda_plan_program(C.byref(d), 22, None, 0, C.byref(n), None, None, None) == abi.INVALID
    assert lib.pacuda_plan_program(C.byref(d), 1, None, 0, None, None, None, None) == abi.INVALID
    small = np.zeros((1, 8), dtype=np.uint32)
    assert lib.pacuda_plan_program(C.byref(d), 1, small.ctypes.data_as(abi.u32p), 1, C.byref(n), None, None,
                                   None) == abi.INVALID
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), abi.KERNEL_F64, 0, 1024, 0, None, None)
    items = (abi.Item * 1)()
    assert lib.pacuda_evaluate(items, 1, C.byref(l)) == abi.INVALID
