# SPDX-License-Identifier: 0BSD
"""GPU parity: librecurrence.so on gfx950 against the oracle, bit for bit.

- the reference's own synthetic-plan client runner/rectest.c (built into oracle/_ref) passes its
  internal host-vs-device KATs against the HIP library for every kernel, and its printed
  amplitudes equal the oracle's emulation of the same v1 calls byte for byte;
- the source wavefunctions (pacman_recurrence_v1_test_sources) equal the oracle's bit for bit on
  the cases of tests/check_sources.py's and rusticol/patches/0009's generators, committed as data
  with their closed-form values (tests/golden/source_cases.npz);
- generated n-gluon plans through the v1 path and the fused path equal the oracle's whole plan
  run bit for bit in every kernel (f64, df64, mixed, mixed64), on ragged point counts."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
import harness
from harness import V1, FusedPlan, DeviceArray, fused_evaluate, bits_equal
import refcpu

pytestmark = pytest.mark.gpu

KERNEL_ARGS = [("f64",), ("f64", "nosort"), ("df64",), (), ("mixed",), ("mixed64",)]


@pytest.fixture(scope="module")
def lib():
    return abi.load()


@pytest.mark.parametrize("args", KERNEL_ARGS, ids=lambda a: "-".join(a) or "default")
def test_rectest_against_hip_library(args):
    print("rectest", args, flush=True)
    code, out, err = harness.rectest(abi.LIBRARY, *args)
    assert code == 0, f"rectest {args}: exit {code}: {err}"
    code_o, out_o, err_o = harness.rectest(refcpu.ABI_LIBRARY, *args)
    assert code_o == 0, err_o
    assert len(out.split()) == 3 * 300 * 2
    assert out == out_o, "HIP amplitudes differ from the oracle's"


def test_capabilities(lib):
    caps = lib.pacman_recurrence_v1_capabilities()
    assert caps & abi.CAP_SOURCES and caps & abi.CAP_FORMS and caps & abi.CAP_FUSED and caps & abi.CAP_SYMJIT
    assert lib.pacman_recurrence_v1_symjit_kernel(b"0" * 64) == -1


def test_sources_bitwise(lib):
    """the source wavefunctions on the GPU equal the oracle's bit for bit and the closed forms of
    convert/CONVENTIONS.md:52-61 within 1e-13, on the committed case vectors
    (tests/golden/source_cases.npz)"""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "source_cases.npz"))
    codes, masses, mom = g["codes"], g["masses"], g["momenta"].reshape(-1)
    n = len(codes)
    print("sources", n, flush=True)
    re, im = np.zeros(4 * n), np.zeros(4 * n)
    st = lib.pacman_recurrence_v1_test_sources(abi.ptr(codes, abi.u32p), abi.ptr(masses), abi.ptr(mom), n,
                                               abi.ptr(re), abi.ptr(im))
    assert st == 0
    _, ore, oim = refcpu.test_sources(codes, masses, mom)
    bad = np.nonzero((re.view(np.uint64) != ore.reshape(-1).view(np.uint64)) |
                     (im.view(np.uint64) != oim.reshape(-1).view(np.uint64)))[0]
    assert len(bad) == 0, f"{len(bad)} components differ, first case {bad[0] // 4}: code {codes[bad[0] // 4]}"
    got = (re + 1j * im).reshape(n, 4)
    want = g["legend_re"] + 1j * g["legend_im"]
    scale = np.maximum(np.max(np.abs(got), axis=1), 1e-300)
    err = np.max(np.abs(got - want), axis=1) / scale
    assert np.max(err) < 1e-13, f"case {int(np.argmax(err))}: {np.max(err)}"


def _points(n, count, seed):
    mom, _ = ps.generate(n, count, seed)
    return mom


@pytest.fixture(scope="module")
def gg_ggg():
    plan = plans.gluon_plan(5)
    mom = _points(5, 16384, 1)
    return plan, ps.external_planes(mom, plan.crossing)


def test_v1_f64_gluon5_matches_oracle(lib, gg_ggg):
    plan, ext = gg_ggg
    v1 = V1(lib)
    st = v1.configure_kernel(abi.KERNEL_F64)
    assert st == 0
    for points, stride in [(16384, 16384), (1000, 1024), (1, 64), (333, 333)]:
        print("v1 f64 points", points, "stride", stride, flush=True)
        got = v1.run_plan(plan, ext, points, stride)
        want, _ = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext[:, :points], points, kernel=1)
        assert bits_equal(got, want), f"points {points}: max |diff| {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("kernel", ["f64", "df64", "mixed", "mixed64", "compensated", "compensated64"])
def test_fused_gluon5_matches_oracle(lib, gg_ggg, kernel):
    plan, ext = gg_ggg
    k = abi.FUSED_KERNELS[kernel]
    fp = FusedPlan(lib, plan)
    points = 16384 - 37  # a ragged last unit
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    outputs = plan.amplitude_planes
    d_res = DeviceArray(lib, points * outputs * 8)
    d_tot = DeviceArray(lib, points * 8)
    print("fused", kernel, points, flush=True)
    fused_evaluate(lib, [(fp, d_ext, points, ext.shape[1], d_res, d_tot)], k)
    res = d_res.download((points, outputs))
    tot = d_tot.download(points)
    want_res, want_tot = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext[:, :points], points, kernel=k)
    assert bits_equal(res, want_res), f"{kernel}: resolved differs, max {np.max(np.abs(res - want_res))}"
    assert bits_equal(tot, want_tot), f"{kernel}: totals differ"
    for d in (d_ext, d_res, d_tot):
        d.free()
    fp.destroy()


def test_fused_multi_plan_queue(lib):
    """several plans and uneven point counts in one persistent launch == the oracle per item"""
    jobs, wants, devs, fps = [], [], [], []
    for n, points, seed in [(4, 70, 11), (5, 1000, 12), (4, 1, 13), (5, 129, 14), (6, 200, 15), (5, 0, 16)]:
        plan = plans.gluon_plan(n)
        ext = ps.external_planes(_points(n, max(points, 1), seed), plan.crossing)
        fp = FusedPlan(lib, plan)
        d_ext = DeviceArray(lib, ext.nbytes)
        d_ext.upload(ext)
        d_tot = DeviceArray(lib, max(points, 1) * 8)
        jobs.append((fp, d_ext, points, ext.shape[1], None, d_tot))
        devs += [d_ext, d_tot]
        fps.append(fp)
        if points:
            _, want = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext[:, :points], points, kernel=1)
        else:
            want = np.zeros(0)
        wants.append(want)
    print("multi-plan launch", flush=True)
    fused_evaluate(lib, jobs, abi.KERNEL_F64, waves_per_cu=1)
    for (fp, _, points, _, _, d_tot), want in zip(jobs, wants):
        got = d_tot.download(points)
        assert bits_equal(got, want), f"{fp.plan.name} {points} points"
    for d in devs:
        d.free()
    for fp in fps:
        fp.destroy()


def test_v1_other_kernels_in_fresh_processes(gg_ggg, tmp_path):
    """the v1 path of df64, mixed and mixed64 (one configuration per process) == oracle"""
    plan, ext = gg_ggg
    points = 4096 + 5
    np.save(tmp_path / "ext.npy", ext[:, :points])
    for kernel in ["df64", "mixed", "mixed64"]:
        print("v1 driver", kernel, flush=True)
        out = tmp_path / f"{kernel}.npy"
        p = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "v1_driver.py"), kernel,
                            str(tmp_path / "ext.npy"), str(out), "5"], capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        got = np.load(out)
        want, _ = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext[:, :points], points,
                                  kernel=abi.KERNELS[kernel])
        assert bits_equal(got, want), f"{kernel}: max |diff| {np.max(np.abs(got - want))}"


def test_recording_replays_bit_exact_on_gpu():
    """tests/golden/gg_ggg_qqbar.pacuda through pacuda-check on the HIP library: every f64 value
    equals the recorded CPU value (device and host reductions, the reference points read back
    from an examples.json the tool writes, and keyed calls, which run on the fused kernels)"""
    from pacuda_amd import build
    golden = os.path.join(os.path.dirname(__file__), "golden", "gg_ggg_qqbar.pacuda")
    examples = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"examples_{os.getpid()}.json")
    p = subprocess.run([build.CHECK_TOOL, golden, "--write-examples", examples], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0, p.stderr
    for args in [(), ("--host-reduce", "--tile", "3"), ("--examples", examples), ("--key",),
                 ("--key", "--host-reduce", "--tile", "5")]:
        print("pacuda-check", args, flush=True)
        p = subprocess.run([build.CHECK_TOOL, golden, "--runner", abi.LIBRARY, *args], capture_output=True, text=True,
                           timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        assert "0 of 256 values differ" in p.stdout and "0 of 128 values differ" in p.stdout
        fused = "runs on the fused kernels: 0" not in p.stdout
        assert fused == ("--key" in args), p.stdout


def test_fused_colour_orderings_match_oracle(lib, gg_ggg):
    """the 12 orderings of g g > g g g (legs 1..4 permuted, up to reflection) as 12 items of one
    plan on permuted external legs in one launch == the oracle's plan run per ordering"""
    import itertools
    plan, ext = gg_ggg
    n = 700
    fp = FusedPlan(lib, plan)
    orders = [(0,) + p for p in itertools.permutations(range(1, 5)) if p[0] < p[-1]]
    assert len(orders) == 12
    jobs, devs, wants = [], [], []
    for order in orders:
        rows = np.concatenate([ext[4 * s:4 * s + 4, :n] for s in order])
        d_e = DeviceArray(lib, rows.nbytes)
        d_e.upload(rows)
        d_t = DeviceArray(lib, n * 8)
        jobs.append((fp, d_e, n, n, None, d_t))
        devs += [d_e, d_t]
        wants.append(refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), rows, n, kernel=1)[1])
    print("orderings launch", flush=True)
    fused_evaluate(lib, jobs, abi.KERNEL_F64)
    for order, job, want in zip(orders, jobs, wants):
        assert bits_equal(job[5].download(n), want), order
    # the orderings differ (a leading-colour sum is not 12 copies of one)
    assert not np.allclose(wants[0], wants[1])
    for d in devs:
        d.free()
    fp.destroy()


@pytest.mark.parametrize("kernel", ["f64", "df64", "mixed", "compensated", "compensated64"])
def test_fused_launch_shapes_bitwise(lib, gg_ggg, kernel):
    """every LDS variant (40, 26, 20, 13 and 10 LDS-resident planes per wavefront, picked from
    the requested occupancy and block) and the launch without LDS residency give the oracle's
    bits"""
    plan, ext = gg_ggg
    k = abi.FUSED_KERNELS[kernel]
    fp = FusedPlan(lib, plan)
    points = 5000 + 3
    d_ext = DeviceArray(lib, ext.nbytes)
    d_ext.upload(ext)
    d_tot = DeviceArray(lib, points * 8)
    _, want = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext[:, :points], points, kernel=k, resolved=False)
    for wpc, block in [(4, 256), (6, 64), (8, 256), (12, 256), (16, 256), (16, 64), (24, 64), (32, 256)]:
        print("fused", kernel, "waves per CU", wpc, "block", block, flush=True)
        fused_evaluate(lib, [(fp, d_ext, points, ext.shape[1], None, d_tot)], k, waves_per_cu=wpc, block=block)
        got = d_tot.download(points)
        assert bits_equal(got, want), f"{kernel} {wpc}x{block}: max |diff| {np.max(np.abs(got - want))}"
    d_ext.free()
    d_tot.free()
    fp.destroy()


def test_fused_pp_jjj_subprocesses_bitwise(lib):
    """every p p > j j j subprocess plan (quark lines: ops 2, 3, 7, 8) in one persistent launch,
    uneven point counts, f64, mixed and both compensated policies, equals the oracle's plan run bit for bit"""
    from pacuda_amd import processes
    subs = processes.subprocesses()
    jobs, devs, fps, exts = [], [], [], []
    for k, sp in enumerate(subs):
        points = 64 * (k % 5) + 17 * k + 1
        mom, _ = ps.generate(sp.plan.legs, points, 300 + k)
        ext = ps.external_planes(mom, sp.plan.crossing)
        fp = FusedPlan(lib, sp.plan)
        d_ext = DeviceArray(lib, ext.nbytes)
        d_ext.upload(ext)
        d_tot = DeviceArray(lib, points * 8)
        jobs.append((fp, d_ext, points, ext.shape[1], None, d_tot))
        devs += [d_ext, d_tot]
        fps.append(fp)
        exts.append(ext)
    for kernel in ("f64", "mixed", "compensated", "compensated64"):
        k = abi.FUSED_KERNELS[kernel]
        print("pp>jjj launch", kernel, flush=True)
        fused_evaluate(lib, jobs, k)
        for sp, (fp, _, points, _, _, d_tot), ext in zip(subs, jobs, exts):
            _, want = refcpu.evaluate(sp.plan.arrays(), harness.plan_counts(sp.plan), ext, points, kernel=k,
                                      resolved=False)
            assert bits_equal(d_tot.download(points), want), f"{kernel} {sp.name}"
    for d in devs:
        d.free()
    for fp in fps:
        fp.destroy()
