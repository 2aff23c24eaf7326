# SPDX-License-Identifier: 0BSD
"""The v1 entry points on the fused kernels, one kernel configuration per process (the first run
fixes it). A sequence of pacman_recurrence_v1_run calls, each followed by its reduction or
amplitude download, goes to the HIP library and to the oracle's emulation of the v1 runner
(oracle/refcpu.cpp) alike; every output must agree bit for bit, and pacuda_v1_fused_runs must
count exactly the calls that qualify for the fused kernels (keyed, forms on the device, nothing
uploaded but the external momenta, no current downloaded, closures writing cleared planes only).
Prints one line per call; exits nonzero on the first difference.
usage: v1_fused_driver.py KERNEL"""
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

import numpy as np  # noqa: E402

from pacuda_amd import abi, plans, processes, phase_space as ps  # noqa: E402
from pacuda_amd.abi import ptr, u32p, u64p  # noqa: E402
from harness import V1  # noqa: E402
import refcpu  # noqa: E402


class Caller:
    """one v1 library (the HIP one or the oracle's) and the host arrays of its calls"""

    def __init__(self, v1):
        self.v1 = v1
        self.hold = []

    def run(self, plan, a, ext, points, stride, key, keep, zero_runs, upload_current=None, scales=None, extra=0):
        n_amp = plan.amplitude_planes + extra
        are = np.zeros(n_amp * stride)
        aim = np.zeros(n_amp * stride)
        cre = np.full(plan.current_planes * stride, 0.25)
        cim = np.full(plan.current_planes * stride, -0.5)
        ext = np.ascontiguousarray(ext[:, :points])
        zr = np.array(zero_runs, dtype=np.uint32).reshape(-1)
        dl = np.array([0, n_amp], dtype=np.uint32)
        sre, sim = (a["scale_re"], a["scale_im"]) if scales is None else scales
        self.hold += [are, aim, cre, cim, ext, zr, dl, sre, sim]
        c = abi.Call()
        c.abi_version, c.struct_size = abi.ABI, C.sizeof(abi.Call)
        c.words, c.words_key, c.word_count = ptr(a["words"], u64p), key, len(a["words"])
        for k in range(8):
            c.width[k] = int(a["width"][k])
        c.current_re, c.current_im, c.amplitude_re, c.amplitude_im = ptr(cre), ptr(cim), ptr(are), ptr(aim)
        c.scale_re, c.scale_im = ptr(sre), ptr(sim)
        c.factor_re, c.factor_im = ptr(a["factor_re"]), ptr(a["factor_im"])
        c.current_planes, c.amplitude_planes, c.momentum_planes = plan.current_planes, n_amp, 4 * len(plan.forms)
        c.scale_count, c.factor_count = len(sre), len(a["factor_re"])
        if not keep:
            c.download_amplitude, c.download_amplitude_runs = ptr(dl, u32p), 1
        if upload_current is not None:
            up = np.array(upload_current, dtype=np.uint32)
            self.hold.append(up)
            c.upload_current, c.upload_current_runs = ptr(up, u32p), len(up) // 2
        c.stride, c.points = stride, points
        c.zero_amplitude, c.zero_amplitude_runs = ptr(zr, u32p), len(zr) // 2
        c.flags = (abi.KEEP_AMPLITUDES if keep else 0) | abi.FORMS_ON_DEVICE
        c.external_momenta, c.external_count = ptr(ext), plan.legs
        c.term_count = len(a["term_slots"])
        c.form_terms, c.term_slots = ptr(a["form_terms"], u32p), ptr(a["term_slots"], u32p)
        c.term_coefficients = ptr(a["term_coefficients"])
        st = self.v1.run(C.byref(c))
        if st != 0:
            raise RuntimeError(f"run: status {st}")
        return are.reshape(n_amp, stride)[:, :points], aim.reshape(n_amp, stride)[:, :points]

    def resolved(self, a, points, extra_plane=None):
        planes, offsets, weights = a["planes"], a["offsets"], a["weights"]
        if extra_plane is not None:  # one more output: a plane no closure writes
            planes = np.append(planes, np.uint32(extra_plane)).astype(np.uint32)
            offsets = np.append(offsets, np.uint32(len(planes))).astype(np.uint32)
            weights = np.append(weights, 1.0)
        outputs = len(offsets) - 1
        out = np.full(points * outputs, np.nan)
        self.hold += [planes, offsets, weights, out]
        r = abi.ReduceResolved(abi.ABI, C.sizeof(abi.ReduceResolved), ptr(planes, u32p), ptr(weights),
                               ptr(offsets, u32p), ptr(out), outputs, points, 1, outputs, abi.OUTPUT_ZERO)
        st = self.v1.reduce_resolved(C.byref(r))
        if st != 0:
            raise RuntimeError(f"reduce_resolved: status {st}")
        return out

    def sample(self, plan, a, points, random, scale):
        outputs = len(a["offsets"]) - 1
        hel = np.arange(outputs, dtype=np.uint32)[::-1].copy()
        tot = np.full(points, np.nan)
        cho = np.zeros(points, dtype=np.uint32)
        self.hold += [hel, tot, cho, random, scale]
        r = abi.ReduceSample(abi.ABI, C.sizeof(abi.ReduceSample), ptr(a["planes"], u32p), ptr(a["weights"]),
                             ptr(a["offsets"], u32p), ptr(hel, u32p), outputs, points, ptr(random), ptr(scale),
                             3.0, 7.0, ptr(tot), ptr(cho, u32p))
        st = self.v1.reduce_sample(C.byref(r))
        if st != 0:
            raise RuntimeError(f"reduce_sample: status {st}")
        return np.concatenate([tot, cho.astype(np.float64)])

    def plain(self, a, points, start):
        out = start.copy()
        self.hold.append(out)
        r = abi.Reduce(abi.ABI, C.sizeof(abi.Reduce), ptr(a["planes"], u32p), ptr(a["weights"]), ptr(out),
                       len(a["planes"]), points)
        st = self.v1.reduce(C.byref(r))
        if st != 0:
            raise RuntimeError(f"reduce: status {st}")
        return out


def inputs(plan, points, seed):
    mom, _ = ps.generate(plan.legs, points, seed=seed)
    return ps.external_planes(mom, plan.crossing)


def main():
    kernel = sys.argv[1]
    lib = abi.load()
    hip = Caller(V1(lib))
    olib = refcpu.load()
    olib.refcpu_reset()
    oracle = Caller(V1(olib, prefix="refcpu_v1_"))
    for c in (hip, oracle):
        assert c.v1.configure_kernel(abi.KERNELS[kernel]) == 0

    g5, g6, q5 = plans.gluon_plan(5), plans.gluon_plan(6), plans.quark_line_plan(5)
    four = processes.four_quark_subprocesses()[3].plan
    a5, a6, aq, a4 = g5.arrays(), g6.arrays(), q5.arrays(), four.arrays()
    n5, s5 = 2053, 2060
    e5 = inputs(g5, n5, 71)
    everything = lambda plan: [0, plan.amplitude_planes]  # noqa: E731
    rng = np.random.default_rng(5)

    def step(name, fused, f):
        before = lib.pacuda_v1_fused_runs()
        got = f(hip)
        after = lib.pacuda_v1_fused_runs()
        want = f(oracle)
        for g, w in zip(got if isinstance(got, tuple) else (got,), want if isinstance(want, tuple) else (want,)):
            g, w = np.ascontiguousarray(g, dtype=np.float64), np.ascontiguousarray(w, dtype=np.float64)
            if g.shape != w.shape or not np.array_equal(g.view(np.uint64), w.view(np.uint64)):
                print(f"{kernel} {name}: {np.count_nonzero(g != w)} values differ", flush=True)
                sys.exit(1)
        if after - before != (1 if fused else 0):
            print(f"{kernel} {name}: fused runs {after - before}, expected {int(fused)}", flush=True)
            sys.exit(1)
        print(f"{kernel} {name}: equal ({'fused kernels' if fused else 'v1 kernels'})", flush=True)

    # 1. g g > g g g, stride beyond the points, two more amplitude planes than the closures
    # write (cleared by the run), kept, reduce_resolved with one of them among the outputs
    extra = g5.amplitude_planes + 1
    step("g g > g g g, kept, reduce_resolved", True,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, [0, extra + 1], extra=2),
                    c.resolved(a5, n5, extra_plane=extra))[1])
    # 2. g g > g g g g, kept, reduce_sample with per-point scales
    n6 = 1541
    e6 = inputs(g6, n6, 72)
    rnd, sc = rng.random(n6), np.where(rng.random(n6) < 0.5, 1.0, 1.25)
    step("g g > g g g g, kept, reduce_sample", True,
         lambda c: (c.run(g6, a6, e6, n6, n6, 0x6001, True, everything(g6)), c.sample(g6, a6, n6, rnd, sc))[1])
    # 3. a quark line, the amplitudes downloaded (not kept)
    nq = 777
    eq = inputs(q5, nq, 73)
    step("q q~ > g g g line, amplitudes downloaded", True,
         lambda c: c.run(q5, aq, eq, nq, nq + 1, 0x7001, False, everything(q5)))
    # 4. a four-quark subprocess (symjit rows), kept, the plain reduction onto given values
    n4 = 999
    e4 = inputs(four, n4, 74)
    start = rng.random(n4)
    step("four-quark subprocess (symjit rows), kept, reduce", True,
         lambda c: (c.run(four, a4, e4, n4, n4, 0x8001, True, everything(four)), c.plain(a4, n4, start))[1])
    # 5. the first key again with other scales: the plan is built again
    sre, sim = a5["scale_re"] * 1.5, a5["scale_im"] - 0.25
    step("g g > g g g, the same key with other scales", True,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, everything(g5), scales=(sre, sim)),
                    c.resolved(a5, n5))[1])
    # 6. the last amplitude plane not cleared: its closures add to what the previous run left
    step("g g > g g g, one written plane not cleared", False,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, [0, g5.amplitude_planes - 1]), c.resolved(a5, n5))[1])
    # 7. unkeyed
    step("g g > g g g, unkeyed", False,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0, True, everything(g5)), c.resolved(a5, n5))[1])
    # 8. keyed, with a current run uploaded (the program overwrites it)
    step("g g > g g g, an uploaded current run", False,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, everything(g5), upload_current=[0, 3]),
                    c.resolved(a5, n5))[1])
    # 9. keyed, the fused kernels switched off in the environment
    os.environ["PACUDA_V1_FUSED"] = "0"
    step("g g > g g g, PACUDA_V1_FUSED=0", False,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, everything(g5)), c.resolved(a5, n5))[1])
    del os.environ["PACUDA_V1_FUSED"]
    # 10. and on again, after a run on the v1 kernels
    step("g g > g g g, fused again", True,
         lambda c: (c.run(g5, a5, e5, n5, s5, 0x5001, True, everything(g5)), c.resolved(a5, n5))[1])
    # 11. enough points for the run's upload to go in chunks, each launched when it has arrived
    nb = (1 << 18) + 4099
    eb = inputs(g5, nb, 75)
    rb, sb = rng.random(nb), np.ones(nb)
    step("g g > g g g, 2^18 + 4099 points (chunked upload), kept, reduce_sample", True,
         lambda c: (c.run(g5, a5, eb, nb, nb + 1, 0x5001, True, everything(g5)), c.sample(g5, a5, nb, rb, sb))[1])
    print(f"{kernel}: all equal", flush=True)


if __name__ == "__main__":
    main()
