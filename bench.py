# SPDX-License-Identifier: 0BSD
"""Benchmark of the recurrence runner on one MI355X.

Headline workload (`value`): g g > g g g, leading colour, one colour ordering with all 32
helicity configurations, f64 (bit-exact to the CPU oracle), 2^22 RAMBO points (seed 2) at
sqrt(s) = 1 TeV with pT > 20 GeV, |eta| < 2.5, dR > 0.4. A step is one pass of the hot path over
the batch: the fused evaluation (forms, sources, the plan's program, the helicity reduction and
sum) of every point, momenta already in HBM, |M|^2 per point back into HBM.

The JSON line also carries: the roofline of the fused f64 kernel (its time from HIP events the
library records on the kernel's stream); the CPU oracle (the same plan interpreter, OpenMP)
timed on a bounded sample of the same points; df64/mixed/mixed64/compensated(64) throughput and error
against f64; g g > g g g g (2^21 points, every mode, the persistent grid tuned by a sweep); the
p p > j j j batch (all 181 subprocesses, four-quark ones included, in one persistent launch, Zipf
point counts summing to 2^22); the headline points through the reference's v1 entry points
(host memory in and out); and surrogate vs plain unweighting of g g > g g g g."""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pacuda_amd import abi, plans, processes, phase_space as ps  # noqa: E402
from pacuda_amd.runtime import (Device, FusedPlan, autotune, evaluate, shape_name, TUNE_SHAPES,  # noqa: E402
                                COOP_TUNE_SHAPES, COOP_TUNE_SHAPES_ANY)
from pacuda_amd.unweight import Generator, Unweighting  # noqa: E402

METRIC = "|M|^2 points/s (f64 bit-exact) + unweighted events/s, surrogate vs plain"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_TBS = 8.0        # MI355X HBM3E


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def points_on_device(dev, plan, n, seed):
    """n cut-passing RAMBO points of stream `seed`, generated and compacted on the GPU"""
    gen = Generator(dev, plan)
    draw = int(n * 1.1) + 4096
    while True:
        ext, feat, ok = gen.draw(seed, 0, draw)
        idx = np.nonzero(ok)[0]
        if len(idx) >= n:
            break
        ext.free()
        feat.free()
        draw = int(draw * 1.2 * n / max(len(idx), 1)) + 4096
    out, _ = gen.gather(ext, None, draw, idx[:n])
    ext.free()
    feat.free()
    return out, int(idx[n - 1]) + 1


def points_from_record_stream(dev, lib, plan, n, seed):
    """n cut-passing points of the reference's own RAMBO stream (convert/src/record.rs:17-52, from
    state phase_space.record_state(seed)), generated on the host by the library and uploaded in the
    runner's layout"""
    import ctypes as C
    mom = np.zeros((n, plan.legs, 4))
    drawn = C.c_uint64()
    st = lib.pacuda_rambo_record(ps.record_state(seed), plan.legs, n, ps.ECM, 20.0, 2.5, 0.4,
                                 mom.ctypes.data_as(abi.f64p), C.byref(drawn))
    if st != 0:
        raise RuntimeError(f"pacuda_rambo_record: status {st}")
    return dev.array(ps.external_planes(mom, plan.crossing)), int(drawn.value)


def host_threads():
    """the CPUs this process may run on, within OMP_NUM_THREADS when that is set"""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() and int(omp) > 0 else max(1, n)


def timed_steps(lib, jobs, kernel, steps, waves_per_cu=0, block=0, unit_waves=0):
    """wall seconds of `steps` launches (synchronised) and the mean kernel ms (HIP events)"""
    ms = []
    t = time.perf_counter()
    for _ in range(steps):
        ms.append(evaluate(lib, jobs, kernel, waves_per_cu=waves_per_cu, block=block, timed=True,
                           unit_waves=unit_waves))
    lib.pacuda_synchronize()
    return time.perf_counter() - t, float(np.mean(ms))


def cpu_baseline(plan, ext_host, gpu_total, sample, runs=5):
    """the oracle's OpenMP plan interpreter on the host cores: median of `runs` timed runs over
    `sample` points (all threads) and over 2^14 points (one thread), after a warm-up"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    counts = dict(current_planes=plan.current_planes, amplitude_planes=plan.amplitude_planes,
                  form_count=len(plan.forms), external_count=plan.legs)
    arrays = plan.arrays()
    threads = host_threads()
    refcpu.evaluate(arrays, counts, ext_host[:, :4096].copy(), 4096, kernel=1, threads=threads, resolved=False)
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        _, tot = refcpu.evaluate(arrays, counts, ext_host, sample, kernel=1, threads=threads, resolved=False)
        times.append(time.perf_counter() - t)
    dt_all = float(np.median(times))
    one = min(sample, 1 << 14)
    ext_one = np.ascontiguousarray(ext_host[:, :one])
    times1 = []
    for _ in range(runs):
        t = time.perf_counter()
        refcpu.evaluate(arrays, counts, ext_one, one, kernel=1, threads=1, resolved=False)
        times1.append(time.perf_counter() - t)
    dt_one = float(np.median(times1))
    exact = bool(np.array_equal(tot.view(np.uint64), gpu_total[:sample].view(np.uint64)))
    return {"value": sample / dt_all, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"the first {sample} of the benchmark's points, f64, oracle/refcpu.cpp plan interpreter "
                      f"(OpenMP, {threads} threads, median of {runs} runs, {dt_all:.2f} s each); single thread "
                      f"{one / dt_one:.4g} points/s on {one} points",
            "single_thread_points_per_s": one / dt_one, "host_cpus_visible": os.cpu_count(),
            "f64_bitexact_on_sample": exact}


def mode_errors(got, ref):
    rel = np.abs(got - ref) / np.abs(ref)
    return {"max_rel_err_vs_f64": float(np.max(rel)), "median_rel_err_vs_f64": float(np.median(rel)),
            "max_err_rel_to_max_f64": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))}


def kernel_name(shape):
    """the f64 kernel a launch shape (waves_per_cu, block, unit_waves) of a plan without symjit
    rows runs, as rocprofv3 names its template: the one-wavefront kernel by its register tier (2
    up to 8 wavefronts per CU, 1 up to 12, else 0), the cooperative kernel by its shape index,
    wavefronts per unit and units per CU"""
    from pacuda_amd.runtime import COOP_SHAPES
    wpc, _, uw = shape
    if uw:
        k = COOP_SHAPES.index((uw, wpc or 16))
        return f"coop_kernel<KF64, {k}, false, {uw}, {(wpc or 16) // uw}>"
    tier = 2 if (wpc or 12) <= 8 else 1 if (wpc or 12) <= 12 else 0
    return f"fused_kernel<KF64, {tier}, false>"


def measured_traffic(kernel, points, rows, launch):
    """HBM bytes per dispatch from the committed rocprofv3 counters (profiles/traffic.json),
    scaled by the copy-kernel calibration, when they were taken on this workload and launch;
    else None"""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not t or (t["points"], t["plan_rows"], t["launch"]) != (points, rows, launch):
        return None
    return 1024.0 * (t["fetch_size_kib"] * t["fetch_scale"] + t["write_size_kib"] * t["write_scale"])


def slot_summary(lib, plan, shape):
    """what the launch's program keeps on chip and moves through its scratch slot per point. The
    f64 default launch (12 wavefronts per CU: register tier 1 with six register forms, 13 LDS
    planes, sources made again): the slot's current planes (its live peak) and the program's slot
    accesses. A cooperative launch: its phases, LDS and slot planes per unit, and the current
    components read or written in LDS and in the slot per point."""
    from pacuda_amd.program_stats import slot_traffic
    from pacuda_amd.runtime import plan_program, coop_program, COOP_SHAPES
    wpc, _, uw = shape
    if uw:
        k = COOP_SHAPES.index((uw, wpc or 16))
        _, info, _ = coop_program(lib, plan, k)
        slot_planes = info["slot_planes"] + plan.amplitude_planes + 2 * len(plan.forms)
        return {"program": f"cooperative, {uw} wavefronts per unit, {(wpc or 16) // uw} units per CU",
                "phases": info["phases"], "lds_planes": info["lds_planes"],
                "slot_current_planes": info["slot_planes"], "slot_kib_per_workgroup": slot_planes * 64 * 16 / 1024,
                "lds_current_components_per_point": info["lds_accesses"],
                "slot_current_components_per_point": info["slot_accesses"],
                "slot_current_bytes_per_point": 16 * info["slot_accesses"]}
    prog, sources, planes, lds = plan_program(lib, plan, 13)
    st = slot_traffic(prog, sources, register_forms=6)
    return {"program": "f64, 13 LDS planes, 6 register forms, sources made again", "slot_planes": planes,
            "slot_kib_per_wavefront": planes * 64 * 16 / 1024, "on_chip_planes": lds,
            "slot_bytes_per_point": st["bytes"], **{k: v for k, v in st.items() if k != "bytes"}}


def v1_abi(lib, plan, ext_host, want, tiles=(1 << 18, 1 << 20, 1 << 22), repeats=3):
    """the reference's calling pattern through the v1 entry points (umami_pacuda.c's batched
    evaluation): per tile of points, the external momenta from pinned host memory,
    pacman_recurrence_v1_run with the forms built on the device and the amplitudes kept (cleared
    on the device, none downloaded), then the reduction into pinned host memory: every helicity
    output from zero (pacman_recurrence_v1_reduce_resolved, 8 B per output and point back) or the
    helicity-sampled totals (pacman_recurrence_v1_reduce_sample, 12 B per point back). Wall time
    over all points of the best of `repeats` passes after one warm-up pass: host-to-device and
    device-to-host copies included, as the entry points take and return host memory. The totals
    are compared bit for bit with the fused f64 launch's. The keyed calls run on the fused
    kernels; the largest tile is measured again on the v1 kernels."""
    C = abi.C
    opt = abi.Options(abi.ABI, C.sizeof(abi.Options), abi.KERNEL_F64, 0, 0, 0)
    if lib.pacman_recurrence_v1_configure(C.byref(opt)) != 0:
        raise RuntimeError("pacman_recurrence_v1_configure refused the f64 kernel")
    a = plan.arrays()
    n = ext_host.shape[1]
    rows = ext_host.shape[0]
    outputs = len(a["offsets"]) - 1
    n_amp = plan.amplitude_planes
    amp_runs = np.array([0, n_amp], dtype=np.uint32)
    helicity = np.arange(outputs, dtype=np.uint32)
    rng = np.random.default_rng(7)

    def pinned(count, dtype=np.float64):
        p = lib.pacman_recurrence_v1_host_alloc(count * np.dtype(dtype).itemsize)
        if not p:
            raise RuntimeError("pacman_recurrence_v1_host_alloc failed")
        buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(p)
        return p, np.frombuffer(buf, dtype=dtype)

    res = {"workload": f"{plan.name}, {n} points, f64, KEEP_AMPLITUDES | FORMS_ON_DEVICE", "tiles": {}}
    # the tiles on the fused kernels, then the largest again on the v1 kernels (PACUDA_V1_FUSED=0)
    for tile in list(tiles) + ["v1_kernels"]:
        v1_kernels = tile == "v1_kernels"
        tile = min(tiles[-1] if v1_kernels else tile, n)
        starts = list(range(0, n, tile))
        # the tiles' momenta staged in pinned memory before the timed passes, each tile's planes of
        # `points` values (the layout the run reads)
        p_ext, ext = pinned(rows * n)
        p_out, out = pinned(outputs * tile)
        p_rand, rand = pinned(tile)
        p_tot, tot = pinned(n)
        p_cho, chosen = pinned(tile, np.uint32)
        hold = []
        try:
            if v1_kernels:
                os.environ["PACUDA_V1_FUSED"] = "0"
            for s in starts:
                pts = min(tile, n - s)
                ext[rows * s:rows * (s + pts)] = ext_host[:, s:s + pts].reshape(-1)
            rand[:] = rng.random(tile)
            calls = []
            for s in starts:
                pts = min(tile, n - s)
                c = abi.Call()
                c.abi_version, c.struct_size = abi.ABI, C.sizeof(abi.Call)
                c.words, c.words_key, c.word_count = abi.ptr(a["words"], abi.u64p), 0x6767676767, len(a["words"])
                for k in range(8):
                    c.width[k] = int(a["width"][k])
                c.scale_re, c.scale_im = abi.ptr(a["scale_re"]), abi.ptr(a["scale_im"])
                c.factor_re, c.factor_im = abi.ptr(a["factor_re"]), abi.ptr(a["factor_im"])
                c.current_planes, c.amplitude_planes = plan.current_planes, n_amp
                c.momentum_planes = 4 * len(plan.forms)
                c.scale_count, c.factor_count = len(a["scale_re"]), len(a["factor_re"])
                c.stride, c.points = pts, pts
                c.zero_amplitude, c.zero_amplitude_runs = abi.ptr(amp_runs, abi.u32p), 1
                c.flags = abi.KEEP_AMPLITUDES | abi.FORMS_ON_DEVICE
                c.external_momenta = C.cast(p_ext + rows * s * 8, abi.f64p)
                c.external_count = plan.legs
                c.term_count = len(a["term_slots"])
                c.form_terms = abi.ptr(a["form_terms"], abi.u32p)
                c.term_slots = abi.ptr(a["term_slots"], abi.u32p)
                c.term_coefficients = abi.ptr(a["term_coefficients"])
                resolved = abi.ReduceResolved(abi.ABI, C.sizeof(abi.ReduceResolved), abi.ptr(a["planes"], abi.u32p),
                                              abi.ptr(a["weights"]), abi.ptr(a["offsets"], abi.u32p),
                                              C.cast(p_out, abi.f64p), outputs, pts, 1, outputs, abi.OUTPUT_ZERO)
                sample = abi.ReduceSample(abi.ABI, C.sizeof(abi.ReduceSample), abi.ptr(a["planes"], abi.u32p),
                                          abi.ptr(a["weights"]), abi.ptr(a["offsets"], abi.u32p),
                                          abi.ptr(helicity, abi.u32p), outputs, pts, C.cast(p_rand, abi.f64p), None,
                                          float(plan.numerator), float(plan.denominator),
                                          C.cast(p_tot + s * 8, abi.f64p), C.cast(p_cho, abi.u32p))
                calls.append((s, pts, c, resolved, sample))
            hold.append(calls)

            def pass_resolved(check=False):
                exact = True
                for s, pts, c, r, _ in calls:
                    if lib.pacman_recurrence_v1_run(C.byref(c)) != 0 or lib.pacman_recurrence_v1_reduce_resolved(
                            C.byref(r)) != 0:
                        raise RuntimeError("v1 run/reduce_resolved failed")
                    if check:
                        v = out[:pts * outputs].reshape(pts, outputs)
                        t = np.zeros(pts)
                        for o in range(outputs):
                            x = v[:, o]
                            if plan.numerator != plan.denominator:
                                x = (x * plan.numerator) / plan.denominator
                            t = t + x
                        exact &= bool(np.array_equal(t.view(np.uint64), want[s:s + pts].view(np.uint64)))
                return exact

            def pass_sample():
                for s, pts, c, _, smp in calls:
                    if lib.pacman_recurrence_v1_run(C.byref(c)) != 0 or lib.pacman_recurrence_v1_reduce_sample(
                            C.byref(smp)) != 0:
                        raise RuntimeError("v1 run/reduce_sample failed")

            entry = {"calls_per_pass": len(calls)}
            entry["resolved_bitexact_vs_fused"] = pass_resolved(check=True)
            best = min(_wall(pass_resolved) for _ in range(repeats))
            entry["resolved_points_per_s"] = n / best
            pass_sample()
            entry["sample_totals_bitexact_vs_fused"] = bool(np.array_equal(tot.view(np.uint64), want.view(np.uint64)))
            best = min(_wall(pass_sample) for _ in range(repeats))
            entry["sample_points_per_s"] = n / best
            if v1_kernels:
                res["v1_kernels"] = dict(entry, tile=tile)
            else:
                res["tiles"][str(tile)] = entry
            log(f"v1 ABI tile {tile}{' (v1 kernels)' if v1_kernels else ''}: resolved "
                f"{entry['resolved_points_per_s']:.4g} points/s, sample {entry['sample_points_per_s']:.4g} points/s "
                f"(bit-exact {entry['resolved_bitexact_vs_fused']}, {entry['sample_totals_bitexact_vs_fused']})")
        finally:
            os.environ.pop("PACUDA_V1_FUSED", None)
            for p in (p_ext, p_out, p_rand, p_tot, p_cho):
                lib.pacman_recurrence_v1_host_free(p)
    return res


def _wall(f):
    t = time.perf_counter()
    f()
    return time.perf_counter() - t


def unweighting_summary(u):
    """the unweighting figures of the JSON line (None when the unweighting did not run)"""
    if u is None:
        return None
    t = u["two_stage"]
    return {"plain": u["plain"]["events_per_s"], "surrogate": t["events_per_s"], "gain": t["gain"],
            "sigma_pb": {"plain": u["plain"]["estimate"], "surrogate": t["estimate"],
                         "weighted": u["plain"]["weighted"]},
            "sigma_pull_vs_plain": t["sigma_pull"], "sigma_pull_vs_weighted": t["sigma_pull_vs_weighted"],
            "gain_network_on_every_candidate": u["two_stage_dense"]["gain"],
            "gain_prior_only": u["two_stage_prior_only"]["gain"],
            "training_seconds": u["training"]["label_seconds"] + u["training"]["train_seconds"],
            # events beyond which two-stage unweighting, its one-off labelling and training included,
            # takes less time than plain unweighting
            "break_even_events": (u["training"]["label_seconds"] + u["training"]["train_seconds"]) /
                                 (1.0 / u["plain"]["events_per_s"] - 1.0 / t["events_per_s"])
                                 if t["events_per_s"] > u["plain"]["events_per_s"] else None}


def copy_calibration(rows=20, count=1 << 22, repeats=3):
    """pacuda_gather with the identity index copies rows x count doubles: it reads
    rows*count*8 + count*4 bytes and writes rows*count*8 (gather_kernel, one dispatch each)"""
    dev = Device()
    lib = dev.lib
    src, dst = dev.empty(rows * count), dev.empty(rows * count)
    idx = dev.array(np.arange(count, dtype=np.uint32))
    for _ in range(repeats):
        if lib.pacuda_gather(src.ptr, count, rows, idx.ptr, count, dst.ptr, None, 0, None) != 0:
            raise RuntimeError("pacuda_gather failed")
    dev.synchronize()
    print(json.dumps({"copy_kernel": "gather_kernel", "bytes_read": rows * count * 8 + count * 4,
                      "bytes_written": rows * count * 8, "dispatches": repeats}), flush=True)
    for a in (src, dst, idx):
        a.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1 << 22)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--launch", default="auto",
                    help="the f64 headline launch: auto (the fastest of the one-wavefront-per-unit default and the "
                         "cooperative shapes, chosen before the warm-up), fused (one wavefront per unit, "
                         "--waves-per-cu), or coop:U:W (U wavefronts per unit, W per CU)")
    ap.add_argument("--stream", choices=("counter", "record"), default="counter",
                    help="headline inputs: the GPU's counter-based RAMBO stream, or the reference's own "
                         "xorshift stream (convert/src/record.rs) generated on the host")
    ap.add_argument("--cpu-sample", type=int, default=1 << 22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-modes", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--only-unweighting", action="store_true", help="of the extra workloads, only the unweighting")
    ap.add_argument("--only-v1", action="store_true", help="of the extra workloads, only the v1 entry points")
    ap.add_argument("--only-gggg", action="store_true", help="of the extra workloads, only g g > g g g g")
    ap.add_argument("--candidates", type=int, default=1 << 24)
    ap.add_argument("--train-points", type=int, default=1 << 20)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--copy-only", action="store_true",
                    help="only a plain copy of known size (pacuda_gather, identity index), to scale HBM counters")
    args = ap.parse_args()
    if args.copy_only:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "points with no exchange during evaluation)", file=sys.stderr)
        sys.exit(2)

    dev = Device()
    lib = dev.lib
    plan = plans.gluon_plan(5)
    n = args.points
    t0 = time.perf_counter()
    if args.stream == "record":
        d_ext, drawn = points_from_record_stream(dev, lib, plan, n, seed=2)
        log(f"{n} points of the record.rs stream generated on the host in {time.perf_counter() - t0:.2f} s "
            f"({drawn} drawn)")
    else:
        d_ext, drawn = points_on_device(dev, plan, n, seed=2)
        log(f"{n} points generated on the GPU in {time.perf_counter() - t0:.2f} s ({drawn} drawn)")
    fp = FusedPlan(lib, plan)
    d_tot = dev.empty(n)
    flops, divs = plans.plan_flops(plan)
    jobs = [(fp, d_ext, n, n, None, d_tot)]
    k64 = abi.KERNEL_F64
    tuned = None
    one_wave = (args.waves_per_cu or 12, 256)  # (the one-wavefront-per-unit launch of the f64 kernel)
    if args.launch == "fused":
        shape = one_wave + (0,)
    elif args.launch.startswith("coop:"):
        _, u, w = args.launch.split(":")
        shape = (int(w), 0, int(u))
    else:
        shape, tuned = autotune(lib, jobs, k64, shapes=(one_wave,) + COOP_TUNE_SHAPES)
        shape = tuple(shape) + (0,) * (3 - len(shape))
        log(f"f64 launch: {shape_name(shape)} ({', '.join(f'{k} {v:.2f} ms' for k, v in tuned.items())})")
    wpc, _, uw = shape
    for _ in range(args.warmup):
        evaluate(lib, jobs, k64, waves_per_cu=wpc, timed=True, unit_waves=uw)
    wall, kernel_ms = timed_steps(lib, jobs, k64, args.steps, wpc, 0, uw)
    total64 = d_tot.download(n)
    value = n * args.steps / wall
    achieved = flops * n / (kernel_ms * 1e-3) / 1e12
    log(f"f64 {value:.4g} points/s, kernel {kernel_ms:.3f} ms, {achieved:.2f} TFLOP/s")

    modes = {}
    if not args.no_modes:
        for name in ("df64", "mixed", "mixed64", "compensated", "compensated64"):
            k = abi.FUSED_KERNELS[name]
            mshape, _ = autotune(lib, jobs, k, shapes=((16, 256),) + COOP_TUNE_SHAPES_ANY, repeats=1)
            _, ms = timed_steps(lib, jobs, k, 3, *(tuple(mshape) + (0,))[:3])
            modes[name] = {"points_per_s": n / (ms * 1e-3), "kernel_ms": ms, "launch": shape_name(mshape),
                           **mode_errors(d_tot.download(n), total64)}
            log(f"{name} {modes[name]['points_per_s']:.4g} points/s ({shape_name(mshape)})")

    extra = {}
    only = args.only_unweighting or args.only_v1 or args.only_gggg
    if not args.no_extra and (not only or args.only_v1):
        # the same points through the reference-compatible entry points, host memory in and out
        extra["v1_abi"] = v1_abi(lib, plan, d_ext.download((plan.legs * 4, n)), total64)
    if not args.no_extra and not only:
        # the whole leading-colour sum of g g > g g g: the 12 orderings (0, s1..s4) of the legs up
        # to reflection, each the same plan on its permutation of the external legs, in one launch
        orders = [(0,) + p for p in itertools.permutations(range(1, plan.legs)) if p[0] < p[-1]]
        ident = dev.array(np.arange(n, dtype=np.uint32))
        lc_jobs, lc_hold = [], []
        for order in orders:
            e = dev.empty(plan.legs * 4 * n)
            for leg, src in enumerate(order):
                st = lib.pacuda_gather(d_ext.ptr + src * 4 * n * 8, n, 4, ident.ptr, n, e.ptr + leg * 4 * n * 8,
                                       None, 0, None)
                assert st == 0
            t = dev.empty(n)
            lc_jobs.append((fp, e, n, n, None, t))
            lc_hold += [e, t]
        ident.free()
        lshape, _ = autotune(lib, lc_jobs, k64, shapes=TUNE_SHAPES + COOP_TUNE_SHAPES, key="lc5")
        _, ms_lc = timed_steps(lib, lc_jobs, k64, 3, *(tuple(lshape) + (0,))[:3])
        extra["gg_ggg_lc_sum"] = {"orderings": len(orders), "points": n, "tuned": shape_name(lshape),
                                  "kernel_ms": ms_lc, "points_per_s": n / (ms_lc * 1e-3),
                                  "plan_runs_per_s": len(orders) * n / (ms_lc * 1e-3)}
        log(f"gg>ggg LC sum over {len(orders)} orderings: {n / (ms_lc * 1e-3):.4g} points/s")
        for a in lc_hold:
            a.free()

    if not args.no_extra and (not only or args.only_gggg):
        # g g > g g g g: 2^21 points, the persistent grid swept (waves per CU, block)
        p6 = plans.gluon_plan(6)
        n6 = 1 << 21
        e6, _ = points_on_device(dev, p6, n6, seed=5)
        f6 = FusedPlan(lib, p6)
        t6 = dev.empty(n6)
        j6 = [(f6, e6, n6, n6, None, t6)]
        shape6, sweep = autotune(lib, j6, k64, shapes=TUNE_SHAPES + COOP_TUNE_SHAPES)
        best = shape_name(shape6)
        _, ms64 = timed_steps(lib, j6, k64, 3, *(tuple(shape6) + (0,))[:3])
        ref6 = t6.download(n6)
        f6flops = plans.plan_flops(p6)[0]
        _, ms_default = timed_steps(lib, j6, k64, 3)  # (the library's default launch)
        res6 = {"points": n6, "plan_rows": len(p6.rows), "flops_per_point": f6flops, "tuned": best, "sweep_ms": sweep,
                "f64_points_per_s": n6 / (ms64 * 1e-3),
                "f64_tflops": f6flops * n6 / (ms64 * 1e-3) / 1e12,
                "f64_default_launch_points_per_s": n6 / (ms_default * 1e-3)}
        for name in ("mixed", "mixed64", "compensated", "compensated64", "df64"):
            k = abi.FUSED_KERNELS[name]
            kshape, _ = autotune(lib, j6, k, shapes=TUNE_SHAPES + COOP_TUNE_SHAPES_ANY)
            _, ms = timed_steps(lib, j6, k, 3, *(tuple(kshape) + (0,))[:3])
            res6[name] = {"points_per_s": n6 / (ms * 1e-3), "tuned": shape_name(kshape),
                          **mode_errors(t6.download(n6), ref6)}
        extra["gg_gggg"] = res6
        log(f"gg>gggg f64 {res6['f64_points_per_s']:.4g} points/s tuned {best}")
        for d in (e6, t6):
            d.free()
        f6.destroy()

    if not args.no_extra and not only:
        # p p > j j j: every subprocess, Zipf(1.2) counts summing to 2^22, one persistent launch
        subs = processes.subprocesses()
        counts = processes.zipf_counts(len(subs), 1 << 22, s=1.2, seed=1)
        pjobs, holds = [], []
        for k, (sp, c) in enumerate(zip(subs, counts)):
            e, _ = points_on_device(dev, sp.plan, int(c), seed=100 + k)
            f = FusedPlan(lib, sp.plan)
            t = dev.empty(int(c))
            pjobs.append((f, e, int(c), int(c), None, t))
            holds.append((f, e, t))
        pshape, _ = autotune(lib, pjobs, k64, shapes=((12, 256),) + COOP_TUNE_SHAPES_ANY, repeats=1)
        _, ms_p = timed_steps(lib, pjobs, k64, 3, *(tuple(pshape) + (0,))[:3])
        refs = [j[5].download(j[2]) for j in pjobs]
        pp = {"subprocesses": len(subs), "points": int(counts.sum()), "f64_points_per_s": int(counts.sum()) / (ms_p * 1e-3),
              "f64_launch": shape_name(pshape), "counts_min_max": [int(counts.min()), int(counts.max())]}
        for name in ("mixed", "mixed64", "compensated", "compensated64"):
            k = abi.FUSED_KERNELS[name]
            kshape, _ = autotune(lib, pjobs, k, shapes=((16, 256),) + COOP_TUNE_SHAPES_ANY, repeats=1)
            _, ms = timed_steps(lib, pjobs, k, 3, *(tuple(kshape) + (0,))[:3])
            per = {sp.name: mode_errors(j[5].download(j[2]), r) for sp, j, r in zip(subs, pjobs, refs)}
            errs = {n: e["max_err_rel_to_max_f64"] for n, e in per.items()}
            pp[name] = {"points_per_s": int(counts.sum()) / (ms * 1e-3), "launch": shape_name(kshape),
                        "worst_subprocess_err": max(errs.values()),
                        "worst_subprocess_median_rel_err": max(e["median_rel_err_vs_f64"] for e in per.values()),
                        "err_rel_to_max_f64_per_subprocess": errs}
        extra["pp_jjj"] = pp
        log(f"pp>jjj f64 {pp['f64_points_per_s']:.4g} points/s over {len(subs)} subprocesses")
        for f, e, t in holds:
            e.free()
            t.free()
            f.destroy()

    if not args.no_extra and (not only or args.only_unweighting):
        # surrogate vs plain unweighting of g g > g g g g
        t = time.perf_counter()
        uw = Unweighting(legs=6, hidden=256, layers=4, batch=args.train_batch, prior=True, lib=lib)
        x, logm2, logp = uw.training_set(args.train_points, seed=3)
        label_s = time.perf_counter() - t
        t = time.perf_counter()
        hist = uw.train(x, logm2, logp, epochs=args.epochs, lr=args.lr)
        train_s = time.perf_counter() - t
        log(f"surrogate: {args.train_points} labels in {label_s:.1f} s, trained in {train_s:.1f} s, "
            f"loss {hist[0]:.3f} -> {hist[-1]:.3f}")
        res = uw.run(args.candidates, seed=4, variants=("network", "network_dense", "prior"))
        res["training"] = {"points": args.train_points, "seed": 3, "epochs": args.epochs, "batch": args.train_batch,
                           "label_seconds": label_s, "train_seconds": train_s,
                           "loss_first_last": [hist[0], hist[-1]],
                           "rms_log_residual_before": uw.std,
                           "rms_log_error": float(np.sqrt(hist[-1]) * uw.std)}
        extra["unweighting"] = res
        for k in ("two_stage", "two_stage_dense", "two_stage_prior_only"):
            r = res[k]
            log(f"unweighting {k}: plain {res['plain']['events_per_s']:.4g} ev/s, {r['events_per_s']:.4g} ev/s, "
                f"gain {r['gain']:.2f}, pull {r['sigma_pull']:.2f} (vs weighted {r['sigma_pull_vs_weighted']:.2f})")
        uw.destroy()

    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        ext_host = d_ext.download((plan.legs * 4, n))[:, :sample].copy()
        cpu = cpu_baseline(plan, ext_host, total64, sample)
        log(f"cpu {cpu['value']:.4g} points/s")

    kname = kernel_name(shape)
    traffic = measured_traffic(kname, n, len(plan.rows), shape_name(shape if uw else (wpc or 12, 256)))
    slot = slot_summary(lib, plan, shape)
    alg_bytes = plan.legs * 4 * 8 + 8  # the momenta in, |M|^2 out, per point
    line = {
        "metric": METRIC, "value": value, "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": ("synthetic (RAMBO, the record.rs xorshift stream on the host, seed 2)" if args.stream == "record"
                 else "synthetic (RAMBO on the GPU, counter-based stream, seed 2)"),
        "config": {"workload": f"g g > g g g LC plan run (one colour ordering, {len(plan.outputs)} helicities summed), "
                               f"f64, {n} RAMBO points, sqrt(s)=1 TeV, pT>20 GeV, |eta|<2.5, dR>0.4",
                   "points": n, "plan_rows": len(plan.rows), "current_planes": plan.current_planes,
                   "flops_per_point": flops, "divides_per_point": divs, "parallelism": "1 GPU, persistent fused kernel",
                   "launch": shape_name(shape if uw else (wpc or 12, 256)), "launch_sweep_ms": tuned},
        # The f64 kernel by its algorithm: 39,448 flops per point against 168 algorithmic bytes
        # (momenta in, |M|^2 out), far above the HBM ridge (78.6 TFLOP/s / 8 TB/s), so the
        # roofline is the FP64 one (the VALU's: the kernel issues no MFMA; MI355X's FP64 vector
        # and matrix peaks are both 78.6 TFLOP/s). Its measured limiters are stated beside it:
        # the HBM traffic of its scratch slots (rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/) and
        # the FMA-free ceiling of its rows.
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_unit": "HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "flops_per_point": flops,
                     # the program rows (the reference's operation order) contain no fused
                     # multiply-adds, so they issue one FP64 instruction per flop: half the
                     # FMA-counted peak is their ceiling; only the reduction uses an fma
                     "fma_free_ceiling_frac": 0.5,
                     "kernel": kname, "kernel_ms": kernel_ms,
                     "hbm": {"achieved": alg_bytes * n / (kernel_ms * 1e-3) / 1e9, "peak": HBM_PEAK_TBS * 1e3,
                             "unit": "GB/s", "frac": alg_bytes * n / (kernel_ms * 1e-3) / 1e9 / (HBM_PEAK_TBS * 1e3),
                             "algorithmic_bytes_per_point": alg_bytes, "algorithmic_bytes_per_dispatch": alg_bytes * n,
                             "traffic_tb_per_s": traffic / (kernel_ms * 1e-3) / 1e12 if traffic else None,
                             "traffic_frac_of_hbm_peak":
                                 traffic / (kernel_ms * 1e-3) / 1e12 / HBM_PEAK_TBS if traffic else None},
                     "slot": slot},
        "cpu_baseline": cpu,
        "modes": modes,
        "unweighted_events_per_s": unweighting_summary(extra.get("unweighting")),
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    d_ext.free()
    d_tot.free()
    fp.destroy()


if __name__ == "__main__":
    main()
