# SPDX-License-Identifier: 0BSD
"""Benchmark of the recurrence runner on one MI355X.

Workload (the headline configuration): g g > g g g, leading colour, one colour ordering with all
32 helicity configurations, f64 (bit-exact to the CPU), 2^22 RAMBO points (seed 2) at
sqrt(s) = 1 TeV with pT > 20 GeV, |eta| < 2.5, dR > 0.4. A step is one pass of the hot path over
the batch: the fused evaluation (forms, sources, the plan's program, the helicity reduction and
sum) of every point, momenta already in HBM, |M|^2 per point back into HBM.

Prints one JSON line: value = points/s; roofline of the fused f64 kernel (its time from HIP events
recorded by the library on the kernel's stream); the CPU oracle (the same plan interpreter,
OpenMP) timed on a bounded sample of the same points on this host; the other precision modes'
throughput and error against f64."""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pacuda_amd import abi, plans, phase_space as ps, runtime  # noqa: E402

METRIC = "|M|^2 points/s (f64 bit-exact) + unweighted events/s, surrogate vs plain"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed)


def cpu_baseline(plan, ext, gpu_total, sample):
    """the oracle over the first `sample` points, all host threads and one thread"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    counts = dict(current_planes=plan.current_planes, amplitude_planes=plan.amplitude_planes,
                  form_count=len(plan.forms), external_count=plan.legs)
    arrays = plan.arrays()
    threads = max(1, min(16, os.cpu_count() or 1))
    sub = np.ascontiguousarray(ext[:, :sample])
    refcpu.evaluate(arrays, counts, sub[:, :4096], 4096, kernel=1, threads=threads, resolved=False)  # warm-up
    t = time.perf_counter()
    _, tot = refcpu.evaluate(arrays, counts, sub, sample, kernel=1, threads=threads, resolved=False)
    dt_all = time.perf_counter() - t
    one = min(sample, 1 << 15)
    t = time.perf_counter()
    refcpu.evaluate(arrays, counts, np.ascontiguousarray(sub[:, :one]), one, kernel=1, threads=1, resolved=False)
    dt_one = time.perf_counter() - t
    exact = bool(np.array_equal(tot.view(np.uint64), gpu_total[:sample].view(np.uint64)))
    return {"value": sample / dt_all, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} of the same points, f64, oracle/refcpu.cpp (OpenMP, {threads} threads); "
                      f"single thread {one / dt_one:.4g} points/s",
            "single_thread_points_per_s": one / dt_one, "f64_bitexact_on_sample": exact}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1 << 22)
    ap.add_argument("--legs", type=int, default=5)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-modes", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "points with no exchange during evaluation, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    plan = plans.gluon_plan(args.legs)
    n = args.points
    t0 = time.perf_counter()
    mom, drawn = ps.generate(args.legs, n, seed=2)
    ext = ps.external_planes(mom, plan.crossing)
    gen_s = time.perf_counter() - t0
    print(f"bench: {n} points generated in {gen_s:.1f} s ({drawn} drawn)", file=sys.stderr, flush=True)

    lib = abi.load()
    dev = runtime.Device(lib)
    fp = runtime.FusedPlan(lib, plan)
    d_ext = dev.array(ext)
    d_tot = dev.empty(n)
    flops, divs = plans.plan_flops(plan)

    def step(kernel, timed=True):
        return runtime.evaluate(lib, [(fp, d_ext, n, n, None, d_tot)], kernel, waves_per_cu=args.waves_per_cu,
                                timed=timed)

    k64 = abi.KERNEL_F64
    for _ in range(args.warmup):
        step(k64)
    dev.synchronize()
    ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        ms.append(step(k64))
    dev.synchronize()
    wall = time.perf_counter() - t
    total64 = d_tot.download(n)
    value = n * args.steps / wall
    kernel_ms = float(np.mean(ms))
    achieved = flops * n / (kernel_ms * 1e-3) / 1e12
    print(f"bench: f64 {value:.4g} points/s, kernel {kernel_ms:.3f} ms", file=sys.stderr, flush=True)

    modes = {}
    if not args.no_modes:
        ref = total64
        for name in ("df64", "mixed", "mixed64"):
            k = abi.KERNELS[name]
            step(k)
            times = [step(k) for _ in range(3)]
            got = d_tot.download(n)
            rel = np.abs(got - ref) / np.abs(ref)
            modes[name] = {"points_per_s": n / (np.mean(times) * 1e-3), "kernel_ms": float(np.mean(times)),
                           "max_rel_err_vs_f64": float(np.max(rel)), "median_rel_err_vs_f64": float(np.median(rel)),
                           "max_err_rel_to_max_f64": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))}
            print(f"bench: {name} {modes[name]['points_per_s']:.4g} points/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(plan, ext, total64, min(args.cpu_sample, n))
        print(f"bench: cpu {cpu['value']:.4g} points/s", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (RAMBO, seed 2)",
        "config": {"workload": f"g g > g g g LC plan run (one colour ordering, {len(plan.outputs)} helicities summed), "
                               f"f64, {n} RAMBO points, sqrt(s)=1 TeV, pT>20 GeV, |eta|<2.5, dR>0.4",
                   "points": n, "plan_rows": len(plan.rows), "current_planes": plan.current_planes,
                   "flops_per_point": flops, "divides_per_point": divs, "parallelism": "1 GPU, persistent fused kernel"},
        "roofline": {"bound": "mfma", "pipe": "fp64 VALU (MI355X FP64 vector peak = matrix peak)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "fused_kernel<KF64>", "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_point": 20 * 8 + 8},
        "cpu_baseline": cpu,
        "modes": modes,
        "unweighting": None,
    }
    print(json.dumps(line), flush=True)
    fp.destroy()


if __name__ == "__main__":
    main()
