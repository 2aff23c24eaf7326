# SPDX-License-Identifier: 0BSD
"""Writes tests/golden/gg_ggg_qqbar.pacuda: a .pacuda v2 recording of two generated processes
(g g > g g g and u u~ > g g g, one colour ordering each, every helicity) with 8 reference points
each (record.rs's RAMBO restated, seeds 0x5eed0008 and 0x5eed0009) whose values are the CPU
oracle's f64 ones (group factor 1/4 on the first process). Run from the repository root:
python tests/golden/make_recording.py"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

from pacuda_amd import pacuda_file, plans, processes, phase_space as ps  # noqa: E402
import refcpu  # noqa: E402


def record(plan, key, seed, numerator=1, denominator=1):
    mom = ps.rambo_xorshift(plan.legs, 8, seed)  # [point][leg][4], the runtime's input order
    ext = ps.external_planes(mom, plan.crossing)
    counts = dict(current_planes=plan.current_planes, amplitude_planes=plan.amplitude_planes,
                  form_count=len(plan.forms), external_count=plan.legs)
    res, _ = refcpu.evaluate(plan.arrays(), counts, ext, 8, kernel=1)
    if numerator != denominator:
        res = res * float(numerator) / float(denominator)
    colour = pacuda_file.colour_record(plan, "1", 64, mom.reshape(-1), res.reshape(-1))
    return dict(key=key, legs=plan.legs, helicities=len(plan.outputs), alpha_s_name="aS", alpha_s=0.118,
                qcd_power=3, parameters=[("aS", 0.118, 0.0)], numerator=numerator, denominator=denominator,
                crossing=plan.crossing, colours=[colour])


def main():
    gg = plans.gluon_plan(5)
    qq = [sp for sp in processes.subprocesses() if sp.name == "u u~ > g g g"][0].plan
    procs = [record(gg, "g g > g g g", 0x5EED0008, 1, 4), record(qq, "u u~ > g g g", 0x5EED0009)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gg_ggg_qqbar.pacuda")
    with open(out, "wb") as f:
        f.write(pacuda_file.write(procs, kernel=1, batch=0))
    print(out)


if __name__ == "__main__":
    main()
