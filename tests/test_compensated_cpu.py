# SPDX-License-Identifier: 0BSD
"""The compensated mixed policies of the fused path (PACUDA_KERNEL_COMPENSATED, df64 arithmetic,
and PACUDA_KERNEL_COMPENSATED64, f64 arithmetic; the oracle's kernels 4 and 5; no counterpart in
the reference) pinned on the CPU by what their definition fixes.

Both keep the reference mixed policy's f32 currents (runner/recurrence_runner.cu:178-218) and do
the arithmetic of every row in a wider type from them, so that each current is rounded to f32
once, when it is stored. Checked on gluon plans and on p p > j j j subprocesses with quark lines
and with symjit rows:

- the two agree to df64's accuracy (about 2^-44 per operation) on the totals: they share the
  roundings and differ only in the arithmetic between them, so a slip in the df64 building
  blocks, a dropped widening (an f32 operation left in) or a constant read at f32 in either would
  move every point by an f32-sized amount;
- against f64 their error is that of the one rounding per current: below the reference's mixed
  policy's on the same points, in median by at least 40 % and at the 99th percentile, and not
  zero (the currents are rounded).

The GPU kernels equal the oracle's kernels 4 and 5 bit for bit (tests/test_gpu_parity.py,
tests/test_gpu_gluon6.py)."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, processes, phase_space as ps
import harness
import refcpu

COMP, COMP64, F64, MIXED = abi.KERNEL_COMPENSATED, abi.KERNEL_COMPENSATED64, abi.KERNEL_F64, abi.KERNEL_MIXED
POINTS = 1500


def _cases():
    subs = {sp.name: sp for sp in processes.subprocesses()}
    out = [("gg>gggg", plans.gluon_plan(6)), ("gg>ggg", plans.gluon_plan(5))]
    quark = [n for n, sp in subs.items() if sp.plan.name.startswith("qqbar+")][:1]
    four = [n for n, sp in subs.items() if sp.plan.name.startswith("qqbarQQbar")][:2]
    return out + [(n, subs[n].plan) for n in quark + four]


CASES = _cases()


def _totals(plan, kernels, seed=5):
    mom, _ = ps.generate(plan.legs, POINTS, seed=seed)
    ext = ps.external_planes(mom, plan.crossing)
    out = {}
    for k in kernels:
        _, out[k] = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, POINTS, kernel=k, threads=8,
                                    resolved=False)
    return out


@pytest.mark.parametrize("name,plan", CASES, ids=[c[0] for c in CASES])
def test_df64_and_f64_compensated_policies_agree(name, plan):
    t = _totals(plan, (COMP, COMP64))
    d = np.abs(t[COMP] / t[COMP64] - 1.0)
    print(name, "max |comp/comp64 - 1|", d.max(), flush=True)
    # an f32 rounding that falls the other way (a df64 result within 2^-44 of a rounding
    # boundary) moves a point by ~1e-7; at most a few points in a thousand may have one
    assert np.mean(d > 1e-12) < 0.005, name
    assert d.max() < 1e-5, name
    assert np.median(d) < 1e-14, name


@pytest.mark.parametrize("name,plan", CASES, ids=[c[0] for c in CASES])
def test_compensated_error_is_one_rounding_per_current(name, plan):
    t = _totals(plan, (F64, MIXED, COMP, COMP64))
    e = {k: np.abs(t[k] / t[F64] - 1.0) for k in (MIXED, COMP, COMP64)}
    med = {k: float(np.median(v)) for k, v in e.items()}
    p99 = {k: float(np.quantile(v, 0.99)) for k, v in e.items()}
    print(name, "median", med, "p99", p99, flush=True)
    for k in (COMP, COMP64):
        assert med[k] < 0.6 * med[MIXED], (name, k)
        assert p99[k] < p99[MIXED], (name, k)
        assert 1e-9 < med[k] < 1.5e-7, (name, k)
