# SPDX-License-Identifier: 0BSD
"""The 5-parton subprocesses of p p -> j j j over g, u, d, s, c and their antiquarks, each as a
plan (one leading-colour ordering, every helicity) and a crossing (which physical leg each plan
leg reads, with its sign).

The runner's op vocabulary (runner/recurrence_runner.cu:15-36) has no vertex that makes a vector
current from a fermion pair, so subprocesses with two quark lines (q q' -> q q' g and the like)
need symjit rows (ops 13/14), which this build does not carry; the list holds every subprocess
with at most one quark line, in a fixed order:

  g g  -> g g g
  g g  -> q qbar g       q in u d s c
  q g  -> q g g, g q -> q g g, qbar g -> qbar g g, g qbar -> qbar g g
  q qbar -> g g g, qbar q -> g g g

Physical legs: 0, 1 the beams (incoming), 2, 3, 4 the final state. Plan legs of a quark line:
0 the antiquark slot, 1..3 gluons, 4 the quark slot (all-outgoing convention), so an incoming
quark sits in the antiquark slot with sign -1 and an incoming antiquark in the quark slot."""
import numpy as np

from . import plans

QUARKS = ["u", "d", "s", "c"]


class Subprocess:
    def __init__(self, name, plan, crossing):
        self.name = name
        self.plan = plan
        self.crossing = crossing  # per plan leg: (physical leg, sign)
        plan.crossing = crossing


def _quark_line(name, slots):
    """slots: plan leg -> physical leg for (antiquark slot, g, g, g, quark slot)"""
    incoming = tuple(k for k, phys in enumerate(slots) if phys in (0, 1))
    plan = plans.quark_line_plan(5, incoming=incoming)
    crossing = [(phys, -1.0 if phys in (0, 1) else 1.0) for phys in slots]
    return Subprocess(name, plan, crossing)


def subprocesses():
    out = []
    g = plans.gluon_plan(5)
    out.append(Subprocess("g g > g g g", g, [(k, -1.0 if k in (0, 1) else 1.0) for k in range(5)]))
    for q in QUARKS:
        # g g > q qbar g: the outgoing quark (phys 2) in the quark slot, the antiquark (phys 3)
        # in the antiquark slot, gluons from the beams and phys 4
        out.append(_quark_line(f"g g > {q} {q}~ g", [3, 0, 1, 4, 2]))
    for q in QUARKS:
        out.append(_quark_line(f"{q} g > {q} g g", [0, 1, 3, 4, 2]))
        out.append(_quark_line(f"g {q} > {q} g g", [1, 0, 3, 4, 2]))
        out.append(_quark_line(f"{q}~ g > {q}~ g g", [2, 1, 3, 4, 0]))
        out.append(_quark_line(f"g {q}~ > {q}~ g g", [2, 0, 3, 4, 1]))
    for q in QUARKS:
        out.append(_quark_line(f"{q} {q}~ > g g g", [0, 2, 3, 4, 1]))
        out.append(_quark_line(f"{q}~ {q} > g g g", [1, 2, 3, 4, 0]))
    return out


def zipf_counts(n_items, total, s=1.2, seed=1):
    """point counts per item drawn Zipf(s) over a seeded random ranking, summing to `total`
    (each at least 1)"""
    rng = np.random.default_rng(seed)
    ranks = rng.permutation(n_items) + 1
    w = 1.0 / ranks.astype(np.float64) ** s
    counts = np.floor(w / w.sum() * (total - n_items)).astype(np.int64) + 1
    counts[np.argmax(w)] += total - counts.sum()
    return counts
