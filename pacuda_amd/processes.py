# SPDX-License-Identifier: 0BSD
"""The 5-parton subprocesses of p p -> j j j over g, u, d, s, c and their antiquarks, each as a
plan (one leading-colour ordering, every helicity) and a crossing (which physical leg each plan
leg reads, with its sign).

Physical legs: 0, 1 the beams (incoming), 2, 3, 4 the final state. In the all-outgoing
convention of the plans an incoming quark is an outgoing antiquark of its line and the reverse.

One quark line (plans.quark_line_plan; plan legs 0 the antiquark slot, 1..3 gluons, 4 the quark
slot), in this order:
  g g  -> g g g
  g g  -> q qbar g       q in u d s c
  q g  -> q g g, g q -> q g g, qbar g -> qbar g g, g qbar -> qbar g g
  q qbar -> g g g, qbar q -> g g g
Two quark lines and one gluon (plans.four_quark_plan; plan legs 0 qbar_M, 1 the gluon, 2 q_S,
3 qbar_S, 4 q_M, lines M and S), then, for q, q' in u d s c (q' != q where named so):
  q q' -> q q' g, q q -> q q g
  q qbar' -> q qbar' g, qbar' q -> qbar' q g
  q qbar -> q' qbar' g, qbar q -> q' qbar' g, q qbar -> q qbar g, qbar q -> qbar q g
  qbar qbar' -> qbar qbar' g, qbar qbar -> qbar qbar g
  q g -> q q' qbar', g q -> q q' qbar', q g -> q q qbar, g q -> q q qbar
  qbar g -> qbar q' qbar', g qbar -> qbar q' qbar', qbar g -> qbar q qbar, g qbar -> qbar q qbar
Each takes one leading-colour ordering, the gluon between qbar_M and q_S; with identical flavours,
one of the two pairings of the quarks into lines (the one listed in its slots)."""
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


def _two_lines(name, slots):
    """slots: plan leg -> physical leg for (qbar_M, g, q_S, qbar_S, q_M)"""
    incoming = tuple(k for k, phys in enumerate(slots) if phys in (0, 1))
    plan = plans.four_quark_plan(incoming=incoming)
    crossing = [(phys, -1.0 if phys in (0, 1) else 1.0) for phys in slots]
    return Subprocess(name, plan, crossing)


def four_quark_subprocesses():
    """the subprocesses with two quark lines, in the order of the module's list"""
    out = []
    pairs = [(q, r) for q in QUARKS for r in QUARKS if r != q]
    # q q' -> q q' g: line q (phys 0 in, phys 2 out), line q' (phys 1 in, phys 3 out)
    for q, r in pairs:
        out.append(_two_lines(f"{q} {r} > {q} {r} g", [0, 4, 3, 1, 2]))
    for q in QUARKS:
        out.append(_two_lines(f"{q} {q} > {q} {q} g", [0, 4, 3, 1, 2]))
    # q qbar' -> q qbar' g: line q (0 in, 2 out), line q' (1: incoming antiquark = q slot, 3)
    for q, r in pairs:
        out.append(_two_lines(f"{q} {r}~ > {q} {r}~ g", [0, 4, 1, 3, 2]))
        out.append(_two_lines(f"{r}~ {q} > {r}~ {q} g", [1, 4, 0, 2, 3]))
    # q qbar -> q' qbar' g: line q (0, 1), line q' (2 out quark, 3 out antiquark)
    for q, r in pairs:
        out.append(_two_lines(f"{q} {q}~ > {r} {r}~ g", [0, 4, 2, 3, 1]))
        out.append(_two_lines(f"{q}~ {q} > {r} {r}~ g", [1, 4, 2, 3, 0]))
    for q in QUARKS:
        out.append(_two_lines(f"{q} {q}~ > {q} {q}~ g", [0, 4, 2, 3, 1]))
        out.append(_two_lines(f"{q}~ {q} > {q}~ {q} g", [1, 4, 3, 2, 0]))
    # qbar qbar' -> qbar qbar' g: line q (0: incoming antiquark = q slot, 2 out antiquark)
    for q, r in pairs:
        out.append(_two_lines(f"{q}~ {r}~ > {q}~ {r}~ g", [2, 4, 1, 3, 0]))
    for q in QUARKS:
        out.append(_two_lines(f"{q}~ {q}~ > {q}~ {q}~ g", [2, 4, 1, 3, 0]))
    # q g -> q q' qbar': line q (0 in, 2 out), line q' (3 out quark, 4 out antiquark), g = 1
    for q, r in pairs + [(q, q) for q in QUARKS]:
        out.append(_two_lines(f"{q} g > {q} {r} {r}~", [0, 1, 3, 4, 2]))
        out.append(_two_lines(f"g {q} > {q} {r} {r}~", [1, 0, 3, 4, 2]))
    # qbar g -> qbar q' qbar': line q (0: q slot, 2 out antiquark), line q' (3, 4)
    for q, r in pairs + [(q, q) for q in QUARKS]:
        out.append(_two_lines(f"{q}~ g > {q}~ {r} {r}~", [2, 1, 3, 4, 0]))
        out.append(_two_lines(f"g {q}~ > {q}~ {r} {r}~", [2, 0, 3, 4, 1]))
    return out


def subprocesses(four_quark=True):
    """every subprocess of the module's list (four_quark=False: those with one quark line)"""
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
    if four_quark:
        out += four_quark_subprocesses()
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
