# SPDX-License-Identifier: 0BSD
"""What the trained network gains, and unbiasedness in distributions (g g > g g g g, fixed
seeds).

- Two-stage unweighting with the trained network (inside the prior's envelope) against the same
  scheme with the Parke-Taylor prior alone, at the benchmark's configuration: the network raises
  the stage-2 efficiency (the share of stage-1 survivors the exact recurrence keeps) and the
  number of unweighted events from the same candidates, and narrows log(w/s); all deterministic
  for fixed seeds. The events/s of both are printed.
- The unweighted samples are unbiased in shape as well as in normalisation: binned in the
  smallest pT and the smallest Delta R of the final-state gluons, the weighted histograms of the
  plain and the two-stage samples agree with the exactly weighted histogram of all candidates
  within their statistical errors (chi^2 per bin)."""
import math

import numpy as np
import pytest

from pacuda_amd import phase_space as ps
from pacuda_amd.unweight import Unweighting

pytestmark = pytest.mark.gpu

CANDIDATES = 1 << 22  # the distribution test's candidates (thresholds at the 0.999 quantiles)
BENCH_CANDIDATES = 1 << 24  # the benchmark's (thresholds at the 0.99999 and 0.999 quantiles)
SEED = 4


@pytest.fixture(scope="module")
def runs():
    # the benchmark's network and training (2^20 labelled points of seed 3, 20 epochs)
    uw = Unweighting(legs=6, hidden=256, layers=4, batch=4096, prior=True)
    x, logm2, logp = uw.training_set(1 << 20, seed=3)
    hist = uw.train(x, logm2, logp, epochs=20, lr=1e-3)
    print("training loss", hist[0], "->", hist[-1], flush=True)
    bench = {v: uw.two_stage_envelope(BENCH_CANDIDATES, seed=SEED, keep_ids=True, network=v == "network")
             for v in ("network", "prior")}
    plain = uw.plain(CANDIDATES, seed=SEED, q=0.999, keep_events=True)
    net = uw.two_stage_envelope(CANDIDATES, seed=SEED, q=0.999, q2=0.999, keep_ids=True, network=True)
    scale = uw._scale()
    uw.destroy()
    return bench, plain, net, scale


def test_network_beats_the_prior_alone(runs):
    """the benchmark's configuration: from the same candidates the trained network yields more
    unweighted events, keeps a larger share of the stage-1 survivors (stage-2 efficiency) and leaves
    a much narrower spread of log(w / s) over them than the prior alone (all deterministic for
    fixed seeds); its events/s (wall clock, printed) must not fall far behind, since it pays for
    the network evaluations and the extra exact evaluations"""
    bench, _, _, _ = runs
    net, prior = bench["network"], bench["prior"]
    rate = {}
    for name, r in (("network", net), ("prior only", prior)):
        rate[name] = r["events"] / r["seconds"]
        print(f"{name}: {r['events']} events, {r['exact_evaluations']} exact evaluations, stage-2 efficiency "
              f"{r['stage2_efficiency']:.3f}, {rate[name]:.4g} events/s", flush=True)
    assert net["events"] > 1.2 * prior["events"]
    assert net["stage2_efficiency"] > prior["stage2_efficiency"] + 0.05
    assert net["events"] / net["exact_evaluations"] > prior["events"] / prior["exact_evaluations"]
    assert rate["network"] > 0.8 * rate["prior only"]
    # the survivors' log(w / s): w/s is the correction the exact recurrence applies in stage 2
    spread = {}
    for name, r in (("network", net), ("prior only", prior)):
        spread[name] = float(np.std(np.log(r["stage1_w"] / r["stage1_s"])))
    print("spread of log(w/s):", spread, flush=True)
    assert spread["network"] < 0.6 * spread["prior only"]


def _histogram(values, weights, edges, n_candidates, scale):
    """per bin: the estimate sum(w) / N and its error sqrt(sum(w^2) / N - mean^2 per bin ...)"""
    b = np.clip(np.searchsorted(edges, values, side="right") - 1, 0, len(edges) - 2)
    s1 = np.bincount(b, weights=weights, minlength=len(edges) - 1) / n_candidates
    s2 = np.bincount(b, weights=weights * weights, minlength=len(edges) - 1) / n_candidates
    err = np.sqrt(np.maximum(s2 - s1 * s1, 0.0) / n_candidates)
    return s1 * scale, err * scale


@pytest.mark.parametrize("observable", ["min_pt", "min_dr"])
def test_unweighted_distributions_are_unbiased(runs, observable):
    _, plain, net, scale = runs
    k = 0 if observable == "min_pt" else 1

    def obs(ids):
        return ps.min_pt_and_dr(ps.rambo_at(6, SEED, ids))[k]

    ref_v = obs(plain["all_ids"])
    edges = np.quantile(ref_v, np.linspace(0, 1, 9))
    edges[0], edges[-1] = -np.inf, np.inf
    ref, ref_err = _histogram(ref_v, plain["all_w"], edges, CANDIDATES, scale)
    for name, r in (("plain", plain), ("two-stage", net)):
        h, err = _histogram(obs(r["event_ids"]), r["event_weights"], edges, CANDIDATES, scale)
        chi2 = float(np.sum((h - ref) ** 2 / (err ** 2 + ref_err ** 2)))
        print(f"{observable} {name}: chi2 {chi2:.2f} over {len(h)} bins; per bin {np.round((h - ref) / err, 2)}",
              flush=True)
        # chi^2 of 8 bins: p = 1e-3 at 26.1
        assert chi2 < 26.1, (name, chi2)
        assert np.all(h > 0)
