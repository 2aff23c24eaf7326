# SPDX-License-Identifier: 0BSD
"""Phase space, weight statistics, selection and the two-stage unweighting, on the GPU.

- pacuda_rambo's momenta equal pacuda_amd.phase_space.rambo's (the same uniforms) to 1e-12
  relative (device and host log/sin/cos differ in the last bits), conserve momentum, are massless,
  its cut flags equal the host's cuts, and its log prior equals phase_space.log_prior;
- pacuda_select keeps exactly the indices the host's hit-or-miss test keeps (the same counter
  stream), in ascending order, on ragged sizes, with ids and flags; pacuda_weight_stats gives the
  numpy max / sum / sum of squares and histogram;
- a small two-stage unweighting of g g > g g g g: the cross sections of the plain and the three
  surrogate unweightings (network inside the prior's envelope, network everywhere, prior alone)
  agree with the weighted Monte Carlo estimate within 3 sigma, the two-stage methods evaluate the
  exact recurrence on fewer points than the plain one, and the envelope's stage 1 keeps exactly
  the candidates that the clipped network evaluated on every candidate keeps."""
import numpy as np
import pytest

from pacuda_amd import abi, plans, phase_space as ps
from pacuda_amd.runtime import Device
from pacuda_amd.unweight import BINS, LOG_RANGE, Generator, Selector, Unweighting, quantile_threshold

pytestmark = pytest.mark.gpu


def test_gpu_rambo_matches_host():
    dev = Device()
    plan = plans.gluon_plan(6)
    gen = Generator(dev, plan)
    count, first = 5000, 1234
    d = gen.draw_device(7, first, count, features=True, prior=True)
    got = d["ext"].download((6 * 4, count))
    ok = d["pass"].download(count, np.uint8).astype(bool)
    mom = ps.rambo(6, count, 7, offset=first)
    want = ps.external_planes(mom, plan.crossing)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-9)
    assert np.array_equal(ok, ps.passes_cuts(mom))
    q = got.reshape(6, 4, count)
    total = q.sum(axis=0)
    assert np.max(np.abs(total)) < 1e-9
    m2 = q[:, 0] ** 2 - q[:, 1] ** 2 - q[:, 2] ** 2 - q[:, 3] ** 2
    assert np.max(np.abs(m2)) < 1e-7 * 500 ** 2
    f = d["feat"].download((count, 14), np.float32)
    assert np.all(f >= 0) and np.all(f <= 1)
    lp = d["logp"].download(count)
    assert np.allclose(lp[ok], ps.log_prior(mom, plan.crossing)[ok], rtol=1e-10, atol=1e-9)
    for a in d.values():
        a.free()


@pytest.mark.parametrize("count", [1, 4095, 4096, 70001])
def test_select_and_stats_match_host(count):
    dev = Device()
    sel = Selector(dev)
    rng = np.random.default_rng(count)
    w = np.exp(rng.normal(0, 3, count)) * (rng.uniform(size=count) < 0.9)
    flags = (rng.uniform(size=count) < 0.7).astype(np.uint8)
    ids = rng.permutation(10 * count)[:count].astype(np.uint32)
    d_w, d_f, d_ids = dev.array(w), dev.array(flags), dev.array(ids)
    print("select", count, flush=True)
    # flags only
    idx, _, n = sel.select(count, flags=d_f)
    assert np.array_equal(idx.download(n, np.uint32), np.nonzero(flags)[0])
    idx.free()
    # hit or miss, with ids and values
    thr, first, seed = float(np.quantile(w, 0.9)) if count > 1 else 1.0, 1000, 99
    idx, val, n = sel.select(count, w=d_w, ids=d_ids, first=first, threshold=thr, seed=seed, values=True)
    u = ps.uniform_at(seed, first + ids.astype(np.uint64))
    want = np.nonzero((w > 0) & (u * thr < w))[0]
    assert np.array_equal(idx.download(n, np.uint32), want)
    assert np.array_equal(val.download(n), w[want])
    idx.free()
    val.free()
    # both
    idx, _, n = sel.select(count, w=d_w, flags=d_f, threshold=thr, seed=seed)
    u = ps.uniform_at(seed, np.arange(count))
    want = np.nonzero(flags.astype(bool) & (w > 0) & (u * thr < w))[0]
    assert np.array_equal(idx.download(n, np.uint32), want)
    idx.free()
    # statistics and histogram
    hist = dev.array(np.zeros(BINS, dtype=np.uint32))
    mx, s1, s2 = sel.stats(d_w, count, hist)
    assert mx == np.max(w) and np.isclose(s1, np.sum(w), rtol=1e-12) and np.isclose(s2, np.sum(w * w), rtol=1e-12)
    h = hist.download(BINS, np.uint32)
    lo, hi = LOG_RANGE
    b = np.clip(np.floor((np.log(w[w > 0]) - lo) * (BINS / (hi - lo))), 0, BINS - 1).astype(int)
    assert np.array_equal(h, np.bincount(b, minlength=BINS))
    thr = quantile_threshold(h, 0.99)
    assert np.sum(w > thr) <= int(0.01 * np.sum(w > 0))
    for a in (d_w, d_f, d_ids, hist):
        a.free()
    sel.free()


def test_two_stage_unweighting_small():
    uw = Unweighting(legs=6, hidden=128, layers=3, batch=1024, prior=True)
    x, logm2, logp = uw.training_set(1 << 15, seed=3)
    hist = uw.train(x, logm2, logp, epochs=6, lr=2e-3)
    assert hist[-1] < hist[0]
    res = uw.run(1 << 19, seed=4, chunk=1 << 17, q=0.9999, q2=0.99, variants=("network", "network_dense", "prior"))
    p = res["plain"]
    ref, ref_err = p["weighted"]
    for k in ("plain", "two_stage", "two_stage_dense", "two_stage_prior_only"):
        sig, err = res[k]["estimate"]
        print(k, res[k]["events"], sig, err, ref, ref_err, flush=True)
        assert abs(sig - ref) < 3 * np.hypot(err, ref_err), k
    for k in ("two_stage", "two_stage_dense", "two_stage_prior_only"):
        assert res[k]["exact_evaluations"] < p["exact_evaluations"]
        assert res[k]["events"] > 0
    assert res["two_stage"]["network_evaluations"] < (1 << 19) // 10

    # the envelope's stage 1 keeps exactly what the clipped network evaluated everywhere keeps
    n, chunk, clamp, seed1 = 1 << 17, 1 << 16, 4.0, 22
    t = uw.two_stage_envelope(n, seed=4, chunk=chunk, q=0.9999, q2=0.99, clamp=clamp, accept_seed=seed1,
                              keep_ids=True)
    want_ids = []
    for start in range(0, n, chunk):
        d = uw.gen.draw_device(4, start, chunk, features=True, prior=True)
        d_s0 = uw.surrogate_weights(d, chunk, False)
        s0 = d_s0.download(chunk)
        d_s0.free()
        y = uw.dev.empty(chunk, np.float32)
        uw.mlp.forward(d["feat"], chunk, y)
        nn = y.download(chunk, np.float32).astype(np.float64)
        s = s0 * np.exp(uw.std * np.clip(nn, -clamp, clamp))
        u = ps.uniform_at(seed1, np.arange(start, start + chunk, dtype=np.uint64))
        want_ids.append(np.nonzero(u * t["threshold"] < s)[0] + start)
        for a in list(d.values()) + [y]:
            if a is not None:
                a.free()
    want_ids = np.concatenate(want_ids)
    assert np.array_equal(t["stage1_ids"], want_ids)
    uw.destroy()
