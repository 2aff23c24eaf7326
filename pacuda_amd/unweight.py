# SPDX-License-Identifier: 0BSD
"""Plain and surrogate two-stage unweighting, on one GPU.

The integrand is w(x) = |M|^2(x) * cut(x) for the plan's colour ordering, summed over
helicities, with x the RAMBO point of the unit hypercube (the flat RAMBO weight and the flux are
constant factors applied to the cross sections). Phase-space points, exact |M|^2 (the fused f64
recurrence), the surrogate, the weight statistics and the hit-or-miss selections all run on the
GPU; the host reads counts and the few accepted values.

Surrogate: s(x) = exp(mu + sigma * NN(features(x)) + log P(x)), NN the library's BF16 MLP on the
14 unit-hypercube features, P the pure-gluon MHV Parke-Taylor structure of the ordering
(sum s_ab^4 / prod s_{l,l+1}); the network learns the normalised residual log(|M|^2 / P).
Without the prior the network learns log|M|^2 itself; without the network s = exp(mu) P.

Thresholds are max reductions: a threshold is the q-quantile of the weights it divides (read off
a device log histogram), and an event above it is kept with the weight w / threshold > 1, so
every estimate below stays unbiased; plain and two-stage use the same q.

plain:      accept candidate k when u_k * w_thr < w_k; the exact recurrence runs on every
            candidate that passes the cuts. Event weight max(1, w / w_thr).
two-stage:  stage 1 accepts when u1_k * s_thr < s_k (s on every candidate); stage 2 runs the
            exact recurrence on the stage-1 survivors only and accepts when u2_k * r_thr < r_k,
            r = w / s, r_thr the q2-quantile of r over the survivors. Event weight
            max(1, s / s_thr) * max(1, r / r_thr); E[s_thr r_thr a c] = w.
Cross sections: sigma = mean over all candidates of the per-candidate estimator (threshold
product times the event weight for accepted candidates, 0 otherwise), with its standard error."""
import math
import time

import numpy as np

from . import abi, plans, phase_space as ps
from .runtime import DeviceArray, FusedPlan, PooledDevice, evaluate
from .surrogate import MLP, he_init

GEV2_TO_PB = 0.3893793721e9
LOG_RANGE = (-230.0, 230.0)  # log-weight histogram range (clamped beyond)
BINS = 8192
CUTS = (20.0, 2.5, 0.4)


class Generator:
    """RAMBO points of a plan's process on the GPU, with cuts, surrogate features and prior"""

    def __init__(self, dev, plan, ecm=ps.ECM):
        self.dev, self.plan, self.ecm = dev, plan, ecm
        self.legs = plan.legs
        self.width = 3 * (self.legs - 2) + 2
        self.src = np.array([c[0] for c in plan.crossing], dtype=np.uint32)
        self.sign = np.array([c[1] for c in plan.crossing], dtype=np.float64)

    def draw_device(self, seed, first, count, features=True, prior=False, momenta=True, index=None):
        """dict of device arrays: ext [legs*4][count] (when momenta), pass [count] u8, feat
        [count][width] f32, logp [count] f64 (the last two when asked for); points first + i, or
        first + index[i] for a device index array"""
        lib, dev = self.dev.lib, self.dev
        out = {"ext": dev.empty(self.legs * 4 * count) if momenta else None, "pass": dev.empty(count, np.uint8)}
        out["feat"] = dev.empty(count * self.width, np.float32) if features else None
        out["logp"] = dev.empty(count) if prior else None
        st = lib.pacuda_rambo(seed, first, count, self.legs, self.ecm, abi.ptr(self.src, abi.u32p), abi.ptr(self.sign),
                              out["ext"].ptr if momenta else None, count, out["feat"].ptr if features else None,
                              out["pass"].ptr, *CUTS, out["logp"].ptr if prior else None,
                              index.ptr if index is not None else None)
        if st != 0:
            raise RuntimeError(f"pacuda_rambo: status {st}")
        return out

    def draw(self, seed, first, count):
        """(external DeviceArray [legs*4][count], features DeviceArray [count][width], pass bool)"""
        d = self.draw_device(seed, first, count)
        ok = d["pass"].download(count, np.uint8).astype(bool)
        d["pass"].free()
        return d["ext"], d["feat"], ok

    def gather(self, ext, feat, stride, index, rows=None):
        """rows of ext (columns index, a host array or a (DeviceArray, count) pair) and feature rows"""
        lib = self.dev.lib
        rows = self.legs * 4 if rows is None else rows
        if isinstance(index, tuple):
            d_idx, n, own = index[0], index[1], False
        elif len(index) == 0:
            return self.dev.empty(rows), (self.dev.empty(self.width, np.float32) if feat is not None else None)
        else:
            d_idx = self.dev.array(np.ascontiguousarray(index, dtype=np.uint32))
            n, own = len(index), True
        out = self.dev.empty(rows * max(n, 1))
        fout = self.dev.empty(max(n, 1) * self.width, np.float32) if feat is not None else None
        st = lib.pacuda_gather(ext.ptr, stride, rows, d_idx.ptr, n, out.ptr,
                               feat.ptr if feat is not None else None, self.width, fout.ptr if fout else None)
        if own:
            d_idx.free()
        if st != 0:
            raise RuntimeError(f"pacuda_gather: status {st}")
        return out, fout


class Selector:
    """pacuda_select / pacuda_weight_stats with their device outputs"""

    def __init__(self, dev):
        self.dev = dev
        self.d_count = dev.empty(1, np.uint32)
        self.d_stats = dev.empty(3 * abi.STATS_BLOCKS)

    def select(self, count, w=None, flags=None, ids=None, first=0, threshold=1.0, seed=0, values=False):
        """(index DeviceArray, values DeviceArray or None, selected count)"""
        lib = self.dev.lib
        idx = self.dev.empty(max(count, 1), np.uint32)
        val = self.dev.empty(max(count, 1)) if values else None
        st = lib.pacuda_select(w.ptr if w is not None else None, flags.ptr if flags is not None else None,
                               ids.ptr if ids is not None else None, first, count, threshold, seed, idx.ptr,
                               val.ptr if val is not None else None, self.d_count.ptr)
        if st != 0:
            raise RuntimeError(f"pacuda_select: status {st}")
        n = int(self.d_count.download(1, np.uint32)[0])
        return idx, val, n

    def stats(self, w, count, hist=None):
        """(max, sum, sum of squares) of w[0..count), reduced in a fixed order; hist accumulates"""
        st = self.dev.lib.pacuda_weight_stats(w.ptr, count, LOG_RANGE[0], LOG_RANGE[1], BINS,
                                              hist.ptr if hist is not None else None, self.d_stats.ptr)
        if st != 0:
            raise RuntimeError(f"pacuda_weight_stats: status {st}")
        s = self.d_stats.download((abi.STATS_BLOCKS, 3))
        return float(np.max(s[:, 0])), float(np.sum(s[:, 1])), float(np.sum(s[:, 2]))

    def free(self):
        self.d_count.free()
        self.d_stats.free()


def quantile_threshold(hist, q):
    """the upper edge of the histogram bin holding the q-quantile of the counted weights"""
    h = np.asarray(hist, dtype=np.int64)
    total = int(h.sum())
    above = int(math.floor((1.0 - q) * total))
    tail = np.cumsum(h[::-1])
    b = BINS - 1 - int(np.searchsorted(tail, above + 1))
    lo, hi = LOG_RANGE
    return math.exp(lo + (b + 1) * (hi - lo) / BINS)


class Unweighting:
    def __init__(self, legs=6, hidden=256, layers=4, batch=4096, prior=True, lib=None):
        # (every buffer of the methods is used on the library's stream only: pooled, so that a
        # method's repeated allocations of candidate-sized buffers do not time the allocator)
        self.dev = PooledDevice(lib)
        self.plan = plans.gluon_plan(legs)
        self.fused = FusedPlan(self.dev.lib, self.plan)
        self.gen = Generator(self.dev, self.plan)
        self.sel = Selector(self.dev)
        self.batch = batch
        self.prior = prior
        self.mlp = MLP(self.gen.width, hidden, layers, max_batch=batch, lib=self.dev.lib)
        self.mean, self.std = 0.0, 1.0
        self.w_ps = ps.rambo_weight(legs)
        self.flux = 1.0 / (2.0 * ps.ECM ** 2)

    # ---- the exact integrand --------------------------------------------------------------
    # (the f64 launch the plain method and the second stage use alike: the cooperative one with
    # 16 wavefronts per 64-point unit, the fastest for g g > g g g g)
    EXACT_LAUNCH = dict(waves_per_cu=16, unit_waves=16)

    def exact(self, ext, count, kernel=abi.KERNEL_F64):
        tot = self.dev.empty(max(count, 1))
        if count:
            launch = self.EXACT_LAUNCH if kernel == abi.KERNEL_F64 else {}
            evaluate(self.dev.lib, [(self.fused, ext, count, count, None, tot)], kernel, **launch)
        return tot

    def surrogate_weights(self, d, count, network=True):
        """device s[count] of a draw (0 where the cuts fail)"""
        lib = self.dev.lib
        y = None
        if network:
            y = self.dev.empty((count + 63) // 64 * 64, np.float32)
            if count % 64:
                raise ValueError("candidate chunks are whole tiles of 64 points")
            self.mlp.forward(d["feat"], count, y)
        s = self.dev.empty(max(count, 1))
        logp = d["logp"] if self.prior else None
        st = lib.pacuda_surrogate_weight(y.ptr if y else None, logp.ptr if logp else None, d["pass"].ptr, count,
                                         self.mean, self.std if network else 0.0, s.ptr)
        if y:
            y.free()
        if st != 0:
            raise RuntimeError(f"pacuda_surrogate_weight: status {st}")
        return s

    # ---- training ---------------------------------------------------------------------------
    def training_set(self, points, seed=3):
        """`points` cut-passing points of stream `seed`: features, exact log|M|^2 (f64 GPU) and
        log P (host arrays)"""
        draw = int(points * 1.25) + 1024
        while True:
            d = self.gen.draw_device(seed, 0, draw, features=True, prior=True)
            idx, _, n = self.sel.select(draw, flags=d["pass"])
            if n >= points:
                break
            for a in d.values():
                a.free()
            idx.free()
            draw = int(draw * 1.2 * points / max(n, 1)) + 1024
        e2, f2 = self.gen.gather(d["ext"], d["feat"], draw, (idx, points))
        lp, _ = self.gen.gather(d["logp"], None, draw, (idx, points), rows=1)
        for a in d.values():
            a.free()
        idx.free()
        tot = self.exact(e2, points)
        m2 = tot.download(points)
        logp = lp.download(points)
        x = f2.download((points, self.gen.width), np.float32)
        for a in (tot, e2, f2, lp):
            a.free()
        return x, np.log(m2), logp

    def train(self, x, logm2, logp=None, epochs=30, lr=1e-3, seed=0, init_seed=5, log=None):
        """fit the network to the normalised residual log(|M|^2 / P) (or log|M|^2 without the
        prior) with Adam and a cosine learning-rate decay; returns the loss of each epoch's last
        batch. The training set stays on the device: each epoch uploads only its permutation and
        gathers the shuffled rows there, and the steps queue without host synchronisation."""
        target = logm2 - logp if (self.prior and logp is not None) else logm2
        self.mean, self.std = float(np.mean(target)), float(np.std(target))
        t = ((target - self.mean) / self.std).astype(np.float32)
        self.mlp.set_params(he_init(self.gen.width, self.mlp.hidden, self.mlp.layers, init_seed))
        n = len(t) // self.batch * self.batch
        if n == 0:
            raise ValueError(f"training needs at least one batch of {self.batch} points")
        rng = np.random.default_rng(seed)
        w = self.gen.width
        lib = self.dev.lib
        d_xall = self.dev.array(np.ascontiguousarray(x, dtype=np.float32))
        d_tall = self.dev.array(t)
        d_x = self.dev.empty(n * w, np.float32)
        d_t = self.dev.empty(n, np.float32)
        d_idx = self.dev.empty(n, np.uint32)
        dummy = self.dev.empty(1)
        history = []
        batches = n // self.batch
        steps = epochs * batches
        step = 0
        order = rng.permutation(len(t))[:n].astype(np.uint32)
        for epoch in range(epochs):
            d_idx.upload(order)
            for src, width, dst in ((d_xall, w, d_x), (d_tall, 1, d_t)):
                st = lib.pacuda_gather(dummy.ptr, 1, 0, d_idx.ptr, n, dummy.ptr, src.ptr, width, dst.ptr)
                if st != 0:
                    raise RuntimeError(f"pacuda_gather: status {st}")
            self.dev.synchronize()  # (the gather runs on the runner's stream, training on the network's)
            # one native call queues the epoch's steps (forward, backward and Adam per batch); the
            # next epoch's permutation is drawn on the host meanwhile
            self.mlp.train_steps(d_x, d_t, self.batch, batches, lr, step, steps, want_loss=False)
            step += batches
            if epoch + 1 < epochs:
                order = rng.permutation(len(t))[:n].astype(np.uint32)
            history.append(self.mlp.last_loss())  # (the steps are done with d_x, d_t after this)
            if log:
                log(f"epoch {epoch}: loss {history[-1]:.4f}")
        self.dev.synchronize()
        for d in (d_xall, d_tall, d_x, d_t, d_idx, dummy):
            d.free()
        return history

    # ---- unweighting ---------------------------------------------------------------------------
    def _timed(self, clock, key, fn, *args, **kw):
        self.dev.synchronize()
        t = time.perf_counter()
        out = fn(*args, **kw)
        self.dev.synchronize()
        clock[key] = clock.get(key, 0.0) + time.perf_counter() - t
        return out

    def plain(self, candidates, seed=4, chunk=1 << 24, q=0.99999, accept_seed=21, keep_events=False):
        """plain unweighting of `candidates` points: dict of counts, times and cross sections (with
        keep_events: the accepted candidates' ids and weights, and every passing candidate's id and
        exact weight, for distributions)"""
        clock, hist = {}, self.dev.array(np.zeros(BINS, dtype=np.uint32))
        kept, sums = [], [0.0, 0.0, 0.0]  # per chunk (w, ids, n, first); sum w, sum w^2, max w
        passed = 0
        for start in range(0, candidates, chunk):
            count = min(chunk, candidates - start)
            d = self._timed(clock, "generate", self.gen.draw_device, seed, start, count, features=False)
            idx, _, n = self._timed(clock, "select", self.sel.select, count, flags=d["pass"])
            e2, _ = self._timed(clock, "gather", self.gen.gather, d["ext"], None, count, (idx, n))
            for a in d.values():
                if a is not None:
                    a.free()
            w = self._timed(clock, "exact", self.exact, e2, n)
            e2.free()
            mx, s1, s2 = self._timed(clock, "stats", self.sel.stats, w, n, hist)
            sums[0] += s1
            sums[1] += s2
            sums[2] = max(sums[2], mx)
            kept.append((w, idx, n, start))
            passed += n
        h = hist.download(BINS, np.uint32)
        hist.free()
        w_thr = quantile_threshold(h, q)
        acc_w, events = [], {"event_ids": [], "all_ids": [], "all_w": []}
        for w, idx, n, start in kept:
            a_idx, a_val, m = self._timed(clock, "accept", self.sel.select, n, w=w, ids=idx, first=start,
                                          threshold=w_thr, seed=accept_seed, values=True)
            acc_w.append(a_val.download(m))
            if keep_events:
                pos = a_idx.download(m, np.uint32)
                all_ids = idx.download(n, np.uint32).astype(np.int64) + start
                events["event_ids"].append(all_ids[pos])
                events["all_ids"].append(all_ids)
                events["all_w"].append(w.download(n))
            for a in (a_idx, a_val, w, idx):
                a.free()
        acc_w = np.concatenate(acc_w)
        x = w_thr * np.maximum(1.0, acc_w / w_thr)
        extra = {}
        if keep_events:
            extra = {k: np.concatenate(v) for k, v in events.items()}
            extra["event_weights"] = x
        return {**extra, "candidates": candidates, "passed_cuts": passed, "exact_evaluations": passed,
                "events": len(acc_w), "overweight_events": int(np.sum(acc_w > w_thr)), "threshold": w_thr,
                "estimate": self._estimate(x, candidates), "weighted": self._weighted(sums, candidates),
                "seconds": sum(clock.values()), "times_s": clock}

    def two_stage(self, candidates, seed=4, chunk=1 << 24, q=0.99999, q2=0.999, network=True, accept_seed=22,
                  stage2_seed=23):
        """two-stage unweighting of `candidates` points (the same points as plain() for one seed)"""
        clock, hist = {}, self.dev.array(np.zeros(BINS, dtype=np.uint32))
        staged = []
        for start in range(0, candidates, chunk):
            count = min(chunk, candidates - start)
            d = self._timed(clock, "generate", self.gen.draw_device, seed, start, count, features=network,
                            prior=self.prior)
            s = self._timed(clock, "surrogate", self.surrogate_weights, d, count, network)
            self._timed(clock, "stats", self.sel.stats, s, count, hist)
            for k in ("pass", "feat", "logp"):
                if d[k] is not None:
                    d[k].free()
            staged.append((d["ext"], s, count, start))
        h = hist.download(BINS, np.uint32)
        hist.free()
        s_thr = quantile_threshold(h, q)
        ids, s1, w1 = [], [], []
        for ext, s, count, start in staged:
            a_idx, a_val, m = self._timed(clock, "stage1", self.sel.select, count, w=s, first=start, threshold=s_thr,
                                          seed=accept_seed, values=True)
            e2, _ = self._timed(clock, "gather", self.gen.gather, ext, None, count, (a_idx, m))
            w = self._timed(clock, "exact", self.exact, e2, m)
            t = time.perf_counter()
            ids.append(a_idx.download(m, np.uint32).astype(np.int64) + start)
            s1.append(a_val.download(m))
            w1.append(w.download(m))
            clock["stage2"] = clock.get("stage2", 0.0) + time.perf_counter() - t
            for a in (a_idx, a_val, e2, w, ext, s):
                a.free()
        t = time.perf_counter()
        ids, s1, w1 = np.concatenate(ids), np.concatenate(s1), np.concatenate(w1)
        r = w1 / s1
        r_thr = float(np.quantile(r, q2)) if len(r) else 1.0
        acc = ps.uniform_at(stage2_seed, ids) * r_thr < r
        clock["stage2"] = clock.get("stage2", 0.0) + time.perf_counter() - t
        x = s_thr * r_thr * np.maximum(1.0, s1[acc] / s_thr) * np.maximum(1.0, r[acc] / r_thr)
        return {"candidates": candidates, "stage1_accepted": len(ids), "exact_evaluations": len(ids),
                "events": int(acc.sum()), "overweight_events": int(np.sum((s1[acc] > s_thr) | (r[acc] > r_thr))),
                "threshold": s_thr, "r_threshold": r_thr, "stage2_efficiency": float(acc.mean()) if len(r) else 0.0,
                "estimate": self._estimate(x, candidates), "seconds": sum(clock.values()), "times_s": clock}

    def two_stage_envelope(self, candidates, seed=4, chunk=1 << 24, q=0.99999, q2=0.999, clamp=4.0, accept_seed=22,
                           stage2_seed=23, keep_ids=False, network=True):
        """two-stage unweighting with the network evaluated only where it can matter.

        The surrogate is s = P e^(mu + sigma clip(NN, -clamp, clamp)): any positive function is a
        valid surrogate, and the clip bounds it by the envelope E = P e^(mu + sigma clamp). Stage 1
        keeps a candidate when u1 s_thr < s; since s <= E, every candidate with u1 s_thr >= E is
        rejected from the prior alone, and the network runs only on the others (the same
        decisions as evaluating it everywhere). s_thr is the q-quantile of P e^mu (the surrogate
        without its network), so it is known before any network evaluation."""
        if not self.prior:
            raise ValueError("the envelope needs the prior")
        clock, hist = {}, self.dev.array(np.zeros(BINS, dtype=np.uint32))
        staged = []
        for start in range(0, candidates, chunk):
            count = min(chunk, candidates - start)
            # cuts and prior only: momenta and features are made again for the envelope's points
            d = self._timed(clock, "generate", self.gen.draw_device, seed, start, count, features=False, prior=True,
                            momenta=False)
            s0 = self._timed(clock, "prior", self.surrogate_weights, d, count, False)
            self._timed(clock, "stats", self.sel.stats, s0, count, hist)
            for k in ("pass", "logp"):
                d[k].free()
            staged.append((s0, count, start))
        h = hist.download(BINS, np.uint32)
        hist.free()
        s_thr = quantile_threshold(h, q)
        lift = math.exp(self.std * clamp) if network else 1.0  # without the network, s = P e^mu exactly
        ids, s1, w1 = [], [], []
        n_env = n_net = 0
        for s0, count, start in staged:
            e_idx, e_val, m = self._timed(clock, "envelope", self.sel.select, count, w=s0, first=start,
                                          threshold=s_thr / lift, seed=accept_seed, values=True)
            n_env += m
            t = time.perf_counter()
            idx = e_idx.download(m, np.uint32)
            base = e_val.download(m)
            keep = np.zeros(m, dtype=bool)
            s = np.zeros(m)
            ext = None
            if m:
                # the envelope's points again, with momenta and features (padded to whole tiles)
                pad = (m + 63) // 64 * 64
                pidx = self.dev.array(np.concatenate([idx, np.repeat(idx[-1:], pad - m)]))
                d = self.gen.draw_device(seed, start, pad, features=network, prior=False, index=pidx)
                ext = d["ext"]
                if network:
                    y = self.dev.empty(pad, np.float32)
                    self.mlp.forward(d["feat"], pad, y)
                    nn = y.download(pad, np.float32)[:m].astype(np.float64)
                    for a in (d["feat"], y):
                        a.free()
                    n_net += pad
                    s = base * np.exp(self.std * np.clip(nn, -clamp, clamp))
                    keep = ps.uniform_at(accept_seed, start + idx.astype(np.uint64)) * s_thr < s
                else:
                    s, keep = base, np.ones(m, dtype=bool)
                for a in (pidx, d["pass"]):
                    a.free()
            self.dev.synchronize()
            clock["network"] = clock.get("network", 0.0) + time.perf_counter() - t
            k_pos = np.nonzero(keep)[0]
            n = len(k_pos)
            e2, _ = self._timed(clock, "gather", self.gen.gather, ext, None, pad if m else 1, k_pos) if ext is not None \
                else (self.dev.empty(self.gen.legs * 4), None)
            w = self._timed(clock, "exact", self.exact, e2, n)
            t = time.perf_counter()
            ids.append(idx[keep].astype(np.int64) + start)
            s1.append(s[keep])
            w1.append(w.download(n))
            clock["stage2"] = clock.get("stage2", 0.0) + time.perf_counter() - t
            for a in (e_idx, e_val, e2, w, s0) + ((ext,) if ext is not None else ()):
                a.free()
        t = time.perf_counter()
        ids, s1, w1 = np.concatenate(ids), np.concatenate(s1), np.concatenate(w1)
        r = w1 / s1
        r_thr = float(np.quantile(r, q2)) if len(r) else 1.0
        acc = ps.uniform_at(stage2_seed, ids) * r_thr < r
        clock["stage2"] = clock.get("stage2", 0.0) + time.perf_counter() - t
        x = s_thr * r_thr * np.maximum(1.0, s1[acc] / s_thr) * np.maximum(1.0, r[acc] / r_thr)
        extra = {"stage1_ids": ids, "stage1_s": s1, "stage1_w": w1, "event_ids": ids[acc], "event_weights": x} \
            if keep_ids else {}
        if network:
            extra["envelope_accepted"] = n_env
        return {**extra, "candidates": candidates, "network_evaluations": n_net,
                "stage1_accepted": len(ids), "exact_evaluations": len(ids),
                "events": int(acc.sum()), "overweight_events": int(np.sum((s1[acc] > s_thr) | (r[acc] > r_thr))),
                "threshold": s_thr, "r_threshold": r_thr, "clamp": clamp,
                "stage2_efficiency": float(acc.mean()) if len(r) else 0.0,
                "estimate": self._estimate(x, candidates), "seconds": sum(clock.values()), "times_s": clock}

    def _scale(self):
        return self.w_ps * self.flux * GEV2_TO_PB

    def _estimate(self, x, n):
        """(sigma, error) in pb of the per-candidate estimator with values x on accepted events"""
        mean = np.sum(x) / n
        var = max(np.sum(x * x) / n - mean * mean, 0.0)
        return float(mean * self._scale()), float(math.sqrt(var / n) * self._scale())

    def _weighted(self, sums, n):
        mean = sums[0] / n
        var = max(sums[1] / n - mean * mean, 0.0)
        return float(mean * self._scale()), float(math.sqrt(var / n) * self._scale())

    VARIANTS = {"network": "two_stage", "network_dense": "two_stage_dense", "prior": "two_stage_prior_only"}

    def run(self, candidates, seed=4, chunk=1 << 24, q=0.99999, q2=0.999, variants=("network", "prior"), repeats=3):
        """plain and two-stage unweighting on the same candidates; `variants` of the surrogate:
        "network" (the trained network with the prior, evaluated inside the envelope only),
        "network_dense" (the same network evaluated on every candidate) and "prior" (the prior
        alone, no network). Pulls: against the plain estimate (errors added in quadrature) and
        against the weighted Monte Carlo estimate of the same candidates (the method's error).
        Each method runs `repeats` times on the same streams (the same events each time) and
        reports its fastest run, with every run's seconds in `seconds_all`."""
        def best(f):
            runs = [f() for _ in range(max(1, repeats))]
            r = min(runs, key=lambda x: x["seconds"])
            r["seconds_all"] = [x["seconds"] for x in runs]
            return r

        out = {"plain": best(lambda: self.plain(candidates, seed, chunk, q))}
        p = out["plain"]
        rate_p = p["events"] / p["seconds"]
        p["events_per_s"] = rate_p
        sig_p, err_p = p["estimate"]
        sig_w = p["weighted"][0]
        p["sigma_pull_vs_weighted"] = (sig_p - sig_w) / err_p
        for v in variants:
            if v in ("prior", "network") and not self.prior:
                if v == "network":
                    v = "network_dense"
                else:
                    continue
            if v in ("network", "prior"):
                t = best(lambda: self.two_stage_envelope(candidates, seed, chunk, q, q2, network=(v == "network")))
            else:
                t = best(lambda: self.two_stage(candidates, seed, chunk, q, q2, network=True))
            t["events_per_s"] = t["events"] / t["seconds"]
            t["gain"] = t["events_per_s"] / rate_p if rate_p else None
            sig, err = t["estimate"]
            t["sigma_pull"] = (sig - sig_p) / math.hypot(err, err_p)
            t["sigma_pull_vs_weighted"] = (sig - sig_w) / err
            out[self.VARIANTS[v]] = t
        return out

    def destroy(self):
        self.sel.free()
        self.mlp.destroy()
        self.fused.destroy()
        self.dev.close()
