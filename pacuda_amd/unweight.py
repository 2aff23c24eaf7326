# SPDX-License-Identifier: 0BSD
"""Plain and surrogate two-stage unweighting on one GPU.

The integrand is w(x) = |M|^2(x) * cut(x) * w_RAMBO for the plan's colour ordering, summed over
helicities (g = 1), with x the RAMBO point of the unit hypercube. Every phase-space point and
every exact |M|^2 is computed on the GPU (pacuda_rambo, pacuda_evaluate in f64); the surrogate
exp(NN(features)) of log|M|^2 is the library's BF16 MLP.

plain:      accept candidate k with probability w_k / w_max (w_max the sample maximum); the
            exact recurrence runs on every candidate that passes the cuts.
two-stage:  stage 1 accepts with probability s_k / s_max (s the surrogate weight, every candidate);
            stage 2 runs the exact recurrence on the stage-1 survivors only and accepts with
            probability min(1, r_k / r_max), r = w / s; an event with r > r_max is kept with the
            correction weight r / r_max. The accepted sample, with its weights, is unbiased:
            E[s_max r_max a_k c_k] = E[w] (a the acceptance, c the correction weight).
Cross sections: sigma = N_candidates^-1 sum of the per-candidate estimator, with its standard
error; plain w_max * accepted / N, two-stage s_max r_max * sum(weights) / N."""
import math
import time

import numpy as np

from . import abi, plans, phase_space as ps
from .runtime import Device, DeviceArray, FusedPlan, evaluate
from .surrogate import MLP, he_init

GEV2_TO_PB = 0.3893793721e9


class Generator:
    """RAMBO points of a plan's process on the GPU, with cuts and surrogate features"""

    def __init__(self, dev, plan, ecm=ps.ECM):
        self.dev, self.plan, self.ecm = dev, plan, ecm
        self.legs = plan.legs
        self.width = 3 * (self.legs - 2) + 2
        self.src = np.array([c[0] for c in plan.crossing], dtype=np.uint32)
        self.sign = np.array([c[1] for c in plan.crossing], dtype=np.float64)

    def draw(self, seed, first, count):
        """(external DeviceArray [legs*4][count], features DeviceArray [count][width], pass bool)"""
        lib = self.dev.lib
        ext = self.dev.empty(self.legs * 4 * count)
        feat = self.dev.empty(count * self.width, np.float32)
        passed = self.dev.empty(count, np.uint8)
        st = lib.pacuda_rambo(seed, first, count, self.legs, self.ecm, abi.ptr(self.src, abi.u32p), abi.ptr(self.sign),
                              ext.ptr, count, feat.ptr, passed.ptr, 20.0, 2.5, 0.4)
        if st != 0:
            raise RuntimeError(f"pacuda_rambo: status {st}")
        ok = passed.download(count, np.uint8).astype(bool)
        passed.free()
        return ext, feat, ok

    def gather(self, ext, feat, stride, index):
        lib = self.dev.lib
        index = np.ascontiguousarray(index, dtype=np.uint32)
        d_idx = self.dev.array(index)
        n = len(index)
        out = self.dev.empty(self.legs * 4 * max(n, 1))
        fout = self.dev.empty(max(n, 1) * self.width, np.float32) if feat is not None else None
        st = lib.pacuda_gather(ext.ptr, stride, self.legs * 4, d_idx.ptr, n, out.ptr,
                               feat.ptr if feat is not None else None, self.width, fout.ptr if fout else None)
        d_idx.free()
        if st != 0:
            raise RuntimeError(f"pacuda_gather: status {st}")
        return out, fout


class Unweighting:
    def __init__(self, legs=6, hidden=256, layers=4, batch=4096, lib=None):
        self.dev = Device(lib)
        self.plan = plans.gluon_plan(legs)
        self.fused = FusedPlan(self.dev.lib, self.plan)
        self.gen = Generator(self.dev, self.plan)
        self.batch = batch
        self.mlp = MLP(self.gen.width, hidden, layers, max_batch=batch, lib=self.dev.lib)
        self.mean, self.std = 0.0, 1.0
        self.w_ps = ps.rambo_weight(legs)
        self.flux = 1.0 / (2.0 * ps.ECM ** 2)

    # ---- the exact integrand --------------------------------------------------------------
    def exact(self, ext, count, kernel=abi.KERNEL_F64):
        tot = self.dev.empty(max(count, 1))
        if count:
            evaluate(self.dev.lib, [(self.fused, ext, count, count, None, tot)], kernel)
        return tot

    def surrogate(self, feat, count):
        """exp of the network, in |M|^2 units (padding to whole tiles of 64)"""
        pad = (count + 63) // 64 * 64
        y = self.dev.empty(pad, np.float32)
        if pad != count:
            # pad the feature rows with a copy of the last ones
            host = feat.download((count, self.gen.width), np.float32)
            host = np.concatenate([host, np.repeat(host[-1:], pad - count, axis=0)])
            f2 = self.dev.array(host)
            self.mlp.forward(f2, pad, y)
            f2.free()
        else:
            self.mlp.forward(feat, pad, y)
        out = y.download(pad, np.float32)[:count].astype(np.float64)
        y.free()
        return np.exp(self.mean + self.std * out)

    # ---- training ---------------------------------------------------------------------------
    def training_set(self, points, seed=3):
        """`points` cut-passing points of stream `seed` with their exact log|M|^2 (f64 GPU)"""
        draw = int(points * 1.1) + 1024
        while True:
            ext, feat, ok = self.gen.draw(seed, 0, draw)
            idx = np.nonzero(ok)[0][:points]
            if len(idx) == points:
                break
            ext.free()
            feat.free()
            draw = int(draw * 1.2 * points / max(len(idx), 1)) + 1024
        e2, f2 = self.gen.gather(ext, feat, draw, idx)
        ext.free()
        feat.free()
        tot = self.exact(e2, points)
        m2 = tot.download(points)
        tot.free()
        e2.free()
        x = f2.download((points, self.gen.width), np.float32)
        f2.free()
        return x, np.log(m2)

    def train(self, x, logm2, epochs=30, lr=1e-3, seed=0, init_seed=5, log=None):
        self.mean, self.std = float(np.mean(logm2)), float(np.std(logm2))
        t = ((logm2 - self.mean) / self.std).astype(np.float32)
        self.mlp.set_params(he_init(self.gen.width, self.mlp.hidden, self.mlp.layers, init_seed))
        n = len(t) // self.batch * self.batch
        rng = np.random.default_rng(seed)
        d_x = self.dev.empty(self.batch * self.gen.width, np.float32)
        d_t = self.dev.empty(self.batch, np.float32)
        d_y = self.dev.empty(self.batch, np.float32)
        history = []
        steps = epochs * (n // self.batch)
        step = 0
        for epoch in range(epochs):
            order = rng.permutation(len(t))[:n]
            losses = []
            for b in range(n // self.batch):
                sel = order[b * self.batch:(b + 1) * self.batch]
                d_x.upload(x[sel])
                d_t.upload(t[sel])
                self.mlp.forward(d_x, self.batch, d_y, train=True)
                losses.append(self.mlp.backward(d_y, d_t, self.batch))
                # cosine learning-rate decay
                self.mlp.adam(lr * 0.5 * (1 + math.cos(math.pi * step / steps)))
                step += 1
            history.append(float(np.mean(losses)))
            if log:
                log(f"epoch {epoch}: loss {history[-1]:.4f}")
        for d in (d_x, d_t, d_y):
            d.free()
        return history

    # ---- unweighting ---------------------------------------------------------------------------
    def run(self, candidates, seed=4, chunk=1 << 22, r_quantile=0.999, rng_seed=11):
        """both unweightings on the same candidates; dict of counts, times, cross sections"""
        rng = np.random.default_rng(rng_seed)
        sync = self.dev.synchronize
        t_gen = t_nn = t_exact_plain = t_exact_stage2 = t_select = 0.0
        s_all, w_all, ok_all = [], [], []
        stage = []  # per chunk: (ext, feat, count) kept for stage 2
        for start in range(0, candidates, chunk):
            count = min(chunk, candidates - start)
            sync()
            t = time.perf_counter()
            ext, feat, ok = self.gen.draw(seed, start, count)
            sync()
            t_gen += time.perf_counter() - t
            t = time.perf_counter()
            s = self.surrogate(feat, count) * ok
            sync()
            t_nn += time.perf_counter() - t
            # plain: the exact integrand on every candidate that passes the cuts
            t = time.perf_counter()
            idx = np.nonzero(ok)[0]
            e2, _ = self.gen.gather(ext, None, count, idx)
            tot = self.exact(e2, len(idx))
            w = np.zeros(count)
            w[idx] = tot.download(len(idx))
            sync()
            t_exact_plain += time.perf_counter() - t
            tot.free()
            e2.free()
            s_all.append(s)
            w_all.append(w)
            ok_all.append(ok)
            stage.append((ext, feat, count))
        s = np.concatenate(s_all)
        w = np.concatenate(w_all)
        N = candidates
        # plain unweighting
        w_max = float(np.max(w))
        u = rng.uniform(size=N)
        plain_acc = u < w / w_max
        n_plain = int(plain_acc.sum())
        x_plain = w_max * plain_acc
        # two-stage: stage 1 on the surrogate
        t = time.perf_counter()
        s_max = float(np.max(s))
        u1 = rng.uniform(size=N)
        acc1 = u1 < s / s_max
        t_select += time.perf_counter() - t
        # stage 2: the exact integrand on the stage-1 survivors only (gathered per chunk)
        w_stage = np.zeros(N)
        at = 0
        for ext, feat, count in stage:
            sync()
            t = time.perf_counter()
            idx = np.nonzero(acc1[at:at + count])[0]
            e2, _ = self.gen.gather(ext, None, count, idx)
            tot = self.exact(e2, len(idx))
            w_stage[at + idx] = tot.download(len(idx))
            sync()
            t_exact_stage2 += time.perf_counter() - t
            tot.free()
            e2.free()
            ext.free()
            feat.free()
            at += count
        t = time.perf_counter()
        r = np.zeros(N)
        r[acc1] = w_stage[acc1] / s[acc1]
        r_max = float(np.quantile(r[acc1], r_quantile)) if acc1.any() else 1.0
        u2 = rng.uniform(size=N)
        acc2 = acc1 & (u2 < r / r_max)
        weight = np.where(acc2, np.maximum(1.0, r / r_max), 0.0)
        t_select += time.perf_counter() - t
        n_two = int(acc2.sum())
        x_two = s_max * r_max * weight
        scale = self.w_ps * self.flux * GEV2_TO_PB

        def est(xs):
            return float(np.mean(xs) * scale), float(np.std(xs) / math.sqrt(N) * scale)

        sig_plain, err_plain = est(x_plain)
        sig_two, err_two = est(x_two)
        sig_w, err_w = est(w)
        time_plain = t_gen + t_exact_plain
        time_two = t_gen + t_nn + t_exact_stage2 + t_select
        return {
            "candidates": N, "passed_cuts": int(np.concatenate(ok_all).sum()),
            "plain": {"events": n_plain, "seconds": time_plain, "events_per_s": n_plain / time_plain,
                      "efficiency": n_plain / N, "exact_evaluations": int(np.concatenate(ok_all).sum()),
                      "sigma_pb": sig_plain, "sigma_err_pb": err_plain},
            "two_stage": {"events": n_two, "seconds": time_two, "events_per_s": n_two / time_two,
                          "stage1_accepted": int(acc1.sum()), "exact_evaluations": int(acc1.sum()),
                          "overweight_events": int((weight > 1).sum()), "r_max": r_max,
                          "sigma_pb": sig_two, "sigma_err_pb": err_two},
            "weighted_mc_sigma_pb": sig_w, "weighted_mc_err_pb": err_w,
            "gain": (n_two / time_two) / (n_plain / time_plain) if n_plain else None,
            "sigma_pull": (sig_two - sig_plain) / math.hypot(err_plain, err_two),
            "times_s": {"generate": t_gen, "surrogate": t_nn, "exact_plain": t_exact_plain,
                        "exact_stage2": t_exact_stage2, "select": t_select},
        }

    def destroy(self):
        self.mlp.destroy()
        self.fused.destroy()
