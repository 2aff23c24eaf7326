# SPDX-License-Identifier: 0BSD
"""Phase space on the GPU and the two-stage unweighting, on the GPU.

- pacuda_rambo's momenta equal pacuda_amd.phase_space.rambo's (the same uniforms) to 1e-12
  relative (device and host log/sin/cos differ in the last bits), conserve momentum, are massless,
  and its cut flags equal the host's cuts;
- a small two-stage unweighting of g g > g g g: the cross sections of the plain and the
  surrogate unweighting agree with the weighted Monte Carlo estimate within 3 sigma, and the
  surrogate stage evaluates the exact recurrence on fewer points than the plain one."""
import numpy as np
import pytest

from pacuda_amd import plans, phase_space as ps
from pacuda_amd.runtime import Device
from pacuda_amd.unweight import Generator, Unweighting

pytestmark = pytest.mark.gpu


def test_gpu_rambo_matches_host():
    dev = Device()
    plan = plans.gluon_plan(6)
    gen = Generator(dev, plan)
    count, first = 5000, 1234
    ext, feat, ok = gen.draw(7, first, count)
    got = ext.download((6 * 4, count))
    mom = ps.rambo(6, count, 7, offset=first)
    want = ps.external_planes(mom, plan.crossing)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-9)
    assert np.array_equal(ok, ps.passes_cuts(mom))
    q = got.reshape(6, 4, count)
    total = q.sum(axis=0)
    assert np.max(np.abs(total)) < 1e-9
    m2 = q[:, 0] ** 2 - q[:, 1] ** 2 - q[:, 2] ** 2 - q[:, 3] ** 2
    assert np.max(np.abs(m2)) < 1e-7 * 500 ** 2
    f = feat.download((count, 14), np.float32)
    assert np.all(f >= 0) and np.all(f <= 1)
    for d in (ext, feat):
        d.free()


def test_two_stage_unweighting_small():
    uw = Unweighting(legs=5, hidden=128, layers=3, batch=1024)
    x, logm2 = uw.training_set(1 << 15, seed=3)
    hist = uw.train(x, logm2, epochs=8, lr=2e-3)
    assert hist[-1] < 0.5 * hist[0]
    res = uw.run(1 << 18, seed=4, chunk=1 << 17)
    p, t = res["plain"], res["two_stage"]
    ref, ref_err = res["weighted_mc_sigma_pb"], res["weighted_mc_err_pb"]
    assert abs(p["sigma_pb"] - ref) < 3 * np.hypot(p["sigma_err_pb"], ref_err)
    assert abs(t["sigma_pb"] - ref) < 3 * np.hypot(t["sigma_err_pb"], ref_err)
    assert t["exact_evaluations"] < p["exact_evaluations"]
    print(res, flush=True)
    uw.destroy()
