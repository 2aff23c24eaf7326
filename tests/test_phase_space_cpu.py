# SPDX-License-Identifier: 0BSD
"""The reference's RAMBO stream (convert/src/record.rs:17-52) as the library computes it on the
host (pacuda_rambo_record) equals its Python restatement (pacuda_amd.phase_space.rambo_xorshift)
bit for bit, cuts and draw count included; the points conserve momentum and are massless."""
import ctypes as C

import numpy as np
import pytest

from pacuda_amd import abi, phase_space as ps


@pytest.mark.parametrize("n,seed", [(4, 1), (5, 2), (6, 3)])
def test_record_stream_matches_restatement(n, seed):
    lib = abi.load()
    state = ps.record_state(seed)
    points = 150
    out = np.zeros((points, n, 4))
    drawn = C.c_uint64()
    assert lib.pacuda_rambo_record(state, n, points, ps.ECM, 20.0, 2.5, 0.4, out.ctypes.data_as(abi.f64p),
                                   C.byref(drawn)) == 0
    m = ps.rambo_xorshift(n, 4 * points + 400, 0, state=state)
    ok = ps.passes_cuts(m)
    assert np.count_nonzero(ok) >= points
    assert drawn.value == np.nonzero(ok)[0][points - 1] + 1
    assert np.array_equal(out.view(np.uint64), m[ok][:points].view(np.uint64))
    # no cuts: the raw stream
    raw = np.zeros((20, n, 4))
    assert lib.pacuda_rambo_record(state, n, 20, ps.ECM, 0.0, 0.0, 0.0, raw.ctypes.data_as(abi.f64p), None) == 0
    assert np.array_equal(raw.view(np.uint64), m[:20].view(np.uint64))
    # 2 -> n-2 massless: sum of outgoing = sum of beams, p^2 = 0 (to rounding)
    total = out[:, 2:].sum(axis=1)
    assert np.allclose(total, [[ps.ECM, 0, 0, 0]], atol=1e-9)
    m2 = out[..., 0] ** 2 - (out[..., 1:] ** 2).sum(axis=-1)
    assert np.all(np.abs(m2) < 1e-8 * ps.ECM ** 2)


def test_record_state_avoids_the_seed_or_one_collision():
    assert (2 | 1) == (3 | 1)
    assert ps.record_state(2) != ps.record_state(3)
    lib = abi.load()
    assert lib.pacuda_rambo_record(0, 5, 1, ps.ECM, 0.0, 0.0, 0.0, np.zeros(20).ctypes.data_as(abi.f64p), None) == 2


def test_points_by_id_equal_the_stream():
    import numpy as np
    a = ps.rambo(6, 40, 4, offset=1000)
    b = ps.rambo_at(6, 4, np.arange(1000, 1040))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    pt, dr = ps.min_pt_and_dr(a)
    assert pt.shape == dr.shape == (40,) and np.all(pt > 0) and np.all(dr > 0)
