# SPDX-License-Identifier: 0BSD
"""Pins of the CPU oracle (oracle/refcpu.cpp), without a GPU.

- runner/rectest.c, the reference's own known-answer client, passes against the oracle's v1
  emulation for every kernel (its host-side reductions and sampling are checked bit for bit
  inside it), and tests/runner_variants.sh's criteria hold: f64 sorted == unsorted byte for byte,
  df64 within 1e-9 and mixed/mixed64 within 1e-4 of the largest f64 value;
- the source wavefunctions equal the closed forms of convert/CONVENTIONS.md:52-61 (restated in
  tests/source_cases.py after the reference's own Python legend) within 1e-13 relative, on the
  reference's two case generators, and satisfy the textbook identities (transversality,
  normalisation, spinor completeness), which do not depend on that legend;
- the momentum forms equal their brute-force sums bit for bit;
- whole generated plans reproduce Parke-Taylor (tests/test_physics.py).
A mistake planted in the oracle's arithmetic (a dropped term, a sign, a swapped operand) fails
at least one of these."""
import math
import os

import numpy as np
import pytest

import harness
import refcpu


def _sc():
    """the restated case generators and legend (CPU only; GPU runs load tests/golden data)"""
    import source_cases
    return source_cases

KERNEL_ARGS = [("f64",), ("f64", "nosort"), ("df64",), (), ("mixed",), ("mixed64",)]


def _have_rectest():
    return os.path.exists(refcpu.RECTEST)


@pytest.fixture(scope="module")
def rectest_outputs():
    if not _have_rectest():
        pytest.skip("oracle/_ref/rectest was not built (no reference checkout)")
    outs = {}
    for args in KERNEL_ARGS:
        code, out, err = harness.rectest(refcpu.ABI_LIBRARY, *args)
        assert code == 0, f"rectest {args} against the oracle: exit {code}: {err}"
        outs[args] = np.array([float(v) for v in out.split()])
    return outs


def test_rectest_kats_pass_on_oracle(rectest_outputs):
    want = rectest_outputs[("f64",)]
    assert len(want) == 3 * 300 * 2 and np.any(want)


def test_runner_variants_criteria(rectest_outputs):
    f64 = rectest_outputs[("f64",)]
    assert np.array_equal(f64.view(np.uint64), rectest_outputs[("f64", "nosort")].view(np.uint64))
    scale = np.max(np.abs(f64))
    for args, tol in [(("df64",), 1e-9), ((), 1e-9), (("mixed",), 1e-4), (("mixed64",), 1e-4)]:
        err = np.max(np.abs(rectest_outputs[args] - f64)) / scale
        assert err <= tol, (args, err)
        assert err > 0, f"{args}: identical to f64, so the precision mode is not in effect"


@pytest.mark.parametrize("which", ["legend", "patch"])
def test_sources_match_legend(which):
    cases = _sc().legend_cases() if which == "legend" else _sc().patch_cases()
    codes = [c[0] for c in cases]
    masses = [c[1] for c in cases]
    mom = np.array([c[2] for c in cases]).reshape(-1)
    got, _, _ = refcpu.test_sources(codes, masses, mom)
    worst = 0.0
    for i, (code, m, k, name, h, ch) in enumerate(cases):
        want = [complex(x) for x in _sc().legend(name, k, h, ch, m)] + [0j] * 4
        scale = max(1e-300, max(abs(x) for x in got[i]))
        err = max(abs(got[i][j] - want[j]) for j in range(4)) / scale
        worst = max(worst, err)
        assert err < 1e-13, (name, k, h, ch, m, got[i], want[:4])
    assert worst < 1e-13


def _mdot(a, b):
    return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3]


def test_polarisation_identities():
    """massless vectors: eps.k = 0, eps.eps* = -1, eps(+).eps(-)* = 0, both energy signs"""
    rng = np.random.default_rng(5)
    codes, masses, moms = [], [], []
    for _ in range(200):
        p3 = rng.uniform(-400, 400, 3)
        e = math.sqrt(float(p3 @ p3))
        sgn = rng.choice([-1.0, 1.0])
        for h in (-1, 1):
            codes.append(_sc().code(3, 0, h, 0))
            masses.append(0.0)
            moms.append([sgn * e, *(sgn * p3)])
    got, _, _ = refcpu.test_sources(codes, masses, np.array(moms).reshape(-1))
    for i in range(0, len(codes), 2):
        k = moms[i]
        em, ep = got[i], got[i + 1]
        assert abs(_mdot(em, k)) < 1e-10 * k[0] ** 2 / abs(k[0])
        assert abs(_mdot(ep, np.conj(ep)) + 1) < 1e-13
        assert abs(_mdot(em, np.conj(em)) + 1) < 1e-13
        assert abs(_mdot(ep, np.conj(em))) < 1e-13


def test_massless_spinor_completeness():
    """Dirac u(p,h) for m = 0: sum over h of u u^dagger = pslash gamma^0 (chiral blocks), E > 0"""
    rng = np.random.default_rng(6)
    sigma = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]
    for _ in range(50):
        p3 = rng.uniform(-300, 300, 3)
        e = math.sqrt(float(p3 @ p3))
        p = [e, *p3]
        us = []
        for h in (-1, 1):
            got, _, _ = refcpu.test_sources([_sc().code(2, 0, h, 0)], [0.0], np.array(p))
            us.append(got[0])
        m = sum(np.outer(u, np.conj(u)) for u in us)
        # pslash gamma^0 in the chiral basis: [[p.sigma, 0], [0, p.sigmabar]] up to the block order
        ps = p[0] * sigma[0] - p[1] * sigma[1] - p[2] * sigma[2] - p[3] * sigma[3]
        psb = p[0] * sigma[0] + p[1] * sigma[1] + p[2] * sigma[2] + p[3] * sigma[3]
        # HELAS's spinors are the complex conjugates of the textbook ones in this basis: compare
        # each chiral block with p.sigma / p.sigmabar up to conjugation
        def same(a, b):
            return np.allclose(a, b, atol=1e-9 * e) or np.allclose(a, np.conj(b), atol=1e-9 * e)
        ok = (same(m[:2, :2], ps) and same(m[2:, 2:], psb)) or (same(m[:2, :2], psb) and same(m[2:, 2:], ps))
        assert ok
        assert np.allclose(m[:2, 2:], 0, atol=1e-9 * e)


def test_forms_brute_force():
    """the oracle's momentum forms (the FORMS_ON_DEVICE computation, runner/recurrence_runner.cu:
    745-764) equal the brute-force sum 0.0 + c0 p(s0) + c1 p(s1) + ... over the external momenta
    (convert/CONVENTIONS.md:44), bit for bit, for a gluon plan, a quark-line plan and a crossed
    p p > j j j subprocess; and the forms of a gluon plan are the momenta flowing through its
    propagators (sums of cyclically consecutive legs), so the full sum is zero"""
    from pacuda_amd import plans, processes, phase_space as ps
    subs = processes.subprocesses()
    for plan in (plans.gluon_plan(5), plans.quark_line_plan(5), subs[len(subs) // 2].plan):
        points = 37
        mom, _ = ps.generate(plan.legs, points, seed=9)
        ext = ps.external_planes(mom, plan.crossing)
        got = refcpu.fill_forms(plan.arrays(), ext, points)
        assert got.shape == (len(plan.forms), 4, points)
        for f, form in enumerate(plan.forms):
            want = 0.0
            for leg, c in form:
                want = want + c * ext[leg * 4:(leg + 1) * 4]
            assert np.array_equal(got[f].view(np.uint64), np.asarray(want).view(np.uint64)), (plan.name, f, form)
    plan = plans.gluon_plan(5)
    mom, _ = ps.generate(5, 8, seed=9)
    ext = ps.external_planes(mom, plan.crossing)
    got = refcpu.fill_forms(plan.arrays(), ext, 8)
    legs = [ext[4 * l:4 * l + 4] for l in range(5)]
    for f, form in enumerate(plan.forms):
        members = sorted(leg for leg, _ in form)
        assert all(c == 1.0 for _, c in form), form
        # a cyclically consecutive run of legs
        assert any(members == sorted((a + k) % 5 for k in range(len(members))) for a in range(5)), members
        assert np.allclose(got[f], sum(legs[l] for l in members), rtol=0, atol=1e-9)
    # conservation: the form of legs 1..n-1 is minus the crossed momentum of leg 0
    full = sum(legs[leg] for leg in range(1, 5))
    assert np.allclose(full, -legs[0], atol=1e-9)


def test_source_fixture_is_current():
    """tests/golden/source_cases.npz (what the GPU test loads) is what its script writes now"""
    import importlib.util
    here = os.path.join(os.path.dirname(__file__), "golden")
    spec = importlib.util.spec_from_file_location("make_source_cases", os.path.join(here, "make_source_cases.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    want = m.build()
    got = np.load(os.path.join(here, "source_cases.npz"))
    for k, v in want.items():
        assert np.array_equal(got[k], v), k
