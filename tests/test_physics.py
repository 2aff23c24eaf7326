# SPDX-License-Identifier: 0BSD
"""The generated plans (pacuda_amd/plans.py) run on the CPU oracle against closed forms.

- every MHV and anti-MHV helicity configuration of 4, 5 and 6 gluons equals Parke-Taylor,
  |A|^2 = s_ij^4 / |s_12 s_23 ... s_n1|, to 1e-9 relative (g = 1);
- all-plus and one-minus configurations vanish (to 1e-20 of the point's total);
- the colour-ordered amplitudes obey reflection, |A(1..n)|^2 = |A(n..1)|^2, and the U(1)
  decoupling identity sum over cyclic insertions of leg 1 of A = 0, both on NMHV configurations
  of six gluons, which Parke-Taylor does not cover;
- the plan's words decode back to its rows (convert/src/program.rs:60-162).
A wrong sign or constant in the three- or four-gluon vertex, a swapped operand in an op, or a
dropped propagator fails the Parke-Taylor check."""
import numpy as np
import pytest

from pacuda_amd import plans, phase_space as ps
import harness
import refcpu


def _mdot(a, b):
    return a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2] - a[..., 3] * b[..., 3]


def _run(plan, mom, kernel=1):
    ext = ps.external_planes(mom, plan.crossing)
    res, tot = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, mom.shape[0], kernel=kernel)
    return ext, res, tot


@pytest.mark.parametrize("n", [4, 5, 6])
def test_parke_taylor(n):
    plan = plans.gluon_plan(n)
    mom, _ = ps.generate(n, 48, seed=100 + n)
    ext, res, tot = _run(plan, mom)
    P = mom.shape[0]
    q = ext.reshape(n, 4, P).transpose(2, 0, 1)
    s = np.empty((P, n, n))
    for i in range(n):
        for j in range(n):
            s[:, i, j] = 2 * _mdot(q[:, i], q[:, j])
    den = np.abs(np.prod([s[:, k, (k + 1) % n] for k in range(n)], axis=0))
    checked = 0
    for o, hel in enumerate(plan.helicities):
        out = [(-h if plan.crossing[l][1] < 0 else h) for l, h in enumerate(hel)]  # outgoing helicities
        neg = [l for l in range(n) if out[l] < 0]
        pos = [l for l in range(n) if out[l] > 0]
        if len(neg) == 2 or len(pos) == 2:
            i, j = neg if len(neg) == 2 else pos
            want = s[:, i, j] ** 4 / den
            assert np.allclose(res[:, o], want, rtol=1e-9, atol=0), (hel, np.max(np.abs(res[:, o] / want - 1)))
            checked += 1
        elif len(neg) <= 1 or len(pos) <= 1:
            assert np.all(np.abs(res[:, o]) <= 1e-20 * np.abs(tot)), hel
    assert checked == (2 * n * (n - 1) // 2 if n > 4 else 6)


def _amplitudes(plan, mom):
    """complex amplitudes of every helicity configuration (v1 emulation, kept amplitudes)"""
    ext = ps.external_planes(mom, plan.crossing)
    lib = refcpu.load()
    lib.refcpu_reset()
    v1 = harness.V1(lib, prefix="refcpu_v1_")
    assert v1.configure_kernel(1) == 0
    re, im = v1.run_plan(plan, ext, mom.shape[0], keep=False)
    lib.refcpu_reset()
    return re + 1j * im  # [amplitude plane][point]


def test_reflection_and_u1_decoupling_six_gluons():
    n = 6
    plan = plans.gluon_plan(n)
    mom, _ = ps.generate(n, 16, seed=7)
    # the amplitude of ordering sigma: permute the momenta so that plan leg k reads leg sigma[k]
    def amps(order):
        return _amplitudes(plan, mom[:, list(order), :])
    # helicity configurations are the source-code helicities of the plan's legs; permuting
    # momenta permutes which physical leg each plan leg carries (crossing signs follow legs 0, 1)
    base = amps(range(n))
    nmhv = [o for o, h in enumerate(plan.helicities) if sum(1 for x in h if x < 0) == 3]
    assert nmhv
    # reflection: A(0, 5, 4, 3, 2, 1) has the same modulus as A(0..5) for the reflected labels
    # (crossing is applied by plan leg, so reflect only among the outgoing legs 2..5 and keep
    # the incoming 0, 1: A(0,1,2,3,4,5) and A(1,0,5,4,3,2) are reflections of each other)
    order = [1, 0, 5, 4, 3, 2]
    plan_r = plans.gluon_plan(n, incoming=(0, 1))
    refl = _amplitudes(plan_r, mom[:, order, :])
    # the momenta at plan leg k are physical leg order[k]; their crossing: legs 0, 1 incoming
    floor = 1e-10 * np.max(np.abs(base))  # vanishing configurations are roundoff
    for o, hel in enumerate(plan.helicities):
        hel_r = tuple(hel[order[k]] for k in range(n))
        o_r = plan.helicities.index(hel_r)
        assert np.allclose(np.abs(base[o]), np.abs(refl[o_r]), rtol=1e-9, atol=floor), hel
    # U(1) decoupling: sum over the insertions of physical leg 1 into (0, 2, 3, 4, 5) cyclically
    others = [0, 2, 3, 4, 5]
    total, largest = None, None
    for pos in range(1, n):
        order = others[:pos] + [1] + others[pos:]
        pl = plans.gluon_plan(n, incoming=tuple(k for k in range(n) if order[k] in (0, 1)))
        a = _amplitudes(pl, mom[:, order, :])
        idx = [pl.helicities.index(tuple(hel[order[k]] for k in range(n))) for hel in plan.helicities]
        a = a[idx]
        total = a if total is None else total + a
        largest = np.abs(a) if largest is None else np.maximum(largest, np.abs(a))
    for o in nmhv:
        # relative to the largest term of the sum, point by point
        assert np.all(np.abs(total[o]) <= 1e-11 * largest[o] + floor), plan.helicities[o]


@pytest.mark.parametrize("n", [4, 5, 6, 7])
def test_words_round_trip(n):
    plan = plans.gluon_plan(n)
    width = plan.widths()
    words = plan.words()
    assert len(words) == len(plan.rows)
    for w, r in zip(words, plan.rows):
        assert w < (1 << 64)
        op, fields = plans.decode_fields(w, width)
        if r[0] == "contrib":
            assert op == r[1] and fields == list(r[2:])
        elif r[0] == "prop":
            assert op == r[1] and fields == list(r[2:])
        elif r[0] == "source":
            assert op == 12 and fields == list(r[1:])
        elif r[0] == "closure":
            assert op == 11 and fields == list(r[1:])
        elif r[0] == "identity":
            assert op == 10 and fields == list(r[1:])


def test_plan_sizes_five_gluons():
    plan = plans.gluon_plan(5)
    c = plan.counts()
    assert (c[6], c[5], c[4], c["identity"], c["prop"], c["closure"], c["source"]) == (92, 44, 96, 28, 28, 32, 10)
    assert plan.current_planes == 384 and plan.amplitude_planes == 32 and len(plan.forms) == 10


@pytest.mark.parametrize("n", [4, 5, 6])
def test_quark_line_mhv(n):
    """q qbar + (n-2) gluons: every non-vanishing helicity amplitude equals the MHV closed form
    |s_{f- j}|^3 |s_{f+ j}| / |s_01 s_12 ... s_(n-1)0| (j the gluon of odd helicity, f-/f+ the
    fermions of negative/positive helicity; Dixon, hep-ph/9601359), and the configurations with
    all gluons of one helicity vanish; both beams crossed as for q qbar -> gluons."""
    plan = plans.quark_line_plan(n, incoming=(0, n - 1))
    mom, _ = ps.generate(n, 32, seed=200 + n)
    # plan legs: 0 antiquark slot <- beam 0, n-1 quark slot <- beam 1, gluons <- the outgoing partons
    order = [0] + list(range(2, n)) + [1]
    mom = mom[:, order, :]
    ext, res, tot = _run(plan, mom)
    P = mom.shape[0]
    q = ext.reshape(n, 4, P).transpose(2, 0, 1)
    s = np.array([[2 * _mdot(q[:, i], q[:, j]) for j in range(n)] for i in range(n)])
    den = np.abs(np.prod([s[k, (k + 1) % n] for k in range(n)], axis=0))
    checked = 0
    for o, hel in enumerate(plan.helicities):
        out = [(-h if plan.crossing[l][1] < 0 else h) for l, h in enumerate(hel)]
        fneg = 0 if out[0] < 0 else n - 1
        fpos = n - 1 - fneg
        assert out[0] == -out[n - 1]
        gl = out[1:n - 1]
        negg = [l + 1 for l, h in enumerate(gl) if h < 0]
        posg = [l + 1 for l, h in enumerate(gl) if h > 0]
        if len(negg) == 1:
            j = negg[0]
            want = np.abs(s[fneg, j]) ** 3 * np.abs(s[fpos, j]) / den
        elif len(posg) == 1:
            j = posg[0]
            want = np.abs(s[fpos, j]) ** 3 * np.abs(s[fneg, j]) / den
        elif not negg or not posg:
            assert np.all(np.abs(res[:, o]) <= 1e-20 * np.abs(tot)), hel
            continue
        else:
            continue
        assert np.allclose(res[:, o], want, rtol=1e-9, atol=0), (hel, np.max(np.abs(res[:, o] / want - 1)))
        checked += 1
    assert checked >= 4


def test_pp_jjj_subprocesses_closed_forms():
    """every subprocess of pacuda_amd/processes.py, with its crossing: each non-vanishing
    helicity amplitude equals Parke-Taylor (gluons), the quark-line MHV form or the two-line
    form (four quarks and a gluon)"""
    from pacuda_amd import processes
    subs = processes.subprocesses()
    assert len(subs) == 29 + 152
    assert len({sp.name for sp in subs}) == len(subs)
    mom, _ = ps.generate(5, 8, seed=300)
    n = 5
    for sp in subs:
        plan = sp.plan
        ext, res, tot = _run(plan, mom)
        assert np.all(tot > 0), sp.name
        q = ext.reshape(n, 4, 8).transpose(2, 0, 1)
        s = np.array([[2 * _mdot(q[:, i], q[:, j]) for j in range(n)] for i in range(n)])
        den = np.abs(np.prod([s[k, (k + 1) % n] for k in range(n)], axis=0))
        if plan.name.startswith("qqbarQQbar"):
            want = _fermion_pair_closed_form(plan, ext, 8)
            assert np.allclose(res, want, rtol=1e-9, atol=0), sp.name
            continue
        quark = plan.name.startswith("qqbar")
        for o, hel in enumerate(plan.helicities):
            out = [(-h if plan.crossing[l][1] < 0 else h) for l, h in enumerate(hel)]
            if quark:
                fneg = 0 if out[0] < 0 else 4
                fpos = 4 - fneg
                negg = [l + 1 for l, h in enumerate(out[1:4]) if h < 0]
                posg = [l + 1 for l, h in enumerate(out[1:4]) if h > 0]
                if len(negg) == 1:
                    want = np.abs(s[fneg, negg[0]]) ** 3 * np.abs(s[fpos, negg[0]]) / den
                elif len(posg) == 1:
                    want = np.abs(s[fpos, posg[0]]) ** 3 * np.abs(s[fneg, posg[0]]) / den
                else:
                    want = np.zeros(8)
            else:
                neg = [l for l in range(n) if out[l] < 0]
                pos = [l for l in range(n) if out[l] > 0]
                if len(neg) == 2 or len(pos) == 2:
                    i, j = neg if len(neg) == 2 else pos
                    want = s[i, j] ** 4 / den
                else:
                    want = np.zeros(8)
            assert np.allclose(res[:, o], want, rtol=1e-9, atol=1e-18 * np.max(tot)), (sp.name, hel)


@pytest.mark.parametrize("n", [4, 5])
def test_surrogate_prior_is_exact_below_six_gluons(n):
    """every 4- and 5-gluon helicity amplitude is MHV or anti-MHV, so the helicity-summed
    leading-colour |M|^2 is a constant times sum_{a<b} s_ab^4 / prod s_{l,l+1}: the oracle's plan
    run divided by exp(phase_space.log_prior) is the same number at every point"""
    plan = plans.gluon_plan(n)
    mom, _ = ps.generate(n, 300, 17)
    ext = ps.external_planes(mom, plan.crossing)
    _, tot = refcpu.evaluate(plan.arrays(), harness.plan_counts(plan), ext, len(mom), kernel=1, resolved=False)
    ratio = tot / np.exp(ps.log_prior(mom, plan.crossing))
    assert np.std(ratio) < 1e-9 * np.mean(ratio)
    # and at six gluons (NMHV configurations) it is not
    plan6 = plans.gluon_plan(6)
    mom6, _ = ps.generate(6, 300, 17)
    _, tot6 = refcpu.evaluate(plan6.arrays(), harness.plan_counts(plan6), ps.external_planes(mom6, plan6.crossing),
                              len(mom6), kernel=1, resolved=False)
    r6 = np.log(tot6) - ps.log_prior(mom6, plan6.crossing)
    assert 0.05 < np.std(r6) < 0.5


def _fermion_pair_closed_form(plan, ext, points):
    """|A|^2 per output of plans.four_quark_plan from its closed form (outgoing helicities)"""
    n = plan.legs
    q = ext.reshape(n, 4, points).transpose(2, 0, 1)
    s = np.array([[2 * _mdot(q[:, i], q[:, j]) for j in range(n)] for i in range(n)])
    qbar_m, q_s, qbar_s, q_m = (0, 2, 3, 4) if n == 5 else (0, 1, 2, 3)
    fermions = (qbar_m, q_s, qbar_s, q_m)
    want = []
    for hel in plan.helicities:
        out = [(-h if plan.crossing[l][1] < 0 else h) for l, h in enumerate(hel)]
        # each line keeps its helicity: its quark and antiquark are of opposite helicity
        assert out[q_m] == -out[qbar_m] and out[q_s] == -out[qbar_s]
        if n == 4:
            a, b = [l for l in fermions if out[l] < 0]
            want.append(s[a, b] ** 2 / s[q_s, qbar_s] ** 2)
        else:
            sign = -1 if out[1] > 0 else 1  # MHV: the negative fermions; anti-MHV: the positive ones
            a, b = [l for l in fermions if sign * out[l] > 0]
            den = np.abs(np.prod([s[k, (k + 1) % n] for k in range(n)], axis=0))
            want.append(s[a, b] ** 2 * np.abs(s[qbar_m, q_s] * s[qbar_s, q_m]) / den)
    return np.array(want).T


@pytest.mark.parametrize("gluon", [False, True])
def test_four_quark_closed_forms(gluon):
    """two quark lines of different flavour (and a gluon): every helicity amplitude equals the
    closed form of plans.four_quark_plan's docstring -- with no gluon s_ab^2 / s_{S pair}^2 (one
    gluon exchanged), with one the MHV form s_ab^2 s(qbar_M, q_S) s(qbar_S, q_M) / prod s_{i,i+1}
    -- with two legs crossed as for q qbar -> Q Qbar (g)"""
    n = 5 if gluon else 4
    incoming = (0, 3) if gluon else (0, 2)
    plan = plans.four_quark_plan(incoming=incoming, gluon=gluon)
    mom, _ = ps.generate(n, 16, seed=400 + n)
    # plan legs <- physical legs: the beams into the crossed slots
    order = [0, 4, 2, 1, 3] if gluon else [0, 2, 1, 3]
    mom = mom[:, order, :]
    ext, res, tot = _run(plan, mom)
    want = _fermion_pair_closed_form(plan, ext, mom.shape[0])
    assert res.shape == want.shape == (16, 8 if gluon else 4)
    assert np.allclose(res, want, rtol=1e-10, atol=0), np.max(np.abs(res / want - 1))


def test_four_quark_gauge_invariance():
    """the colour-ordered amplitude vanishes when the gluon's polarisation is replaced by its
    momentum (the current k made by a symjit row copying the gluon's form), for every helicity
    configuration of the fermions: the relative signs of the gluon on the main line, on the S
    line and on the exchanged gluon (three-gluon vertex) are fixed by this"""
    plan = plans.four_quark_plan(incoming=(0, 3))
    mom, _ = ps.generate(5, 8, seed=410)
    mom = mom[:, [0, 4, 2, 1, 3], :]
    ext = ps.external_planes(mom, plan.crossing)
    gauge = plans.four_quark_plan(incoming=(0, 3))
    rows = [r for r in gauge.rows if not (r[0] == "source" and r[5] & 7 == plans.FAMILY_VECTOR)]
    gluon_bases = [r[1] for r in gauge.rows if r[0] == "source" and r[5] & 7 == plans.FAMILY_VECTOR]
    n_src = sum(1 for r in rows if r[0] == "source")
    codes = []
    for m in range(4):  # (k_m, +0) pairs: the form's component and the real factor's imaginary part
        codes += [plans.binding_code(plans.SB_MOMENTUM, 0, 0, m), plans.binding_code(plans.SB_FACTOR, 0, 1, 0)]
    k_rows = [("binding", plans.KERNEL_BINDING_PROBE, codes, [0, 1, 2, 3], 1)]
    for base in gluon_bases:
        k_rows.append(("symjit", 1, 0, 0, base, gauge.form([1]), gauge.form([1]), gauge.factor(1.0)))
    gauge.rows = rows[:n_src] + k_rows + rows[n_src:]

    def amplitudes(p):
        lib = refcpu.load()
        lib.refcpu_reset()
        v1 = harness.V1(lib, prefix="refcpu_v1_")
        assert v1.configure_kernel(1) == 0
        are, aim = v1.run_plan(p, ext, ext.shape[1], keep=False)
        lib.refcpu_reset()
        return are + 1j * aim

    a, ak = amplitudes(plan), amplitudes(gauge)
    energy = np.abs(ext[4])  # the gluon's energy (leg 1)
    assert np.max(np.abs(ak) / (np.max(np.abs(a), axis=0) * energy)) < 1e-12
    assert np.min(np.max(np.abs(a), axis=0)) > 0
