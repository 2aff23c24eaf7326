# SPDX-License-Identifier: 0BSD
"""Recurrence plans in the runner's op vocabulary, and their 64-bit words.

The reference records its plans from pyAmpliCol/RustiCol LC artifacts (convert/src/record.rs),
neither of which exists here; this module generates the same kind of plan itself: a
colour-ordered Berends-Giele recursion over all helicity configurations of one colour ordering,
written with the rows of runner/recurrence_runner.cu:15-36 (legend convert/CONVENTIONS.md):

  op 12 source        the external wavefunction of a leg and helicity (HELAS conventions)
  op 6  three-vector  the colour-ordered three-gluon vertex of two sub-currents
  op 5  wedge         J1 ^ J2, the auxiliary antisymmetric-tensor current of the four-gluon vertex
  op 10 identity      the tensor's constant propagator (i/2)
  op 4  tensor-vector T . J, the tensor's coupling back to a vector current
  op 9  propagator    -i / P^2 (Feynman gauge)
  op 11 closure       A = eps(leg 0) . K(1..n-1), K the amputated current of the other legs

The four-gluon vertex V4(J1, J2, J3) = (i/2)[2 (J1.J3) J2 - (J1.J2) J3 - (J2.J3) J1] equals
(i/2)[op4(J1^J2, J3) - op4(J2^J3, J1)] (op4(A^B, C) = B (A.C) - A (B.C)), so the sum over the
three-way splits of a range becomes tensor currents of the sub-ranges. The three-gluon vertex in
Dixon's colour-ordered normalisation is (i/sqrt 2) times minus op 6 for conserved currents.
With these constants |A|^2 of every MHV configuration equals the Parke-Taylor value
s_ij^4 / |s_12 s_23 ... s_n1| (g = 1); tests/test_physics.py checks it.

Word layout and widths follow runner/recurrence_runner.cu:10-36 and convert/src/program.rs:117-162
(widths as RustiCol sets them: rusticol/patches/0004-...patch:498-504)."""
import itertools
import math

import numpy as np

SQRT_HALF = math.sqrt(0.5)

# the 10-bit source template (runner/recurrence_runner.cu:25-28)
FAMILY_SCALAR, FAMILY_WEYL, FAMILY_DIRAC, FAMILY_VECTOR = 0, 1, 2, 3


def source_code(family, orientation=0, helicity=0, chirality=0):
    return family | orientation << 3 | (helicity + 2) << 5 | (chirality + 1) << 8


def bits(v):
    """64 - clz(max(v, 1)): the width of a field holding values below v (or v itself)."""
    return max(int(v), 1).bit_length()


class Plan:
    """A plan run template: rows, table sizes, constants, forms, reduction, helicity labels."""

    def __init__(self, legs):
        self.legs = legs
        self.rows = []
        self.current_planes = 0
        self.amplitude_planes = 0
        self.forms = []          # per form: [(leg, coefficient)]
        self._form_index = {}
        self.scales = []         # complex
        self.factors = []        # complex
        self._factor_index = {}
        self.outputs = []        # per output: [(amplitude plane, weight)]
        self.helicities = []     # per output: the helicity of each leg (source-code helicities)
        self.numerator = 1
        self.denominator = 1
        self.crossing = None     # per leg: (input leg, sign)
        self.name = ""

    # ---- tables ----
    def current(self, components):
        base = self.current_planes
        self.current_planes += components
        return base

    def amplitude(self):
        a = self.amplitude_planes
        self.amplitude_planes += 1
        return a

    def form(self, legs):
        key = tuple(legs)
        if key not in self._form_index:
            self._form_index[key] = len(self.forms)
            self.forms.append([(leg, 1.0) for leg in key])
        return self._form_index[key]

    def scale(self, value):
        self.scales.append(complex(value))
        return len(self.scales) - 1

    def factor(self, value):
        value = complex(value)
        key = (value.real, value.imag)
        if key not in self._factor_index:
            self._factor_index[key] = len(self.factors)
            self.factors.append(value)
        return self._factor_index[key]

    def factor_run(self, values):
        """consecutive factors (a closure's component factors), not shared"""
        first = len(self.factors)
        for v in values:
            self.factors.append(complex(v))
        return first

    # ---- words ----
    def widths(self):
        w = [0] * 8
        w[0] = bits(self.current_planes)
        w[1] = bits(max(self.amplitude_planes - 1, 0))
        w[2] = bits(max(len(self.forms) - 1, 0))
        w[3] = bits(max(len(self.factors) - 1, 0))
        w[4] = bits(max(len(self.scales) - 1, 0))
        w[5] = 4
        return w

    def words(self):
        return encode(self.rows, self.widths())

    # ---- the arrays the runner takes ----
    def arrays(self):
        """numpy arrays of the tables (scales, factors, forms, reduction)."""
        form_terms, slots, coeffs = [], [], []
        for f in self.forms:
            form_terms += [len(slots), len(f)]
            for leg, c in f:
                slots.append(leg)
                coeffs.append(c)
        offsets, planes, weights = [0], [], []
        for out in self.outputs:
            for p, w in out:
                planes.append(p)
                weights.append(w)
            offsets.append(len(planes))
        return dict(
            words=np.array(self.words(), dtype=np.uint64),
            width=np.array(self.widths(), dtype=np.uint8),
            scale_re=np.array([s.real for s in self.scales]), scale_im=np.array([s.imag for s in self.scales]),
            factor_re=np.array([f.real for f in self.factors]), factor_im=np.array([f.imag for f in self.factors]),
            form_terms=np.array(form_terms, dtype=np.uint32), term_slots=np.array(slots, dtype=np.uint32),
            term_coefficients=np.array(coeffs, dtype=np.float64),
            offsets=np.array(offsets, dtype=np.uint32), planes=np.array(planes, dtype=np.uint32),
            weights=np.array(weights, dtype=np.float64),
        )

    def counts(self):
        ops = {}
        for r in self.rows:
            key = r[1] if r[0] == "contrib" else r[0]
            ops[key] = ops.get(key, 0) + 1
        return ops


def encode(rows, width):
    """rows -> 64-bit words, fields LSB first after the 4-bit op (convert/src/program.rs:117-162)."""
    P, A, F, X, S, K = width[:6]
    words = []
    for n, r in enumerate(rows):
        kind = r[0]
        if kind == "clear":
            op, fields = 1, [(r[1], P), (r[2], P)]
        elif kind == "contrib":
            _, op, init, slot, p0, p1, dst, f0, f1, x = r
            fields = [(init, 1), (slot, S), (p0, P), (p1, P), (dst, P), (f0, F), (f1, F), (x, X)]
        elif kind == "prop":
            _, op, base, form, x = r
            fields = [(base, P), (form, F), (x, X)]
        elif kind == "identity":
            op, fields = 10, [(r[1], P), (r[2], K), (r[3], X)]
        elif kind == "closure":
            _, p0, p1, amp, x, first, count = r
            op, fields = 11, [(p0, P), (p1, P), (amp, A), (x, X), (first, X), (count, K)]
        elif kind == "source":
            _, dst, form, x, mass, code = r
            op, fields = 12, [(dst, P), (form, F), (x, X), (mass, S), (code, 10)]
        elif kind == "symjit":
            _, init, p0, p1, dst, f0, f1, x = r
            op, fields = 13, [(init, 1), (p0, P), (p1, P), (dst, P), (f0, F), (f1, F), (x, X)]
        elif kind == "binding":
            # kernel, input codes, output components, direct: the word, then the codes two per
            # data word, low half first
            _, kernel, codes, components, direct = r
            op, fields = 14, [(kernel, 16), (len(codes), 8), (len(components), 8), (direct, 1)]
            data = list(codes) + list(components)
        else:
            raise ValueError(f"row {n}: {kind}")
        word, at = op, 4
        for value, b in fields:
            if value >> b:
                raise ValueError(f"row {n}: {value} does not fit {b} bits")
            word |= int(value) << at
            at += b
        if at > 64:
            raise ValueError(f"row {n}: {at} bits")
        words.append(word)
        if kind == "binding":
            for k in range(0, len(data), 2):
                lo = data[k]
                hi = data[k + 1] if k + 1 < len(data) else 0
                words.append(int(lo) | int(hi) << 32)
    return words


# a symjit binding's input codes (runner/recurrence_runner.cu:580-582): kind, which operand,
# imaginary part, component
SB_PARENT, SB_MOMENTUM, SB_DESTINATION, SB_SCALE, SB_FACTOR = range(5)


def binding_code(kind, which=0, imag=0, index=0):
    return kind | which << 3 | imag << 4 | index << 5


# the generated kernels' indices (the sorted file names of pacuda_amd/kernels/*.pk)
KERNEL_BINDING_PROBE, KERNEL_FERMION_PAIR_SIGMA, KERNEL_FERMION_PAIR_SIGMABAR = range(3)


def weyl_pair_codes():
    """the inputs of the fermion-pair kernels: re, im of parent 0's two components, then parent 1's"""
    return [binding_code(SB_PARENT, w, im, c) for w in (0, 1) for c in (0, 1) for im in (0, 1)]


def decode_fields(word, width):
    """the op and fields of one word (for tests); inverse of encode for ops 1-12"""
    P, A, F, X, S, K = width[:6]
    op = word & 15
    w = word >> 4
    layout = {1: [P, P], 10: [P, K, X], 11: [P, P, A, X, X, K], 12: [P, F, X, S, 10]}
    for o in range(2, 7):
        layout[o] = [1, S, P, P, P, F, F, X]
    for o in range(7, 10):
        layout[o] = [P, F, X]
    out = []
    for b in layout[op]:
        out.append(w & ((1 << b) - 1))
        w >>= b
    return op, out


# ---- n-gluon colour-ordered plan ------------------------------------------------------------

def gluon_plan(n, incoming=(0, 1), weight=1.0):
    """The leading-colour plan of the colour ordering (0, 1, ..., n-1) of n gluons, every helicity.

    Leg 0 closes the recursion. Output o has helicities h[leg] = +1 if bit (n-1-leg) of o is set,
    else -1 (source-code helicities: an incoming leg's physical helicity; the source flips it for
    the negative energy of the crossed momentum). Output weights are `weight`."""
    if n < 3:
        raise ValueError("n >= 3")
    p = Plan(n)
    p.name = f"{n}g"
    p.crossing = [(leg, -1.0 if leg in incoming else 1.0) for leg in range(n)]
    mass0 = p.scale(0.0)
    g3 = p.scale(complex(0.0, -SQRT_HALF))   # (i/sqrt 2) x (-1): op 6 is minus the vertex bracket
    unit = p.scale(1.0)
    one = p.factor(1.0)
    minus_one = p.factor(-1.0)
    tensor_prop = p.factor(complex(0.0, 0.5))  # the four-gluon vertex's i/2
    hels = (-1, 1)
    rows = p.rows
    # sources: every leg, both helicities
    J = {}
    for leg in range(n):
        f = p.form([leg])
        for h in hels:
            base = p.current(4)
            rows.append(("source", base, f, one, mass0, source_code(FAMILY_VECTOR, 0, h, 0)))
            J[(leg, leg, (h,))] = base
    first, last = 1, n - 1
    _gluon_currents(p, rows, J, first, last, True, unit, g3, one, minus_one, tensor_prop)
    # closures: A(h0, H) = eps0(h0) . K(1..n-1, H), metric (+, -, -, -) as component factors
    metric = p.factor_run([1.0, -1.0, -1.0, -1.0])
    amp_of = {}
    for h0 in hels:
        for H in itertools.product(hels, repeat=n - 1):
            a = p.amplitude()
            amp_of[(h0,) + H] = a
            rows.append(("closure", J[(0, 0, (h0,))], J[(first, last, H)], a, one, metric, 4))
    for o in range(2 ** n):
        hel = tuple(1 if (o >> (n - 1 - leg)) & 1 else -1 for leg in range(n))
        p.outputs.append([(amp_of[hel], weight)])
        p.helicities.append(hel)
    return p


def _gluon_currents(p, rows_out, J, first, last, full_is_amputated, unit, g3, one, minus_one, tensor_prop):
    """Berends-Giele gluon currents J[(a, b, H)] of every range inside legs first..last (sources
    already in J), level by level: tensor currents of sub-ranges (op 5 + op 10), vector currents
    from three-vertex splits (op 6) and tensor-vector terms (op 4), propagators (op 9). The full
    range is propagated unless `full_is_amputated`."""
    hels = (-1, 1)
    T = {}
    for L in range(2, last - first + 2):
        contrib, identities, props = [], [], []
        full = L == last - first + 1
        for a in range(first, last - L + 2):
            b = a + L - 1
            for H in itertools.product(hels, repeat=L):
                if not full:
                    t = p.current(6)
                    init = 1
                    for m in range(a, b):
                        hl, hr = H[: m - a + 1], H[m - a + 1:]
                        contrib.append(("contrib", 5, init, unit, J[(a, m, hl)], J[(m + 1, b, hr)], t, 0, 0, one))
                        init = 0
                    T[(a, b, H)] = t
                    identities.append(("identity", t, 6, tensor_prop))
                j = p.current(4)
                init = 1
                for m in range(a, b):
                    hl, hr = H[: m - a + 1], H[m - a + 1:]
                    contrib.append(("contrib", 6, init, g3, J[(a, m, hl)], J[(m + 1, b, hr)], j,
                                    p.form(range(a, m + 1)), p.form(range(m + 1, b + 1)), one))
                    init = 0
                for m in range(a, b):
                    hl, hr = H[: m - a + 1], H[m - a + 1:]
                    if m - a + 1 >= 2:
                        contrib.append(("contrib", 4, 0, unit, T[(a, m, hl)], J[(m + 1, b, hr)], j, 0, 0, one))
                    if b - m >= 2:
                        contrib.append(("contrib", 4, 0, unit, T[(m + 1, b, hr)], J[(a, m, hl)], j, 0, 0, minus_one))
                J[(a, b, H)] = j
                if not (full and full_is_amputated):
                    props.append(("prop", 9, j, p.form(range(a, b + 1)), one))
        rows_out += contrib + identities + props
    return T


QUARK_GLUON = SQRT_HALF  # the colour-ordered quark-gluon coupling (Dixon's normalisation, g = 1)


def quark_line_plan(n, incoming=(0, 1), weight=1.0):
    """The colour ordering (qbar, g_1, ..., g_{n-2}, q) of one massless quark line and n-2 gluons,
    every helicity configuration, as Weyl currents (ops 2/3, 7/8; convert/CONVENTIONS.md:80-90).

    Legs in the all-outgoing convention: 0 the antiquark (closing the recursion), 1..n-2 the
    gluons, n-1 the quark; `incoming` legs get crossed (negative-energy) momenta. The fermion
    current is a row spinor grown from the quark: a line of chirality +1 alternates vertex op 2
    (sigma) and propagator op 8 (sigma-bar); chirality -1 alternates op 3 and op 7; the antiquark
    closes with the opposite chirality (op 11 over 2 components, factors 1). With the coupling
    1/sqrt 2 every non-vanishing amplitude has |A|^2 = |s_{f- j}|^3 |s_{f+ j}| / |s_01 s_12 ...
    s_(n-1)0| (MHV, j the odd gluon; tests/test_physics.py). Output o: chirality -1 for o < 2^(n-2),
    then +1; within, gluon helicities as bits (leg 1 most significant)."""
    if n < 3:
        raise ValueError("n >= 3")
    p = Plan(n)
    p.name = f"qqbar+{n - 2}g"
    p.crossing = [(leg, -1.0 if leg in incoming else 1.0) for leg in range(n)]
    mass0 = p.scale(0.0)
    g3 = p.scale(complex(0.0, -SQRT_HALF))
    unit = p.scale(1.0)
    gq = p.scale(QUARK_GLUON)
    one = p.factor(1.0)
    minus_one = p.factor(-1.0)
    tensor_prop = p.factor(complex(0.0, 0.5))
    rows = p.rows
    J = {}
    for leg in range(1, n - 1):
        f = p.form([leg])
        for h in (-1, 1):
            base = p.current(4)
            rows.append(("source", base, f, one, mass0, source_code(FAMILY_VECTOR, 0, h, 0)))
            J[(leg, leg, (h,))] = base
    quark, antiquark, code_hel = {}, {}, {}
    for ch in (-1, 1):
        hq = -ch if (n - 1) in incoming else ch
        base = p.current(2)
        rows.append(("source", base, p.form([n - 1]), one, mass0, source_code(FAMILY_WEYL, 0, hq, ch)))
        quark[ch] = base
        cb = -ch
        hb = -cb if 0 in incoming else cb
        base = p.current(2)
        rows.append(("source", base, p.form([0]), one, mass0, source_code(FAMILY_WEYL, 1, hb, cb)))
        antiquark[ch] = base
        code_hel[ch] = (hb, hq)
    if n > 3:
        _gluon_currents(p, rows, J, 1, n - 2, False, unit, g3, one, minus_one, tensor_prop)
    psi = {(n - 1, ch, ()): quark[ch] for ch in (-1, 1)}
    for a in range(n - 2, 0, -1):
        contrib, props = [], []
        for ch in (-1, 1):
            vertex, propagator = (2, 8) if ch == 1 else (3, 7)
            for H in itertools.product((-1, 1), repeat=n - 1 - a):
                d = p.current(2)
                init = 1
                for m in range(a, n - 1):
                    hl, hr = H[: m - a + 1], H[m - a + 1:]
                    contrib.append(("contrib", vertex, init, gq, psi[(m + 1, ch, hr)], J[(a, m, hl)], d, 0, 0, one))
                    init = 0
                psi[(a, ch, H)] = d
                if a > 1:
                    props.append(("prop", propagator, d, p.form(range(a, n)), one))
        rows += contrib + props
    components = p.factor_run([1.0, 1.0])
    for ch in (-1, 1):
        for H in itertools.product((-1, 1), repeat=n - 2):
            amp = p.amplitude()
            rows.append(("closure", antiquark[ch], psi[(1, ch, H)], amp, one, components, 2))
            p.outputs.append([(amp, weight)])
            hb, hq = code_hel[ch]
            p.helicities.append((hb,) + H + (hq,))
    return p


# ---- two quark lines (four quarks and at most one gluon) ------------------------------------

# the symjit row factor of the fermion-pair current: the vertex's i times the coupling, so that
# the current replaces a gluon's polarisation in the closure of its line
FERMION_PAIR = 1j * QUARK_GLUON


def _fermion_pair(rows, chirality, left, right, dst, factor, init):
    """K = factor * (l^T sigma^mu r) (chirality +1, the line of vertex op 2) or l^T sigmabar^mu r
    (chirality -1, op 3): a symjit row (op 13) after its binding (op 14)"""
    kernel = KERNEL_FERMION_PAIR_SIGMA if chirality == 1 else KERNEL_FERMION_PAIR_SIGMABAR
    rows.append(("binding", kernel, weyl_pair_codes(), [0, 1, 2, 3], 0))
    rows.append(("symjit", init, left, right, dst, 0, 0, factor))


def four_quark_plan(incoming=(0, 3), gluon=True, weight=1.0):
    """The colour ordering (qbar_M, g, q_S, qbar_S, q_M) of two massless quark lines of different
    flavour (M, the main line, and S) and one gluon, every helicity configuration; without the
    gluon (gluon=False) the four-quark ordering (qbar_M, q_S, qbar_S, q_M).

    Legs (all outgoing; `incoming` legs get crossed momenta): 0 qbar_M, 1 the gluon, 2 q_S,
    3 qbar_S, 4 q_M (without the gluon: 0, 1 q_S, 2 qbar_S, 3 q_M). The main line's current grows
    from q_M as in quark_line_plan and absorbs the vector currents of the elements to its left:
    the gluon's, the S pair's and that of the gluon and the S pair together. The S pair's vector
    current is its fermion-pair current (a symjit row: the lines' two Weyl currents, the kernel
    l^T sigma^mu r or its sigmabar twin, factor i/sqrt 2), propagated by op 9; with the gluon it
    also takes the S line with the gluon absorbed (its quark propagated) and the three-gluon
    vertex of the gluon and the S pair (op 6). Leading colour: the gluon sits on the colour line
    from q_S to qbar_M.

    Every helicity configuration is MHV or anti-MHV; |A|^2 = s_ab^2 s(qbar_M, q_S) s(qbar_S, q_M)
    / prod s_{i,i+1} (a, b the two negative-helicity fermions for a positive gluon, the two
    positive ones for a negative gluon; with no gluon s_ab^2 / s(q_S, qbar_S)^2), and the
    amplitude vanishes when the gluon's polarisation is replaced by its momentum
    (tests/test_physics.py). Output o: the main line's chirality (o < 4: -1), the S line's, the
    gluon's helicity (bits from the most significant)."""
    n = 5 if gluon else 4
    p = Plan(n)
    p.name = "qqbarQQbar+g" if gluon else "qqbarQQbar"
    p.crossing = [(leg, -1.0 if leg in incoming else 1.0) for leg in range(n)]
    mass0 = p.scale(0.0)
    g3 = p.scale(complex(0.0, -SQRT_HALF))
    gq = p.scale(QUARK_GLUON)
    one = p.factor(1.0)
    pair = p.factor(FERMION_PAIR)
    rows = p.rows
    qbar_m, q_s, qbar_s, q_m = (0, 2, 3, 4) if gluon else (0, 1, 2, 3)
    J = {}
    if gluon:
        f = p.form([1])
        for h in (-1, 1):
            base = p.current(4)
            rows.append(("source", base, f, one, mass0, source_code(FAMILY_VECTOR, 0, h, 0)))
            J[h] = base

    def helicity(leg, chirality):
        return -chirality if leg in incoming else chirality

    src, code_hel = {}, {}
    for ch in (-1, 1):
        # a quark of chirality ch and its line's antiquark of chirality -ch
        for leg, orientation, chirality in ((q_m, 0, ch), (qbar_m, 1, -ch), (q_s, 0, ch), (qbar_s, 1, -ch)):
            base = p.current(2)
            h = helicity(leg, chirality)
            rows.append(("source", base, p.form([leg]), one, mass0, source_code(FAMILY_WEYL, orientation, h, chirality)))
            src[(leg, ch)] = base
            code_hel[(leg, ch)] = h
    components = p.factor_run([1.0, 1.0])
    for ch_m in (-1, 1):
        vertex_m, propagator_m = (2, 8) if ch_m == 1 else (3, 7)
        for ch_s in (-1, 1):
            vertex_s, propagator_s = (2, 8) if ch_s == 1 else (3, 7)
            for h in ((-1, 1) if gluon else (None,)):
                # the S pair's vector current, propagated
                k_s = p.current(4)
                _fermion_pair(rows, ch_s, src[(q_s, ch_s)], src[(qbar_s, ch_s)], k_s, pair, 1)
                rows.append(("prop", 9, k_s, p.form([q_s, qbar_s]), one))
                full = p.current(2)
                if gluon:
                    # the S line with the gluon absorbed, then the vector current of (g, S pair):
                    # its fermion pair and the three-gluon vertex, propagated
                    psi_s = p.current(2)
                    rows.append(("contrib", vertex_s, 1, gq, src[(q_s, ch_s)], J[h], psi_s, 0, 0, one))
                    rows.append(("prop", propagator_s, psi_s, p.form([1, q_s]), one))
                    k_gs = p.current(4)
                    _fermion_pair(rows, ch_s, psi_s, src[(qbar_s, ch_s)], k_gs, pair, 1)
                    rows.append(("contrib", 6, 0, g3, J[h], k_s, k_gs, p.form([1]), p.form([q_s, qbar_s]), one))
                    rows.append(("prop", 9, k_gs, p.form([1, q_s, qbar_s]), one))
                    # the main line: the S pair next to q_M, then the gluon; or both at once
                    psi_m = p.current(2)
                    rows.append(("contrib", vertex_m, 1, gq, src[(q_m, ch_m)], k_s, psi_m, 0, 0, one))
                    rows.append(("prop", propagator_m, psi_m, p.form([q_s, qbar_s, q_m]), one))
                    rows.append(("contrib", vertex_m, 1, gq, psi_m, J[h], full, 0, 0, one))
                    rows.append(("contrib", vertex_m, 0, gq, src[(q_m, ch_m)], k_gs, full, 0, 0, one))
                else:
                    rows.append(("contrib", vertex_m, 1, gq, src[(q_m, ch_m)], k_s, full, 0, 0, one))
                amp = p.amplitude()
                rows.append(("closure", src[(qbar_m, ch_m)], full, amp, one, components, 2))
                p.outputs.append([(amp, weight)])
                hel = [0] * n
                hel[qbar_m], hel[q_m] = code_hel[(qbar_m, ch_m)], code_hel[(q_m, ch_m)]
                hel[q_s], hel[qbar_s] = code_hel[(q_s, ch_s)], code_hel[(qbar_s, ch_s)]
                if gluon:
                    hel[1] = h
                p.helicities.append(tuple(hel))
    return p


# ---- flop count (per point) ----------------------------------------------------------------
# A real add, subtract or multiply counts 1, an fma 2, a divide or square root 1 (also counted
# apart); negation and multiplication by i count 0 (the operations of runner/recurrence_runner.cu
# :229-330 as written).
CONTRIB_VALUES = {2: 60, 3: 60, 4: 88, 5: 84, 6: 142}
CONTRIB_OUT = {2: 2, 3: 2, 4: 4, 5: 6, 6: 4}
SOURCE_FLOPS = {FAMILY_SCALAR: 6, FAMILY_WEYL: 20, FAMILY_DIRAC: 30, FAMILY_VECTOR: 22}


def kernel_flops():
    """flops per call of each generated symjit kernel, counted from its plane-kernel file (the
    rules above: copy, neg and constants free)"""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernels")
    cost = {"add": 1, "sub": 1, "mul": 1, "div": 1, "sqrt": 1, "fma": 2}
    out = []
    for name in sorted(f for f in os.listdir(here) if f.endswith(".pk")):
        n = 0
        for line in open(os.path.join(here, name)):
            w = line.split("#")[0].split()
            if len(w) >= 3 and w[1] == "=":
                n += cost.get(w[2], 0)
        out.append(n)
    return out


def plan_flops(plan):
    """(flops, divides) per point of one plan run: forms, sources, program and reduction."""
    flops, divs = 0, 0
    kflops = kernel_flops()
    binding = None
    for f in plan.forms:
        flops += 4 * (2 * len(f) - 1)
    for r in plan.rows:
        k = r[0]
        if k == "contrib":
            op, init = r[1], r[2]
            cd = CONTRIB_OUT[op]
            flops += 6 + CONTRIB_VALUES[op] + 6 * cd + (0 if init else 2 * cd)
        elif k == "prop":
            op = r[1]
            flops += 7 + (8 + 24 if op == 9 else 2 + 28 + 12)
            divs += 1
        elif k == "identity":
            flops += 6 * r[2]
        elif k == "closure":
            flops += 20 * r[6]
        elif k == "source":
            code = r[5]
            fam = code & 7
            comps = {0: 1, 1: 2, 2: 4, 3: 4}[fam]
            flops += SOURCE_FLOPS[fam] + 6 * comps
            divs += 3
        elif k == "binding":
            binding = r
        elif k == "symjit":
            # the kernel, then each output pair times the factor (6) and, accumulating, added (2)
            pairs = len(binding[3])
            flops += kflops[binding[1]] + 6 * pairs + (0 if r[1] else 2 * pairs)
    for out in plan.outputs:
        flops += 5 * len(out) + 1
    return flops, divs
