# SPDX-License-Identifier: 0BSD
"""A plan whose symjit rows (op 13, bindings op 14) exercise every kind of binding input code and
both output modes, and its expected currents computed in numpy.

Rows (after the sources: a Weyl quark W on leg 0, a gluon V on leg 1, a Weyl antiquark B on
leg 2):
  A  probe kernel (the identity on 8 values), init, factor x1: inputs W[0] re, W[0] im, V[2] re,
     V[3] im, form 0 component 1, form 1 component 3, scale s (real part), the factor's imaginary
     part -> D0 components 0..3, each pair times x1
  B  probe, accumulate, factor x2: inputs D0[0] re, D0[0] im, D0[3] re, D0[3] im, the factor's
     real part, W[1] re, V[0] im, form 0 component 0 -> D0 components 1 and 2 (+=), then D0[3]
     and D0[0] from the last two pairs (+=)
  C  probe, direct, factor exactly 1: inputs W[1] im, W[1] re, V[1] re, V[1] im -> D1
     components 1, 0 stored as they are (no add, although the row does not initialise)
  D  probe, direct, factor x1 (not 1): the same inputs -> D2 components 0, 1 scaled, init
  E  fermion_pair_sigma, init, factor x3: W x B -> J (4 components)
  F  fermion_pair_sigmabar, accumulate, factor x1: W x B -> J (+=)
The expected values follow runner/recurrence_runner.cu:631-684 operation by operation in f64
(numpy's float64 arithmetic rounds each operation once, as the kernels do)."""
import numpy as np

from pacuda_amd import plans
from pacuda_amd.plans import binding_code as code, SB_PARENT, SB_MOMENTUM, SB_DESTINATION, SB_SCALE, SB_FACTOR

X1 = 0.75 - 0.5j
X2 = -1.25 + 2.0j
X3 = 0.5 + 0.25j
SCALE = 1.5 + 7.0j


def probe_plan():
    p = plans.Plan(3)
    p.name = "symjit-probe"
    p.crossing = [(0, 1.0), (1, 1.0), (2, 1.0)]
    mass0 = p.scale(0.0)
    s = p.scale(SCALE)
    one = p.factor(1.0)
    x1, x2, x3 = p.factor(X1), p.factor(X2), p.factor(X3)
    f0, f1, f2 = p.form([0]), p.form([1]), p.form([2])
    W = p.current(2)
    V = p.current(4)
    B = p.current(2)
    D0 = p.current(4)
    D1 = p.current(2)
    D2 = p.current(2)
    J = p.current(4)
    p.rows += [
        ("source", W, f0, one, mass0, plans.source_code(plans.FAMILY_WEYL, 0, 1, 1)),
        ("source", V, f1, one, mass0, plans.source_code(plans.FAMILY_VECTOR, 0, -1, 0)),
        ("source", B, f2, one, mass0, plans.source_code(plans.FAMILY_WEYL, 1, -1, -1)),
    ]
    probe = plans.KERNEL_BINDING_PROBE
    # A
    p.rows.append(("binding", probe, [code(SB_PARENT, 0, 0, 0), code(SB_PARENT, 0, 1, 0), code(SB_PARENT, 1, 0, 2),
                                      code(SB_PARENT, 1, 1, 3), code(SB_MOMENTUM, 0, 0, 1), code(SB_MOMENTUM, 1, 0, 3),
                                      code(SB_SCALE, 0, 0, s), code(SB_FACTOR, 0, 1, 0)], [0, 1, 2, 3], 0))
    p.rows.append(("symjit", 1, W, V, D0, f0, f1, x1))
    # B
    p.rows.append(("binding", probe, [code(SB_DESTINATION, 0, 0, 0), code(SB_DESTINATION, 0, 1, 0),
                                      code(SB_DESTINATION, 0, 0, 3), code(SB_DESTINATION, 0, 1, 3),
                                      code(SB_FACTOR, 0, 0, 0), code(SB_PARENT, 0, 0, 1), code(SB_PARENT, 1, 1, 0),
                                      code(SB_MOMENTUM, 0, 0, 0)], [1, 2, 3, 0], 0))
    p.rows.append(("symjit", 0, W, V, D0, f0, f1, x2))
    # C and D
    cd_codes = [code(SB_PARENT, 0, 1, 1), code(SB_PARENT, 0, 0, 1), code(SB_PARENT, 1, 0, 1), code(SB_PARENT, 1, 1, 1)]
    p.rows.append(("binding", probe, cd_codes, [1, 0], 1))
    p.rows.append(("symjit", 0, W, V, D1, f0, f1, one))
    p.rows.append(("binding", probe, cd_codes, [0, 1], 1))
    p.rows.append(("symjit", 1, W, V, D2, f0, f1, x1))
    # E and F
    p.rows.append(("binding", plans.KERNEL_FERMION_PAIR_SIGMA, plans.weyl_pair_codes(), [0, 1, 2, 3], 0))
    p.rows.append(("symjit", 1, W, B, J, f0, f2, x3))
    p.rows.append(("binding", plans.KERNEL_FERMION_PAIR_SIGMABAR, plans.weyl_pair_codes(), [0, 1, 2, 3], 0))
    p.rows.append(("symjit", 0, W, B, J, f0, f2, x1))
    # closures, one amplitude each (so the fused path, which returns amplitudes only, sees every
    # current the symjit rows wrote): D0 . V and J . V with the metric, D1 . W, D2 . B
    metric = p.factor_run([1.0, -1.0, -1.0, -1.0])
    pair = p.factor_run([1.0, 1.0])
    for l, r, comps, fac in ((D0, V, 4, metric), (D1, W, 2, pair), (D2, B, 2, pair), (J, V, 4, metric)):
        amp = p.amplitude()
        p.rows.append(("closure", l, r, amp, one, fac, comps))
        p.outputs.append([(amp, 1.0)])
    p.layout = dict(W=W, V=V, B=B, D0=D0, D1=D1, D2=D2, J=J)
    return p


def _scaled(out, f):
    """(re fr - im fi, re fi + im fr) of each (re, im) pair, each operation rounded"""
    re, im = out[0::2], out[1::2]
    return re * f.real - im * f.imag, re * f.imag + im * f.real


def expected(sources_of, forms):
    """the f64 currents D0, D1, D2, J (complex [components][points]) from the source currents
    (sources_of: name -> complex [components][points]) and the forms ([form][4][points])"""
    W, V, B = sources_of["W"], sources_of["V"], sources_of["B"]
    n = W.shape[1]
    # A: init, x1
    a_in = np.array([W[0].real, W[0].imag, V[2].real, V[3].imag, forms[0][1], forms[1][3],
                     np.full(n, SCALE.real), np.full(n, X1.imag)])
    re, im = _scaled(a_in, X1)
    D0r, D0i = re.copy(), im.copy()
    # B: accumulate, x2 (the inputs read before any output is written)
    b_in = np.array([D0r[0], D0i[0], D0r[3], D0i[3], np.full(n, X2.real), W[1].real, V[0].imag, forms[0][0]])
    re, im = _scaled(b_in, X2)
    for j, comp in enumerate([1, 2, 3, 0]):
        D0r[comp] = D0r[comp] + re[j]
        D0i[comp] = D0i[comp] + im[j]
    # C: direct with factor 1: stored as they are; D: direct, x1, init
    c_in = np.array([W[1].imag, W[1].real, V[1].real, V[1].imag])
    D1r, D1i = np.zeros((2, n)), np.zeros((2, n))
    D1r[1], D1i[1] = c_in[0], c_in[1]
    D1r[0], D1i[0] = c_in[2], c_in[3]
    re, im = _scaled(c_in, X1)
    D2r, D2i = np.array([re[0], re[1]]), np.array([im[0], im[1]])
    # E, F: l^T sigma^mu r and l^T sigmabar^mu r with explicit complex products
    l = [(W[0].real, W[0].imag), (W[1].real, W[1].imag)]
    r = [(B[0].real, B[0].imag), (B[1].real, B[1].imag)]

    def cmul(a, b):
        return a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0]

    p00, p11, p01, p10 = cmul(l[0], r[0]), cmul(l[1], r[1]), cmul(l[0], r[1]), cmul(l[1], r[0])
    add = lambda a, b: (a[0] + b[0], a[1] + b[1])  # noqa: E731
    sub = lambda a, b: (a[0] - b[0], a[1] - b[1])  # noqa: E731
    d = sub(p10, p01)
    sigma = [add(p00, p11), add(p01, p10), (-d[1], d[0]), sub(p00, p11)]
    sm = add(p01, p10)
    e = sub(p01, p10)
    sigmabar = [add(p00, p11), (-sm[0], -sm[1]), (-e[1], e[0]), sub(p11, p00)]
    out = np.array([x for c in sigma for x in c])
    Jr, Ji = _scaled(out, X3)
    out = np.array([x for c in sigmabar for x in c])
    re, im = _scaled(out, X1)
    Jr, Ji = Jr + re, Ji + im
    return {"D0": D0r + 1j * D0i, "D1": D1r + 1j * D1i, "D2": D2r + 1j * D2i, "J": Jr + 1j * Ji}


def run_currents(v1, plan, ext, points, stride=None, kernel=None):
    """one v1 run of `plan` (forms on the device) downloading every current plane:
    complex [current_planes][points]"""
    import ctypes as C
    from pacuda_amd import abi
    from pacuda_amd.abi import ptr, u32p, u64p
    a = plan.arrays()
    stride = stride or points
    cre = np.zeros(plan.current_planes * stride)
    cim = np.zeros(plan.current_planes * stride)
    n_amp = max(plan.amplitude_planes, 1)
    are, aim = np.zeros(n_amp * stride), np.zeros(n_amp * stride)
    runs = np.array([0, plan.current_planes], dtype=np.uint32)
    ext = np.ascontiguousarray(ext[:, :points])
    c = abi.Call()
    c.abi_version, c.struct_size = abi.ABI, C.sizeof(abi.Call)
    c.words, c.word_count = ptr(a["words"], u64p), len(a["words"])
    for k in range(8):
        c.width[k] = int(a["width"][k])
    c.current_re, c.current_im, c.amplitude_re, c.amplitude_im = ptr(cre), ptr(cim), ptr(are), ptr(aim)
    c.scale_re, c.scale_im = ptr(a["scale_re"]), ptr(a["scale_im"])
    c.factor_re, c.factor_im = ptr(a["factor_re"]), ptr(a["factor_im"])
    c.download_current, c.download_current_runs = ptr(runs, u32p), 1
    c.current_planes, c.amplitude_planes = plan.current_planes, plan.amplitude_planes
    c.momentum_planes = 4 * len(plan.forms)
    c.scale_count, c.factor_count = len(a["scale_re"]), len(a["factor_re"])
    c.stride, c.points = stride, points
    c.flags = abi.FORMS_ON_DEVICE
    c.external_momenta, c.external_count = ptr(ext), plan.legs
    c.term_count = len(a["term_slots"])
    c.form_terms, c.term_slots = ptr(a["form_terms"], u32p), ptr(a["term_slots"], u32p)
    c.term_coefficients = ptr(a["term_coefficients"])
    st = v1.run(C.byref(c))
    if st != 0:
        raise RuntimeError(f"run: status {st}")
    return (cre + 1j * cim).reshape(plan.current_planes, stride)[:, :points]
