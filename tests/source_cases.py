# SPDX-License-Identifier: 0BSD
"""Source (op 12) test cases and the wavefunction legend.

legend_cases: the case generator of tests/check_sources.py:75-90 restated (Python's random with
  seed 3, 300 cases per family: beam-axis, zero and random three-momenta, both energy signs).
patch_cases: the generator of rusticol/patches/0009-...patch:64-117 restated (xorshift64 from
  0x9e3779b97f4a7c15, 40 momenta per mass in both energy signs, the beam axis first).
legend(): the closed forms of convert/CONVENTIONS.md:52-61 (HELAS-style spinors and polarisation
  vectors) in Python complex arithmetic, for a 1e-13 check. It follows the branch structure of the
  reference's own Python legend (tests/check_sources.py:10-72), so it pins the oracle to that
  legend, not to an independent derivation; the physical identities in tests/test_oracle.py
  (transversality, normalisation, spinor completeness) are the independent pins.
CPU only: tests/golden/make_source_cases.py writes these cases and values into a data file, which
the GPU tests load."""
import math
import random

SQ2 = math.sqrt(2.0)


def code(family, orientation, helicity, chirality):
    return family | orientation << 3 | (helicity + 2) << 5 | (chirality + 1) << 8


FAMILIES = {"scalar": (0, 0), "weyl particle": (1, 0), "weyl antiparticle": (1, 1), "dirac particle": (2, 0),
            "dirac antiparticle": (2, 1), "vector (massless)": (3, 0), "vector (massive)": (3, 0)}


def legend_cases(seed=3, per_family=300):
    """[(code, mass, [E, px, py, pz], name, helicity, chirality)]"""
    rnd = random.Random(seed)
    cases = []
    for name, (fam, ori) in FAMILIES.items():
        for _ in range(per_family):
            if name == "vector (massless)" or name.startswith(("weyl", "scalar")):
                m = 0.0
            elif name.startswith("dirac"):
                m = rnd.choice([0.0, 4.7, 91.188])
            else:
                m = rnd.choice([80.4, 91.188])
            kind = rnd.random()
            if kind < 0.15:
                p3 = [0.0, 0.0, rnd.choice([-1, 1]) * rnd.uniform(1, 500)]
            elif kind < 0.2:
                p3 = [0.0, 0.0, 0.0]
            else:
                p3 = [rnd.uniform(-500, 500) for _ in range(3)]
            E = math.sqrt(sum(x * x for x in p3) + m * m)
            if E == 0:
                continue
            sgn = rnd.choice([1, -1])
            k = [sgn * E] + [sgn * x for x in p3]
            hs = [0] if fam == 0 else ([-1, 0, 1] if name == "vector (massive)" else [-1, 1])
            h = rnd.choice(hs)
            ch = rnd.choice([-1, 1]) if fam == 1 else 0
            cases.append((code(fam, ori, h, ch), m, k, name, h, ch))
    return cases


def patch_cases():
    state = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1

    def uniform():
        nonlocal state
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        return (state >> 11) / float(1 << 53)

    def momenta_for(mass):
        out = []
        for k in range(40):
            if k < 4:
                px, py, pz = 0.0, 0.0, (250.0 if k % 2 == 0 else -250.0)
            else:
                r = 1.0 + 499.0 * uniform()
                c = 2.0 * uniform() - 1.0
                s = math.sqrt(1.0 - c * c)
                f = 2.0 * math.pi * uniform()
                px, py, pz = r * s * math.cos(f), r * s * math.sin(f), r * c
            e = math.sqrt(px * px + py * py + pz * pz + mass * mass)
            out.append([e, px, py, pz])
            out.append([-e, -px, -py, -pz])
        return out

    specs = [(0, 0, 0, 0, 0.0)]
    for orientation in range(2):
        for helicity in (-1, 1):
            for chirality in (-1, 1):
                specs.append((1, orientation, helicity, chirality, 0.0))
            for mass in (0.0, 1.0e-9, 4.7, 173.0):
                specs.append((2, orientation, helicity, 0, mass))
    for helicity in (-1, 0, 1):
        for mass in (0.0, 80.379, 91.1876):
            if helicity == 0 and mass == 0.0:
                continue
            specs.append((3, 0, helicity, 0, mass))
    cases = []
    for family, orientation, helicity, chirality, mass in specs:
        for p in momenta_for(mass):
            side = " antiparticle" if orientation else " particle"
            name = {0: "scalar", 1: "weyl" + side, 2: "dirac" + side,
                    3: "vector (massless)" if mass == 0.0 else "vector (massive)"}[family]
            cases.append((code(family, orientation, helicity, chirality), mass, p, name, helicity, chirality))
    return cases


def _sign(x):
    return math.copysign(1.0, x)


def _ab_weyl(k, h, anti):
    E, px, py, pz = k
    if E > 0:
        s = 0.0 if (px == 0 and py == 0 and pz < 0) else math.sqrt(max(E + pz, 0))
        if anti:
            s = -s if s != 0 else 0.0
        if s == 0:
            b = -h * math.sqrt(2 * E)
        else:
            b = (-h * px / s + 1j * py / s) if anti else (h * px / s - 1j * py / s)
    else:
        s = 0.0 if (px == 0 and py == 0 and pz > 0) else math.sqrt(max(-(E + pz), 0))
        if not anti:
            s = -s if s != 0 else 0.0
        if s == 0:
            b = -h * math.sqrt(2 * abs(E))
        else:
            b = (-h * px / s - 1j * py / s) if anti else (h * px / s + 1j * py / s)
    return complex(s), complex(b)


def legend(name, k, h, ch, m):
    """convert/CONVENTIONS.md:52-61, the wavefunction w of a source (without its factor)"""
    E, px, py, pz = k
    pt = math.hypot(px, py)
    ak = math.sqrt(px * px + py * py + pz * pz)
    if name == "scalar":
        return [1]
    if name.startswith("weyl"):
        anti = "anti" in name
        a, b = _ab_weyl(k, h, anti)
        if not anti:
            if E > 0:
                return [a, b] if (h, ch) == (1, 1) else [b, a] if (h, ch) == (-1, -1) else [0, 0]
            return [a, b] if (h, ch) == (-1, 1) else [b, a] if (h, ch) == (1, -1) else [0, 0]
        if E > 0:
            return [b, a] if (h, ch) == (1, 1) else [a, b] if (h, ch) == (-1, -1) else [0, 0]
        return [b, a] if (h, ch) == (-1, 1) else [a, b] if (h, ch) == (1, -1) else [0, 0]
    if name.startswith("dirac"):
        anti = "anti" in name
        if abs(m) < 1e-8:
            a, b = _ab_weyl(k, h, anti)
            if not anti:
                cond = (h == 1) if E > 0 else (h == -1)
                return [a, b, 0, 0] if cond else [0, 0, b, a]
            cond = (h == -1) if E > 0 else (h == 1)
            return [0, 0, a, b] if cond else [b, a, 0, 0]
        n = (1 if E > 0 else -1) if not anti else (-1 if E > 0 else 1)
        nh = n * h
        om = [math.sqrt(abs(E) + ak), m / math.sqrt(abs(E) + ak)]
        f1 = (1 + n + (1 - n) * nh) / 2
        f2 = (1 + n - (1 - n) * nh) / 2
        ip, im_ = int((1 + nh) // 2), int((1 - nh) // 2)
        g = [f1 * om[ip], f2 * om[im_]]
        sgn_e = 1 if E > 0 else -1
        t = [sgn_e * px, sgn_e * py, sgn_e * pz]
        r = max(ak + t[2], 0)
        c1 = 1.0 if ak == 0 else math.sqrt(r / (2 * ak))
        if r == 0 or ak == 0:
            c2 = -nh
        else:
            c2 = ((nh * t[0] + 1j * t[1]) if anti else (nh * t[0] - 1j * t[1])) / math.sqrt(2 * ak * r)
        chi = [c1, c2]
        if not anti:
            return [chi[im_] * g[1], chi[ip] * g[1], chi[im_] * g[0], chi[ip] * g[0]]
        return [chi[im_] * g[0], chi[ip] * g[0], chi[im_] * g[1], chi[ip] * g[1]]
    if name == "vector (massless)" or (name == "vector" and m == 0.0):
        if E > 0:
            if pt == 0:
                return [0, -h / SQ2, 1j * _sign(pz) / SQ2, 0]
            return [0, (-h * px * pz / (E * pt) - 1j * py / pt) / SQ2, (-h * py * pz / (E * pt) + 1j * px / pt) / SQ2,
                    h * pt / (E * SQ2)]
        g_, P = -h, -E
        if pt == 0:
            return [0, -g_ / SQ2, -1j * _sign(pz) / SQ2, 0]
        return [0, (-g_ * px * pz / (P * pt) + 1j * py / pt) / SQ2, (-g_ * py * pz / (P * pt) - 1j * px / pt) / SQ2,
                g_ * pt / (P * SQ2)]
    # massive vector
    if E < 0:
        E, px, py, pz, h = -E, -px, -py, -pz, -h
    P = min(E, math.sqrt(px * px + py * py + pz * pz))
    q = min(P, math.hypot(px, py))
    l = 1 - abs(h)
    n = abs(h)
    if P == 0:
        return [0, -h / SQ2, 1j * n / SQ2, l]
    e = E / (m * P)
    if q == 0:
        mid = [-h / SQ2, 1j * n * _sign(pz) / SQ2]
    else:
        z = pz * h / (P * q * SQ2)
        mid = [l * px * e - px * z - 1j * n * py / (q * SQ2), l * py * e - py * z + 1j * n * px / (q * SQ2)]
    return [l * P / m, mid[0], mid[1], l * pz * e + h * q / (P * SQ2)]
