# SPDX-License-Identifier: 0BSD
"""Phase space: RAMBO momenta for 2 -> n-2 massless partons, the analysis cuts, the runner's
external-momentum layout.

`rambo_xorshift` restates convert/src/record.rs:17-52 (xorshift64 with shifts 13, 7, 17 from
state seed | 1, u = ((s >> 11) + 0.5) / 2^53; per outgoing parton c = 2u - 1, phi = 2 pi u,
E = -ln(u u'); boost and scale to sqrt(s) = 1000 GeV; beams (500, 0, 0, +-500)), point by point
in Python, for small samples. `rambo` is the same construction on a counter-based stream
(uniform k = splitmix64 finaliser of seed_state + k * golden gamma, seed_state = splitmix64(seed))
so that millions of points are drawn vectorised and any seed gives an independent stream
(record.rs's seed | 1 maps seeds 2 and 3 onto one stream).

Momenta are point-major [point][leg][E, px, py, pz], incoming legs first with positive energy;
`external_planes` applies the crossing signs and lays them out as the runner reads them,
external[(leg * 4 + k) * points + i]."""
import math

import numpy as np

ECM = 1000.0
GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def seed_state(seed):
    with np.errstate(over="ignore"):
        return _mix(np.uint64(seed) + GAMMA)


def uniforms(seed, count, offset=0):
    """count uniforms in (0, 1) of the stream `seed`, from position `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = _mix(seed_state(seed) + k * GAMMA)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)


def uniform_at(seed, ids):
    """the uniforms number ids + 1 of the stream `seed` (the device's hit-or-miss numbers)"""
    with np.errstate(over="ignore"):
        k = np.asarray(ids, dtype=np.uint64) + np.uint64(1)
        z = _mix(seed_state(seed) + k * GAMMA)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)


def log_prior(mom, crossing):
    """log(sum_{a<b} s_ab^4 / prod_l s_{l,l+1}), s_ab = |2 p_a.p_b| over the plan's legs in
    crossing order (cyclic); mom [points][legs][4] physical momenta"""
    legs = len(crossing)
    p = np.stack([mom[:, c[0]] for c in crossing], axis=1)

    def s(a, b):
        x, y = p[:, a], p[:, b]
        return np.abs(2 * (x[:, 0] * y[:, 0] - x[:, 1] * y[:, 1] - x[:, 2] * y[:, 2] - x[:, 3] * y[:, 3]))
    num = sum(s(a, b) ** 4 for a in range(legs) for b in range(a + 1, legs))
    den = np.prod([s(a, (a + 1) % legs) for a in range(legs)], axis=0)
    return np.log(num) - np.log(den)


def _rambo_from(u, n, ecm):
    """RAMBO from uniforms u[point, particle, 4] (c, phi, e1, e2), as record.rs builds them"""
    m = n - 2
    c = 2.0 * u[:, :, 0] - 1.0
    s = np.sqrt(1.0 - c * c)
    f = 2.0 * math.pi * u[:, :, 1]
    e = -np.log(u[:, :, 2] * u[:, :, 3])
    q = np.stack([e, e * s * np.cos(f), e * s * np.sin(f), e * c], axis=-1)  # [P, m, 4]
    total = q.sum(axis=1)
    mass = np.sqrt(total[:, 0] ** 2 - total[:, 1] ** 2 - total[:, 2] ** 2 - total[:, 3] ** 2)
    b = -total[:, 1:] / mass[:, None]
    x = ecm / mass
    g = total[:, 0] / mass
    a = 1.0 / (1.0 + g)
    bq = (b[:, None, :] * q[:, :, 1:]).sum(axis=-1)
    out = np.empty((u.shape[0], n, 4))
    out[:, 0] = [ecm / 2, 0.0, 0.0, ecm / 2]
    out[:, 1] = [ecm / 2, 0.0, 0.0, -ecm / 2]
    out[:, 2:, 0] = x[:, None] * (g[:, None] * q[:, :, 0] + bq)
    out[:, 2:, 1:] = x[:, None, None] * (q[:, :, 1:] + b[:, None, :] * (q[:, :, :1] + a[:, None, None] * bq[:, :, None]))
    del m
    return out


def rambo(n, points, seed, offset=0, ecm=ECM):
    """`points` RAMBO points of n legs from the counter-based stream `seed` at point `offset`."""
    m = n - 2
    u = uniforms(seed, points * m * 4, offset * m * 4).reshape(points, m, 4)
    return _rambo_from(u, n, ecm)


def record_state(seed):
    """the xorshift state a seed starts: record.rs takes seed | 1 (so seeds 2 and 3 would share a
    stream); here the seed goes through the splitmix64 finaliser first, then | 1"""
    z = (seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return (z ^ (z >> 31)) | 1


def rambo_at(n, seed, ids, ecm=ECM):
    """the RAMBO points number `ids` of the counter-based stream `seed` (as rambo() draws them)"""
    ids = np.asarray(ids, dtype=np.uint64)
    per = np.uint64(4 * (n - 2))
    pos = ids[:, None] * per + np.arange(int(per), dtype=np.uint64)[None, :]
    u = uniform_at(seed, pos.reshape(-1)).reshape(len(ids), n - 2, 4)
    return _rambo_from(u, n, ecm)


def min_pt_and_dr(mom):
    """per point: the smallest pT and the smallest Delta R among the outgoing partons"""
    out = mom[:, 2:]
    px, py, pz = out[..., 1], out[..., 2], out[..., 3]
    pt = np.hypot(px, py)
    p = np.sqrt(px * px + py * py + pz * pz)
    eta = 0.5 * np.log((p + pz) / (p - pz))
    phi = np.arctan2(py, px)
    dr = np.full(len(mom), np.inf)
    m = out.shape[1]
    for i in range(m):
        for j in range(i + 1, m):
            dphi = np.abs(phi[:, i] - phi[:, j])
            dphi = np.where(dphi > math.pi, 2 * math.pi - dphi, dphi)
            dr = np.minimum(dr, np.hypot(eta[:, i] - eta[:, j], dphi))
    return pt.min(axis=1), dr


def rambo_xorshift(n, points, seed, ecm=ECM, state=None):
    """convert/src/record.rs:17-52 restated: the reference's own generator, point by point, from
    state seed | 1 (record.rs) or the given state."""
    state = (seed | 1) & 0xFFFFFFFFFFFFFFFF if state is None else state
    mask = 0xFFFFFFFFFFFFFFFF

    def uniform():
        nonlocal state
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        return ((state >> 11) + 0.5) / float(1 << 53)

    out = []
    for _ in range(points):
        q = []
        for _ in range(n - 2):
            c = 2.0 * uniform() - 1.0
            s = math.sqrt(1.0 - c * c)
            f = 2.0 * math.pi * uniform()
            e = -math.log(uniform() * uniform())
            q.append([e, e * s * math.cos(f), e * s * math.sin(f), e * c])
        total = [sum(p[k] for p in q) for k in range(4)]
        mass = math.sqrt(total[0] * total[0] - total[1] * total[1] - total[2] * total[2] - total[3] * total[3])
        b = [-total[1] / mass, -total[2] / mass, -total[3] / mass]
        x, g = ecm / mass, total[0] / mass
        a = 1.0 / (1.0 + g)
        point = [[ecm / 2, 0.0, 0.0, ecm / 2], [ecm / 2, 0.0, 0.0, -ecm / 2]]
        for p in q:
            bq = b[0] * p[1] + b[1] * p[2] + b[2] * p[3]
            point.append([x * (g * p[0] + bq)] + [x * (p[k + 1] + b[k] * (p[0] + a * bq)) for k in range(3)])
        out.append(point)
    return np.array(out)


def rambo_weight(n, ecm=ECM):
    """the constant phase-space volume of n-2 massless partons (RAMBO's weight), GeV^(2(n-2)-4)"""
    m = n - 2
    s = ecm * ecm
    return (2 * math.pi) ** (4 - 3 * m) * (math.pi / 2) ** (m - 1) * s ** (m - 2) / (math.factorial(m - 1) * math.factorial(m - 2))


def passes_cuts(mom, pt_min=20.0, eta_max=2.5, dr_min=0.4):
    """pT > 20 GeV, |eta| < 2.5 for every outgoing parton and Delta R > 0.4 between each pair"""
    out = mom[:, 2:]
    px, py, pz = out[..., 1], out[..., 2], out[..., 3]
    pt = np.hypot(px, py)
    p = np.sqrt(px * px + py * py + pz * pz)
    with np.errstate(divide="ignore", invalid="ignore"):
        eta = 0.5 * np.log((p + pz) / (p - pz))
    phi = np.arctan2(py, px)
    ok = np.all(pt > pt_min, axis=1) & np.all(np.abs(eta) < eta_max, axis=1)
    m = out.shape[1]
    for i in range(m):
        for j in range(i + 1, m):
            dphi = np.abs(phi[:, i] - phi[:, j])
            dphi = np.where(dphi > math.pi, 2 * math.pi - dphi, dphi)
            deta = eta[:, i] - eta[:, j]
            ok &= deta * deta + dphi * dphi > dr_min * dr_min
    return ok


def generate(n, points, seed, batch=1 << 20, ecm=ECM):
    """`points` RAMBO points passing the cuts, and the number drawn to get them. (The stream is
    counter based, so the batch size changes only how many are drawn at a time, never which.)"""
    batch = min(batch, max(4096, 2 * points))
    kept, drawn, offset = [], 0, 0
    have = 0
    while have < points:
        m = rambo(n, batch, seed, offset, ecm)
        offset += batch
        ok = passes_cuts(m)
        good = m[ok]
        need = points - have
        if len(good) >= need:
            # count the draws up to the last point kept
            idx = np.nonzero(ok)[0][need - 1]
            drawn += idx + 1
            kept.append(good[:need])
            have += need
        else:
            drawn += batch
            kept.append(good)
            have += len(good)
    return np.concatenate(kept), drawn


def external_planes(mom, crossing, legs=None):
    """the runner's external momenta: sign * p[point][leg][k] at [(l * 4 + k) * points + i]"""
    P, n, _ = mom.shape
    legs = n if legs is None else legs
    ext = np.empty((legs, 4, P))
    for l in range(legs):
        src, sign = crossing[l] if crossing else (l, 1.0)
        ext[l] = sign * mom[:, src, :].T
    return np.ascontiguousarray(ext.reshape(legs * 4, P))
