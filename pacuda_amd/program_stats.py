# SPDX-License-Identifier: 0BSD
"""What a fused program (pacuda_plan_program's instruction stream) moves through its wavefront's
scratch slot per point: current components loaded and stored (16 B each in f64, 8 B in the
mixed kernels; on-chip planes and known-zero components take none), momentum forms read from the
slot (32 B; the kernel's register forms take none), amplitude loads and stores (16 B), and the
source rows made again. For inspection and reports (bench.py); the kernels do the same
accesses."""
LDS_PLANE, DUAL_FILL, DUAL_THROUGH = 1 << 24, 1 << 31, 1 << 30
FORM_SLOT_SHIFT = 24
K_CLEAR, K_CONTRIB, K_PROP, K_IDENTITY, K_CLOSURE, K_SOURCE = 1, 2, 3, 4, 5, 6
F_INIT, F_FIRST, F_LAST, F_HOLD, F_FROMACC, F_ZERO = 1, 2, 4, 8, 16, 32
# (parent 0, parent 1, destination) components per contribution op
SHAPE = {2: (2, 4, 2), 3: (2, 4, 2), 4: (6, 4, 4), 5: (4, 4, 6), 6: (4, 4, 4)}
SOURCE_COMPONENTS = {0: 1, 1: 2}  # by family (w[5] & 7); vector families: 4


def slot_traffic(prog, sources, register_forms=0, current_bytes=16):
    """per point: {"current_loads", "current_stores", "form_loads", "amplitude_loads",
    "amplitude_stores", "remade_sources", "bytes"} of the rows after the leading sources (the
    sources phase stores its currents once per point, counted in "source_stores")."""
    st = dict(current_loads=0, current_stores=0, form_loads=0, amplitude_loads=0, amplitude_stores=0,
              remade_sources=0, source_stores=0)

    def count(field, n, zero, load=True):
        if field >= LDS_PLANE and not (field & (DUAL_FILL if load else DUAL_THROUGH)):
            return 0
        return sum(1 for j in range(n) if not (zero >> j) & 1)

    def form(field):
        slot = field >> FORM_SLOT_SHIFT
        return 0 if 0 < slot <= register_forms else 1

    for n, w in enumerate(prog):
        w = [int(x) for x in w]
        kind, op, flags = w[0] & 0xff, (w[0] >> 8) & 0xff, w[0] >> 16
        if kind == K_SOURCE:
            comps = SOURCE_COMPONENTS.get(w[5] & 7, 4)
            if n < sources:
                st["source_stores"] += count(w[1], comps, w[6], False)
            else:
                st["remade_sources"] += 1
                st["form_loads"] += form(w[2])
                st["current_stores"] += count(w[1], comps, w[6], False)
        elif kind == K_CONTRIB:
            c0, c1, cd = SHAPE[op]
            st["current_loads"] += count(w[2], c0, (w[0] >> 22) & 0x3f) + count(w[3], c1, (w[0] >> 28) & 0xf)
            if (flags & F_FIRST) and not (flags & F_INIT):
                st["current_loads"] += count(w[4], cd, (w[1] >> 16) & 0x3f)
            if (flags & F_LAST) and not (flags & F_HOLD):
                st["current_stores"] += count(w[4], cd, (w[7] >> 16) & 0x3f, False)
            if op == 6:
                st["form_loads"] += form(w[5]) + form(w[6])
        elif kind == K_PROP:
            m = 4 if op == 9 else 2
            if not flags & F_FROMACC:
                st["current_loads"] += count(w[1], m, w[4] & 0xff)
            if not flags & F_HOLD:
                st["current_stores"] += count(w[1], m, (w[4] >> 8) & 0xff, False)
            st["form_loads"] += form(w[2])
        elif kind == K_IDENTITY:
            if not flags & F_FROMACC:
                st["current_loads"] += count(w[1], w[2], w[4] & 0xffff)
            if not flags & F_HOLD:
                st["current_stores"] += count(w[1], w[2], w[4] >> 16, False)
        elif kind == K_CLOSURE:
            st["current_loads"] += count(w[1], w[6], w[7] & 0xff) + count(w[2], w[6], (w[7] >> 8) & 0xff)
            st["amplitude_stores"] += 1
            st["amplitude_loads"] += 0 if flags & F_ZERO else 1
        elif kind == K_CLEAR:
            st["current_stores"] += count(w[1], w[2], 0, False)
    st["bytes"] = ((st["current_loads"] + st["current_stores"] + st["source_stores"]) * current_bytes
                   + st["form_loads"] * 32 + (st["amplitude_loads"] + st["amplitude_stores"]) * 16)
    return st
