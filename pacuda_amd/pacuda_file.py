# SPDX-License-Identifier: 0BSD
"""The .pacuda v2 recording: writer and reader of the grammar of convert/src/format.rs:6-29
(little endian, no padding, counted arrays, u32-length strings; the reader rejects trailing
bytes), so generated plans can be stored as recordings and replayed by pacuda-check or read
by umami/umami_pacuda.c."""
import struct

import numpy as np

MAGIC = b"PACUDA\0\x01"
VERSION = 2


class Out:
    def __init__(self):
        self.b = bytearray()

    def raw(self, v):
        self.b += v

    def u32(self, v):
        self.b += struct.pack("<I", int(v))

    def i32(self, v):
        self.b += struct.pack("<i", int(v))

    def i64(self, v):
        self.b += struct.pack("<q", int(v))

    def f64(self, v):
        self.b += struct.pack("<d", float(v))

    def str(self, s):
        e = s.encode()
        self.u32(len(e))
        self.b += e

    def arr(self, a, dtype):
        self.b += np.ascontiguousarray(a, dtype=dtype).tobytes()

    def runs(self, runs):
        self.u32(len(runs) // 2)
        self.arr(runs, "<u4")


def colour_record(plan, colour_id, tile, ref_momenta, ref_values):
    """one colour of a process: the plan's template, reduction and reference points"""
    a = plan.arrays()
    return dict(id=colour_id, tile=tile, words=a["words"], width=a["width"], current_planes=plan.current_planes,
                amplitude_planes=plan.amplitude_planes, momentum_planes=4 * len(plan.forms),
                scale_re=a["scale_re"], scale_im=a["scale_im"], factor_re=a["factor_re"], factor_im=a["factor_im"],
                zero_amplitude=[0, plan.amplitude_planes], download_amplitude=[0, plan.amplitude_planes],
                external_count=plan.legs, form_terms=a["form_terms"], term_slots=a["term_slots"],
                term_coefficients=a["term_coefficients"], offsets=a["offsets"], planes=a["planes"], weights=a["weights"],
                points=np.asarray(ref_momenta).size // (4 * plan.legs), momenta=np.asarray(ref_momenta).reshape(-1),
                values=np.asarray(ref_values).reshape(-1))


def write(processes, kernel=1, batch=0, sha256=()):
    """processes: [dict(key, legs, helicities, alpha_s_name, alpha_s, qcd_power, parameters,
    numerator, denominator, crossing [(leg, sign)] or None, colours [colour_record])]"""
    o = Out()
    o.raw(MAGIC)
    o.u32(VERSION)
    o.u32(kernel)
    o.u32(batch)
    o.u32(len(sha256))
    for s in sha256:
        assert len(s) == 64
        o.raw(s.encode())
    o.u32(len(processes))
    for p in processes:
        o.str(p["key"])
        o.u32(p["legs"])
        o.u32(p["helicities"])
        o.str(p.get("alpha_s_name", ""))
        o.f64(p.get("alpha_s", 0.0))
        o.i32(p.get("qcd_power", 0))
        params = p.get("parameters", [])
        o.u32(len(params))
        for name, re, im in params:
            o.str(name)
            o.f64(re)
            o.f64(im)
        o.i64(p.get("numerator", 1))
        o.i64(p.get("denominator", 1))
        crossing = p.get("crossing")
        if crossing is None:
            o.u32(0)
        else:
            o.u32(len(crossing))
            for leg, sign in crossing:
                o.u32(leg)
                o.f64(sign)
        o.u32(len(p["colours"]))
        for c in p["colours"]:
            o.str(c["id"])
            o.u32(c["tile"])
            o.u32(len(c["words"]))
            o.arr(c["words"], "<u8")
            o.arr(c["width"], "u1")
            o.u32(c["current_planes"])
            o.u32(c["amplitude_planes"])
            o.u32(c["momentum_planes"])
            o.u32(len(c["scale_re"]))
            o.arr(c["scale_re"], "<f8")
            o.arr(c["scale_im"], "<f8")
            o.u32(len(c["factor_re"]))
            o.arr(c["factor_re"], "<f8")
            o.arr(c["factor_im"], "<f8")
            o.runs(c["zero_amplitude"])
            o.runs(c["download_amplitude"])
            o.u32(c["external_count"])
            o.runs(c["form_terms"])
            o.u32(len(c["term_slots"]))
            o.arr(c["term_slots"], "<u4")
            o.arr(c["term_coefficients"], "<f8")
            o.u32(len(c["offsets"]) - 1)
            o.arr(c["offsets"], "<u4")
            o.arr(c["planes"], "<u4")
            o.arr(c["weights"], "<f8")
            o.u32(c["points"])
            o.arr(c["momenta"], "<f8")
            o.arr(c["values"], "<f8")
    return bytes(o.b)


class _In:
    def __init__(self, b):
        self.b, self.at = b, 0

    def take(self, n):
        if n > len(self.b) - self.at:
            raise ValueError("truncated .pacuda file")
        v = self.b[self.at:self.at + n]
        self.at += n
        return v

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def i64(self):
        return struct.unpack("<q", self.take(8))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def str(self):
        return self.take(self.u32()).decode()

    def arr(self, n, dtype):
        return np.frombuffer(self.take(n * np.dtype(dtype).itemsize), dtype=dtype).copy()

    def runs(self):
        return self.arr(2 * self.u32(), "<u4")


def read(b):
    r = _In(b)
    if r.take(8) != MAGIC:
        raise ValueError("not a .pacuda file")
    if r.u32() != VERSION:
        raise ValueError("not a version 2 .pacuda file")
    f = dict(kernel=r.u32(), batch=r.u32())
    f["sha256"] = [r.take(64).decode() for _ in range(r.u32())]
    procs = []
    for _ in range(r.u32()):
        p = dict(key=r.str(), legs=r.u32(), helicities=r.u32(), alpha_s_name=r.str(), alpha_s=r.f64(), qcd_power=r.i32())
        p["parameters"] = [(r.str(), r.f64(), r.f64()) for _ in range(r.u32())]
        p["numerator"], p["denominator"] = r.i64(), r.i64()
        n = r.u32()
        p["crossing"] = [(r.u32(), r.f64()) for _ in range(n)] if n else None
        cols = []
        for _ in range(r.u32()):
            c = dict(id=r.str(), tile=r.u32())
            c["words"] = r.arr(r.u32(), "<u8")
            c["width"] = r.arr(8, "u1")
            c["current_planes"], c["amplitude_planes"], c["momentum_planes"] = r.u32(), r.u32(), r.u32()
            ns = r.u32()
            c["scale_re"], c["scale_im"] = r.arr(ns, "<f8"), r.arr(ns, "<f8")
            nf = r.u32()
            c["factor_re"], c["factor_im"] = r.arr(nf, "<f8"), r.arr(nf, "<f8")
            c["zero_amplitude"], c["download_amplitude"] = r.runs(), r.runs()
            c["external_count"] = r.u32()
            c["form_terms"] = r.runs()
            nt = r.u32()
            c["term_slots"], c["term_coefficients"] = r.arr(nt, "<u4"), r.arr(nt, "<f8")
            outputs = r.u32()
            c["offsets"] = r.arr(outputs + 1, "<u4")
            count = int(c["offsets"][-1])
            c["planes"], c["weights"] = r.arr(count, "<u4"), r.arr(count, "<f8")
            c["points"] = r.u32()
            c["momenta"] = r.arr(c["points"] * p["legs"] * 4, "<f8")
            c["values"] = r.arr(c["points"] * p["helicities"], "<f8")
            if outputs != p["helicities"] or c["external_count"] != p["legs"] or c["tile"] == 0 or c["offsets"][0] != 0:
                raise ValueError(f"process {p['key']} colour {c['id']}: inconsistent template")
            cols.append(c)
        p["colours"] = cols
        procs.append(p)
    if r.at != len(b):
        raise ValueError("trailing bytes")
    f["processes"] = procs
    return f
