# SPDX-License-Identifier: 0BSD
"""ctypes binding of the CPU oracle (oracle/librefcpu.so). TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker; the product
never loads it."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "librefcpu.so")
ABI_LIBRARY = os.path.join(HERE, "librecurrence_oracle.so")
RECTEST = os.path.join(HERE, "_ref", "rectest")

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: run `make -C oracle`")
        lib = C.CDLL(LIBRARY)
        lib.refcpu_evaluate.restype = C.c_int
        lib.refcpu_evaluate.argtypes = [
            C.c_uint32, u64p, C.c_uint32, u8p, C.c_uint32, C.c_uint32, C.c_uint32,
            f64p, f64p, C.c_uint32, f64p, f64p, C.c_uint32,
            C.c_uint32, u32p, u32p, f64p, C.c_uint32, u32p, u32p, f64p, C.c_double, C.c_double,
            f64p, C.c_uint32, f64p, f64p, C.c_uint32, C.c_int]
        lib.refcpu_evaluate_program.restype = C.c_int
        lib.refcpu_evaluate_program.argtypes = [
            C.c_uint32, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
            f64p, f64p, C.c_uint32, f64p, f64p, C.c_uint32,
            u32p, u32p, f64p, C.c_uint32, u32p, u32p, f64p, C.c_double, C.c_double,
            f64p, C.c_uint32, f64p, f64p, C.c_uint32, C.c_int, u32p, C.c_uint32]
        lib.refcpu_test_sources.restype = C.c_int
        lib.refcpu_test_sources.argtypes = [u32p, f64p, f64p, C.c_uint32, f64p, f64p]
        lib.refcpu_plane_kernel.restype = C.c_int
        lib.refcpu_plane_kernel.argtypes = [C.c_uint32, f64p, f64p]
        lib.refcpu_fill_forms.restype = C.c_int
        lib.refcpu_fill_forms.argtypes = [f64p, C.c_uint32, u32p, C.c_uint32, u32p, f64p, f64p]
        lib.refcpu_reset.restype = None
        for name in ("run", "reduce", "reduce_resolved", "reduce_sample", "configure"):
            getattr(lib, "refcpu_v1_" + name).restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def evaluate(plan_arrays, counts, external, points, kernel=1, block=64, threads=1, resolved=True, total=True,
             numerator=1.0, denominator=1.0):
    """The whole plan run per point on the CPU: (resolved [points][outputs], total [points]).
    kernel: 0 df64, 1 f64, 2 mixed, 3 mixed64 (the v1 ids), 4 and 5 the library's compensated
    mixed policies (df64 and f64 arithmetic over f32 currents)."""
    lib = load()
    a = plan_arrays
    outputs = len(a["offsets"]) - 1
    res = np.zeros((points, outputs)) if resolved else None
    tot = np.zeros(points) if total else None
    ext = np.ascontiguousarray(external, dtype=np.float64)
    st = lib.refcpu_evaluate(
        kernel, _p(a["words"], u64p), len(a["words"]), _p(a["width"], u8p),
        counts["current_planes"], counts["amplitude_planes"], counts["form_count"],
        _p(a["scale_re"], f64p), _p(a["scale_im"], f64p), len(a["scale_re"]),
        _p(a["factor_re"], f64p), _p(a["factor_im"], f64p), len(a["factor_re"]),
        counts["external_count"], _p(a["form_terms"], u32p), _p(a["term_slots"], u32p),
        _p(a["term_coefficients"], f64p), outputs, _p(a["offsets"], u32p), _p(a["planes"], u32p),
        _p(a["weights"], f64p), numerator, denominator, _p(ext, f64p), points,
        None if res is None else _p(res, f64p), None if tot is None else _p(tot, f64p), block, threads)
    if st != 0:
        raise RuntimeError(f"refcpu_evaluate: status {st}")
    return res, tot


def evaluate_program(instr, sources, slot_planes, lds_planes, plan_arrays, counts, external, points, kernel=1,
                     block=64, threads=1, numerator=1.0, denominator=1.0, bindings=None):
    """The whole plan run per point from an instruction stream of the library's fused path
    (uint32 [count][8]; symjit rows index `bindings`, the plan's binding table): (resolved
    [points][outputs], total [points])."""
    lib = load()
    a = plan_arrays
    outputs = len(a["offsets"]) - 1
    res, tot = np.zeros((points, outputs)), np.zeros(points)
    ins = np.ascontiguousarray(instr, dtype=np.uint32)
    ext = np.ascontiguousarray(external, dtype=np.float64)
    bind = None if bindings is None else np.ascontiguousarray(bindings, dtype=np.uint32)
    st = lib.refcpu_evaluate_program(
        kernel, _p(ins, u32p), len(ins), sources, slot_planes, lds_planes, counts["amplitude_planes"],
        counts["form_count"], _p(a["scale_re"], f64p), _p(a["scale_im"], f64p), len(a["scale_re"]),
        _p(a["factor_re"], f64p), _p(a["factor_im"], f64p), len(a["factor_re"]), _p(a["form_terms"], u32p),
        _p(a["term_slots"], u32p), _p(a["term_coefficients"], f64p), outputs, _p(a["offsets"], u32p),
        _p(a["planes"], u32p), _p(a["weights"], f64p), numerator, denominator, _p(ext, f64p), points,
        _p(res, f64p), _p(tot, f64p), block, threads, _p(bind, u32p) if bind is not None else None,
        0 if bind is None else len(bind))
    if st != 0:
        raise RuntimeError(f"refcpu_evaluate_program: status {st}")
    return res, tot


def plane_kernel(k, inputs, outputs):
    """generated symjit kernel k on `inputs` (f64): its first `outputs` results"""
    lib = load()
    a = np.zeros(64)
    a[:len(inputs)] = inputs
    out = np.zeros(64)
    assert lib.refcpu_plane_kernel(k, _p(a, f64p), _p(out, f64p)) == 0
    return out[:outputs]


def fill_forms(plan_arrays, external, points):
    """the plan's momentum forms of `points` points: [form][4][points]"""
    lib = load()
    a = plan_arrays
    count = len(a["form_terms"]) // 2
    ext = np.ascontiguousarray(external, dtype=np.float64)
    mom = np.zeros((count, 4, points))
    assert lib.refcpu_fill_forms(_p(ext, f64p), points, _p(a["form_terms"], u32p), count, _p(a["term_slots"], u32p),
                                 _p(a["term_coefficients"], f64p), _p(mom, f64p)) == 0
    return mom


def test_sources(codes, masses, momenta):
    lib = load()
    n = len(codes)
    codes = np.ascontiguousarray(codes, dtype=np.uint32)
    masses = np.ascontiguousarray(masses, dtype=np.float64)
    momenta = np.ascontiguousarray(momenta, dtype=np.float64)
    re, im = np.zeros(4 * n), np.zeros(4 * n)
    assert lib.refcpu_test_sources(_p(codes, u32p), _p(masses, f64p), _p(momenta, f64p), n, _p(re, f64p), _p(im, f64p)) == 0
    return (re + 1j * im).reshape(n, 4), re.reshape(n, 4), im.reshape(n, 4)
