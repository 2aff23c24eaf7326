# SPDX-License-Identifier: 0BSD
"""Host-side helpers over the fused entry points of librecurrence.so (include/pacuda_amd.h):
device buffers, compiled plans, and one-launch evaluation of many (plan, points) items."""
import ctypes as C

import numpy as np

from . import abi
from .abi import ptr, u32p, u64p


class DeviceArray:
    """A device buffer owned by the HIP library (pacuda_device_alloc)."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.nbytes = int(nbytes)
        self.ptr = lib.pacuda_device_alloc(self.nbytes)
        if not self.ptr:
            raise RuntimeError(f"device allocation of {nbytes} bytes failed")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if self.lib.pacuda_copy_to_device(self.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError("copy to device failed")

    def download(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("shape larger than the buffer")
        if self.lib.pacuda_copy_to_host(out.ctypes.data, self.ptr, out.nbytes) != 0:
            raise RuntimeError("copy to host failed")
        return out

    def free(self):
        if self.ptr and self.owner:
            self.lib.pacuda_device_free(self.ptr)
        self.ptr = None

    owner = True

    @classmethod
    def view(cls, base, offset_bytes, nbytes=None):
        """a non-owning window into `base` starting `offset_bytes` in"""
        v = cls.__new__(cls)
        v.lib = base.lib
        v.nbytes = base.nbytes - offset_bytes if nbytes is None else nbytes
        if offset_bytes < 0 or offset_bytes + v.nbytes > base.nbytes:
            raise ValueError("view outside the buffer")
        v.ptr = base.ptr + offset_bytes
        v.owner = False
        return v


class Device:
    def __init__(self, lib=None):
        self.lib = lib or abi.load()

    def array(self, a):
        a = np.ascontiguousarray(a)
        d = self.empty(a.nbytes, np.uint8)
        d.upload(a)
        return d

    def empty(self, count, dtype=np.float64):
        return DeviceArray(self.lib, int(count) * np.dtype(dtype).itemsize)

    def synchronize(self):
        if self.lib.pacuda_synchronize() != 0:
            raise RuntimeError("device synchronisation failed")


class PooledDevice(Device):
    """A Device whose buffers come from a pool and go back to it when freed, for drivers that
    allocate and free large buffers on every call. Only for buffers whose every use is queued on
    the library's stream (the library's kernels and copies run there): a buffer freed while work
    that uses it is still queued may be handed out again at once, and the stream orders the new
    work after the old. A free buffer serves a request of at least half its size less 512 KiB
    (the smallest that fits). close() releases the pool (after the device has finished with it)."""

    def __init__(self, lib=None):
        super().__init__(lib)
        self._free = []  # (nbytes, base DeviceArray) of the buffers not handed out
        self._closed = False

    def empty(self, count, dtype=np.float64):
        nbytes = max(1, int(count) * np.dtype(dtype).itemsize)
        fits = [k for k, (size, _) in enumerate(self._free) if nbytes <= size <= 2 * nbytes + (1 << 20)]
        if fits:
            k = min(fits, key=lambda j: self._free[j][0])
            _, base = self._free.pop(k)
        else:
            base = DeviceArray(self.lib, nbytes)
        out = DeviceArray.view(base, 0, nbytes)

        def give_back(a=out, b=base):
            if a.ptr is None:
                return
            a.ptr = None
            if self._closed:  # (freed after close: straight back to the library)
                b.free()
            else:
                self._free.append((b.nbytes, b))
        out.free = give_back
        return out

    def close(self):
        self.synchronize()
        for _, b in self._free:
            b.free()
        self._free = []
        self._closed = True


def plan_desc(plan, arrays, no_sort=0):
    """the pacuda_plan_desc of a plan (pacuda_amd.plans.Plan) over its arrays (kept by the caller)"""
    a = arrays
    d = abi.PlanDesc()
    d.abi_version = abi.ABI
    d.struct_size = C.sizeof(abi.PlanDesc)
    d.words = ptr(a["words"], u64p)
    d.word_count = len(a["words"])
    for k in range(8):
        d.width[k] = int(a["width"][k])
    d.current_planes = plan.current_planes
    d.amplitude_planes = plan.amplitude_planes
    d.form_count = len(plan.forms)
    d.scale_re, d.scale_im, d.scale_count = ptr(a["scale_re"]), ptr(a["scale_im"]), len(a["scale_re"])
    d.factor_re, d.factor_im, d.factor_count = ptr(a["factor_re"]), ptr(a["factor_im"]), len(a["factor_re"])
    d.external_count = plan.legs
    d.term_count = len(a["term_slots"])
    d.form_terms, d.term_slots = ptr(a["form_terms"], u32p), ptr(a["term_slots"], u32p)
    d.term_coefficients = ptr(a["term_coefficients"])
    d.outputs = len(a["offsets"]) - 1
    d.offsets, d.planes, d.weights = ptr(a["offsets"], u32p), ptr(a["planes"], u32p), ptr(a["weights"])
    d.numerator, d.denominator = float(plan.numerator), float(plan.denominator)
    d.no_sort = no_sort
    return d


# the cooperative shapes (pacuda_plan_coop_program): (wavefronts per unit, waves per CU)
COOP_SHAPES = ((16, 16), (8, 8), (8, 16), (4, 16))


def plan_bindings(lib, plan):
    """the binding table a plan's programs index (pacuda_plan_bindings, host only): uint32 array"""
    a = plan.arrays()
    d = plan_desc(plan, a)
    n = C.c_uint32()
    if lib.pacuda_plan_bindings(C.byref(d), None, 0, C.byref(n)) != 0:
        raise RuntimeError("pacuda_plan_bindings failed")
    out = np.zeros(max(n.value, 1), dtype=np.uint32)
    if lib.pacuda_plan_bindings(C.byref(d), out.ctypes.data_as(u32p), len(out), C.byref(n)) != 0:
        raise RuntimeError("pacuda_plan_bindings failed")
    return out[:n.value]


def coop_program(lib, plan, shape, serial=False, f32=False):
    """a plan's cooperative program for shape index `shape` (pacuda_plan_coop_program, host
    only), for 16-byte currents or (f32) the f32 ones: (uint32 [count][8], info dict, stream
    starts), or None when the plan has none"""
    shape = shape + (4 if f32 else 0)
    a = plan.arrays()
    d = plan_desc(plan, a)
    n, info, starts = C.c_uint32(), (C.c_uint32 * 6)(), (C.c_uint32 * 17)()
    if lib.pacuda_plan_coop_program(C.byref(d), shape, int(serial), None, 0, C.byref(n), info, starts) != 0:
        return None
    out = np.zeros((n.value, 8), dtype=np.uint32)
    st = lib.pacuda_plan_coop_program(C.byref(d), shape, int(serial), out.ctypes.data_as(u32p), n.value, C.byref(n),
                                      info, starts)
    if st != 0:
        raise RuntimeError(f"pacuda_plan_coop_program: status {st}")
    keys = ("waves", "phases", "lds_planes", "slot_planes", "slot_accesses", "lds_accesses")
    info = dict(zip(keys, list(info)))
    return out, info, list(starts)[:info["waves"] + 1]


def plan_program(lib, plan, variant):
    """the fused path's instruction stream of a plan (pacuda_plan_program, host only): (uint32
    [count][8], leading sources, slot planes, LDS planes)"""
    a = plan.arrays()
    d = plan_desc(plan, a)
    n, src, planes, lds = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    # (one compilation when a generous buffer holds the program, else a count query first)
    guess = np.zeros((4 * len(plan.rows) + 64, 8), dtype=np.uint32)
    st = lib.pacuda_plan_program(C.byref(d), variant, guess.ctypes.data_as(u32p), len(guess), C.byref(n), C.byref(src),
                                 C.byref(planes), C.byref(lds))
    if st == 0:
        return guess[:n.value].copy(), src.value, planes.value, lds.value
    st = lib.pacuda_plan_program(C.byref(d), variant, None, 0, C.byref(n), None, None, None)
    if st != 0:
        raise RuntimeError(f"pacuda_plan_program: status {st}")
    out = np.zeros((n.value, 8), dtype=np.uint32)
    st = lib.pacuda_plan_program(C.byref(d), variant, out.ctypes.data_as(u32p), n.value, C.byref(n), C.byref(src),
                                 C.byref(planes), C.byref(lds))
    if st != 0:
        raise RuntimeError(f"pacuda_plan_program: status {st}")
    return out, src.value, planes.value, lds.value


class FusedPlan:
    """A plan (pacuda_amd.plans.Plan) compiled into the library (pacuda_plan_create)."""

    def __init__(self, lib, plan, no_sort=0):
        self.lib = lib
        self.plan = plan
        a = plan.arrays()
        self.arrays = a
        d = plan_desc(plan, a, no_sort)
        h = C.c_void_p()
        st = lib.pacuda_plan_create(C.byref(d), C.byref(h))
        if st != 0:
            raise RuntimeError(f"pacuda_plan_create: status {st}")
        self.handle = h

    def info(self):
        n, slot = C.c_uint32(), C.c_uint64()
        self.lib.pacuda_plan_info(self.handle, C.byref(n), C.byref(slot))
        return n.value, slot.value

    def destroy(self):
        if self.handle:
            self.lib.pacuda_plan_destroy(self.handle)
            self.handle = None


def evaluate(lib, jobs, kernel, waves_per_cu=0, block=0, timed=False, unit_waves=0):
    """jobs: [(FusedPlan, DeviceArray external, points, ext_stride, DeviceArray|None resolved,
    DeviceArray|None total)]; one persistent launch (cooperative when unit_waves is 16 or 8: that
    many wavefronts share each 64-point unit). Returns the kernel's time in ms (HIP events) when
    timed, else waits for the device and returns 0."""
    items = (abi.Item * max(1, len(jobs)))()
    for k, (fp, ext, points, stride, res, tot) in enumerate(jobs):
        items[k].plan = fp.handle
        items[k].external = ext.ptr
        items[k].points = points
        items[k].ext_stride = stride
        items[k].resolved = res.ptr if res is not None else None
        items[k].total = tot.ptr if tot is not None else None
    ms = C.c_float(0)
    l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), kernel, waves_per_cu, block, unit_waves, None,
                   C.pointer(ms) if timed else None)
    st = lib.pacuda_evaluate(items, len(jobs), C.byref(l))
    if st != 0:
        raise RuntimeError(f"pacuda_evaluate: status {st}")
    if not timed:
        if lib.pacuda_synchronize() != 0:
            raise RuntimeError("device synchronisation failed")
    return ms.value


class Sweep:
    """A batch of items captured once into a HIP graph (pacuda_sweep_create) and replayed"""

    def __init__(self, lib, jobs, kernel, waves_per_cu=0, block=0, unit_waves=0):
        self.lib = lib
        items = (abi.Item * len(jobs))()
        for k, (fp, ext, points, stride, res, tot) in enumerate(jobs):
            items[k].plan = fp.handle
            items[k].external = ext.ptr
            items[k].points = points
            items[k].ext_stride = stride
            items[k].resolved = res.ptr if res is not None else None
            items[k].total = tot.ptr if tot is not None else None
        l = abi.Launch(abi.ABI, C.sizeof(abi.Launch), kernel, waves_per_cu, block, unit_waves, None, None)
        h = C.c_void_p()
        st = lib.pacuda_sweep_create(items, len(jobs), C.byref(l), C.byref(h))
        if st != 0:
            raise RuntimeError(f"pacuda_sweep_create: status {st}")
        self.handle = h

    def run(self, timed=False):
        ms = C.c_float(0)
        st = self.lib.pacuda_sweep_run(self.handle, None, C.pointer(ms) if timed else None)
        if st != 0:
            raise RuntimeError(f"pacuda_sweep_run: status {st}")
        return ms.value

    def destroy(self):
        if self.handle:
            self.lib.pacuda_sweep_destroy(self.handle)
            self.handle = None


# (waves per CU, block) pairs the autotuner tries: 2 to 6 persistent waves per SIMD, blocks of
# one or four waves
TUNE_SHAPES = ((6, 64), (8, 128), (8, 256), (12, 256), (16, 256), (16, 64))
# ... and the cooperative launches (waves per CU, block ignored, wavefronts per unit), for plans
# without symjit rows: the f64 kernel's, and the two the other kernels run
COOP_TUNE_SHAPES = tuple((per_cu, 0, uw) for uw, per_cu in COOP_SHAPES)
COOP_TUNE_SHAPES_ANY = ((16, 0, 16), (16, 0, 8))
_tuned = {}


def shape_name(shape):
    """"WxB" for (waves_per_cu, block), "coop UxW" for a cooperative launch (wavefronts per unit
    U, waves per CU W)"""
    return f"coop {shape[2]}x{shape[0]}" if len(shape) > 2 and shape[2] else f"{shape[0]}x{shape[1]}"


def autotune(lib, jobs, kernel, shapes=TUNE_SHAPES, repeats=2, key=None):
    """the launch shape of `shapes` -- (waves_per_cu, block) or (waves_per_cu, block, unit_waves)
    -- with the shortest persistent launch of `jobs` (the minimum of `repeats` timed launches
    after one warm-up each); cached under `key` when given. Returns (shape, {name: ms})."""
    if key is not None and key in _tuned:
        return _tuned[key]
    table, by_name = {}, {}
    for shape in shapes:
        wpc, block = shape[0], shape[1]
        uw = shape[2] if len(shape) > 2 else 0
        evaluate(lib, jobs, kernel, waves_per_cu=wpc, block=block, timed=True, unit_waves=uw)
        name = shape_name(shape)
        table[name] = min(evaluate(lib, jobs, kernel, waves_per_cu=wpc, block=block, timed=True, unit_waves=uw)
                          for _ in range(repeats))
        by_name[name] = tuple(shape)
    best = min(table, key=table.get)
    out = (by_name[best], table)
    if key is not None:
        _tuned[key] = out
    return out
