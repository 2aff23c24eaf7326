# SPDX-License-Identifier: 0BSD
"""ctypes bindings of librecurrence.so's C ABI (include/pacman_recurrence.h, include/pacuda_amd.h).

The structures mirror runner/pacman_recurrence.h:17-145 field for field (the layout callers such
as umami/umami_pacuda.c compile against); `load()` opens the in-tree HIP library and raises if it
is missing, so nothing here can fall back to another implementation."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "lib", "librecurrence.so")

ABI = 1
OK, INVALID, DEVICE_ERROR = 0, 2, 3
KEEP_AMPLITUDES, FORMS_ON_DEVICE = 1, 2
OUTPUT_ZERO = 1
CAP_SOURCES, CAP_FORMS, CAP_SYMJIT, CAP_FUSED = 1, 2, 4, 8
KERNEL_DF64, KERNEL_F64, KERNEL_MIXED, KERNEL_MIXED64 = 0, 1, 2, 3
KERNELS = {"df64": KERNEL_DF64, "f64": KERNEL_F64, "mixed": KERNEL_MIXED, "mixed64": KERNEL_MIXED64}
# the fused entry points' compensated mixed policies (PACUDA_KERNEL_COMPENSATED: df64 arithmetic,
# PACUDA_KERNEL_COMPENSATED64: f64 arithmetic, both over f32 currents); the v1 configure call
# takes the four above
KERNEL_COMPENSATED, KERNEL_COMPENSATED64 = 4, 5
FUSED_KERNELS = dict(KERNELS, compensated=KERNEL_COMPENSATED, compensated64=KERNEL_COMPENSATED64)
W_P, W_A, W_F, W_X, W_S, W_K = range(6)

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)


class Call(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("struct_size", C.c_uint32),
        ("words", u64p), ("words_key", C.c_uint64), ("word_count", C.c_uint32),
        ("width", C.c_uint8 * 8),
        ("current_re", f64p), ("current_im", f64p), ("amplitude_re", f64p), ("amplitude_im", f64p),
        ("momenta", f64p), ("scale_re", f64p), ("scale_im", f64p), ("factor_re", f64p), ("factor_im", f64p),
        ("upload_current", u32p), ("upload_amplitude", u32p), ("download_amplitude", u32p),
        ("download_current", u32p),
        ("current_planes", C.c_uint32), ("amplitude_planes", C.c_uint32), ("momentum_planes", C.c_uint32),
        ("scale_count", C.c_uint32), ("factor_count", C.c_uint32),
        ("upload_current_runs", C.c_uint32), ("upload_amplitude_runs", C.c_uint32),
        ("download_amplitude_runs", C.c_uint32), ("download_current_runs", C.c_uint32),
        ("stride", C.c_uint32), ("points", C.c_uint32),
        ("zero_amplitude", u32p), ("zero_amplitude_runs", C.c_uint32), ("flags", C.c_uint32),
        ("external_momenta", f64p), ("external_count", C.c_uint32), ("term_count", C.c_uint32),
        ("form_terms", u32p), ("term_slots", u32p), ("term_coefficients", f64p),
    ]


class Reduce(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("planes", u32p), ("weights", f64p),
                ("output", f64p), ("count", C.c_uint32), ("points", C.c_uint32)]


class ReduceResolved(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("planes", u32p), ("weights", f64p),
                ("offsets", u32p), ("output", f64p), ("outputs", C.c_uint32), ("points", C.c_uint32),
                ("output_stride", C.c_uint32), ("point_stride", C.c_uint32), ("flags", C.c_uint32)]


class ReduceSample(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("planes", u32p), ("weights", f64p),
                ("offsets", u32p), ("helicity", u32p), ("outputs", C.c_uint32), ("points", C.c_uint32),
                ("random", f64p), ("scale", f64p), ("numerator", C.c_double), ("denominator", C.c_double),
                ("total", f64p), ("chosen", u32p)]


class Options(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("kernel", C.c_uint32),
                ("block", C.c_uint32), ("no_sort", C.c_uint32), ("stats", C.c_uint32)]


class PlanDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("struct_size", C.c_uint32),
        ("words", u64p), ("word_count", C.c_uint32), ("width", C.c_uint8 * 8),
        ("current_planes", C.c_uint32), ("amplitude_planes", C.c_uint32), ("form_count", C.c_uint32),
        ("scale_re", f64p), ("scale_im", f64p), ("scale_count", C.c_uint32),
        ("factor_re", f64p), ("factor_im", f64p), ("factor_count", C.c_uint32),
        ("external_count", C.c_uint32), ("term_count", C.c_uint32),
        ("form_terms", u32p), ("term_slots", u32p), ("term_coefficients", f64p),
        ("outputs", C.c_uint32), ("offsets", u32p), ("planes", u32p), ("weights", f64p),
        ("numerator", C.c_double), ("denominator", C.c_double), ("no_sort", C.c_uint32),
    ]


class Item(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("external", C.c_void_p), ("points", C.c_uint32), ("ext_stride", C.c_uint32),
                ("resolved", C.c_void_p), ("total", C.c_void_p)]


class Launch(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("struct_size", C.c_uint32), ("kernel", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("block", C.c_uint32), ("unit_waves", C.c_uint32),
                ("stream", C.c_void_p), ("kernel_ms", C.POINTER(C.c_float))]


# the v1 entry points and the extension entry points: (name, restype, argtypes)
V1_SYMBOLS = [
    ("pacman_recurrence_v1_run", C.c_int, [C.POINTER(Call)]),
    ("pacman_recurrence_v1_reduce_resolved", C.c_int, [C.POINTER(ReduceResolved)]),
    ("pacman_recurrence_v1_reduce_sample", C.c_int, [C.POINTER(ReduceSample)]),
    ("pacman_recurrence_v1_host_alloc", C.c_void_p, [C.c_size_t]),
    ("pacman_recurrence_v1_host_free", None, [C.c_void_p]),
    ("pacman_recurrence_v1_reduce", C.c_int, [C.POINTER(Reduce)]),
    ("pacman_recurrence_v1_configure", C.c_int, [C.POINTER(Options)]),
    ("pacman_recurrence_v1_capabilities", C.c_uint32, []),
    ("pacman_recurrence_v1_symjit_kernel", C.c_int, [C.c_char_p]),
]
EXTENSION_SYMBOLS = [
    ("pacman_recurrence_v1_test_sources", C.c_int, [u32p, f64p, f64p, C.c_uint32, f64p, f64p]),
    ("pacuda_plan_program", C.c_int, [C.POINTER(PlanDesc), C.c_uint32, u32p, C.c_uint32, u32p, u32p, u32p, u32p]),
    ("pacuda_plan_coop_program", C.c_int, [C.POINTER(PlanDesc), C.c_uint32, C.c_uint32, u32p, C.c_uint32, u32p, u32p, u32p]),
    ("pacuda_plan_bindings", C.c_int, [C.POINTER(PlanDesc), u32p, C.c_uint32, u32p]),
    ("pacuda_plan_create", C.c_int, [C.POINTER(PlanDesc), C.POINTER(C.c_void_p)]),
    ("pacuda_plan_destroy", None, [C.c_void_p]),
    ("pacuda_plan_info", C.c_int, [C.c_void_p, u32p, u64p]),
    ("pacuda_evaluate", C.c_int, [C.POINTER(Item), C.c_uint32, C.POINTER(Launch)]),
    ("pacuda_evaluate_static", C.c_int, [C.POINTER(Item), C.c_uint32, C.POINTER(Launch)]),
    ("pacuda_v1_fused_runs", C.c_uint64, []),
    ("pacuda_device_alloc", C.c_void_p, [C.c_size_t]),
    ("pacuda_device_free", None, [C.c_void_p]),
    ("pacuda_copy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("pacuda_copy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("pacuda_synchronize", C.c_int, []),
    ("pacuda_device_compute_units", C.c_int, []),
    ("pacuda_mlp_create", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("pacuda_mlp_destroy", None, [C.c_void_p]),
    ("pacuda_mlp_param_count", C.c_uint64, [C.c_void_p]),
    ("pacuda_mlp_set_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint64]),
    ("pacuda_mlp_get_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint64]),
    ("pacuda_mlp_get_grads", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint64]),
    ("pacuda_mlp_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
    ("pacuda_mlp_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    ("pacuda_mlp_adam", C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32]),
    ("pacuda_mlp_last_loss", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("pacuda_mlp_train_steps", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_double,
                                         C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    ("pacuda_sweep_create", C.c_int, [C.POINTER(Item), C.c_uint32, C.POINTER(Launch), C.POINTER(C.c_void_p)]),
    ("pacuda_sweep_run", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    ("pacuda_sweep_destroy", None, [C.c_void_p]),
    ("pacuda_rambo", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, u32p, f64p, C.c_void_p,
                               C.c_uint32, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p,
                               C.c_void_p]),
    ("pacuda_rambo_record", C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                      C.c_double, f64p, C.POINTER(C.c_uint64)]),
    ("pacuda_gather", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                C.c_uint32, C.c_void_p]),
    ("pacuda_surrogate_weight", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                          C.c_void_p]),
    ("pacuda_weight_stats", C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.c_void_p,
                                      C.c_void_p]),
    ("pacuda_select", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_uint64,
                                C.c_void_p, C.c_void_p, C.c_void_p]),
]
STATS_BLOCKS = 256
MAX_BINS = 8192


def bind(lib, symbols):
    for name, res, args in symbols:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def load(path=None):
    """The HIP runner library (built in-tree by pacuda_amd/build.py); raises when it is absent."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    p = path or LIBRARY
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python pacuda_amd/build.py` (hipcc, gfx950)")
    lib = bind(C.CDLL(p), V1_SYMBOLS + EXTENSION_SYMBOLS)
    if path is None:
        _lib = lib
    return lib


def ptr(a, t=f64p):
    """A ctypes pointer to a numpy array's data (None for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(t)
