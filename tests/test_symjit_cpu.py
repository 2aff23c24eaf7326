# SPDX-License-Identifier: 0BSD
"""Symjit rows (op 13) and their bindings (op 14) without a GPU.

- the generated kernels (pacuda_amd/kernels/*.pk through pacuda-gen) compute what their files
  say: the fermion-pair kernels equal l^T sigma^mu r and l^T sigmabar^mu r built from the Pauli
  matrices in numpy, the probe kernel is the identity;
- the oracle's symjit rows follow runner/recurrence_runner.cu:631-684: on the probe plan of
  tests/symjit_cases.py (every binding code kind, accumulate, direct) its f64 currents equal the
  numpy restatement bit for bit;
- the header the build generates is what pacuda-gen writes from the committed kernel files, and
  names each kernel by its file's SHA-256;
- the library's plan compiler turns op 13/14 into K_SYMJIT instructions and binding tables and
  refuses a symjit row without a valid binding, as the reference does."""
import ctypes as C
import hashlib
import os
import subprocess

import numpy as np
import pytest

from pacuda_amd import abi, build, plans
from pacuda_amd.runtime import plan_desc
import harness
import refcpu
import symjit_cases

SIGMA = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]


def _random_spinors(rng, n):
    return rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))


def test_fermion_pair_kernels_are_the_pauli_bilinears():
    rng = np.random.default_rng(7)
    ls, rs = _random_spinors(rng, 50), _random_spinors(rng, 50)
    for l, r in zip(ls, rs):
        inputs = [l[0].real, l[0].imag, l[1].real, l[1].imag, r[0].real, r[0].imag, r[1].real, r[1].imag]
        for k, sign in ((plans.KERNEL_FERMION_PAIR_SIGMA, 1), (plans.KERNEL_FERMION_PAIR_SIGMABAR, -1)):
            out = refcpu.plane_kernel(k, inputs, 8)
            got = out[0::2] + 1j * out[1::2]
            want = np.array([l @ SIGMA[0] @ r] + [sign * (l @ SIGMA[m] @ r) for m in (1, 2, 3)])
            assert np.allclose(got, want, rtol=0, atol=1e-14 * max(1.0, np.max(np.abs(want)))), (k, got, want)
    x = rng.normal(size=8)
    assert np.array_equal(refcpu.plane_kernel(plans.KERNEL_BINDING_PROBE, x, 8), x)


def test_generated_header_is_current(tmp_path):
    out = tmp_path / "plane_kernels.hpp"
    subprocess.run([build.GEN_TOOL, "hip", *build.kernel_files(), "-o", str(out)], check=True)
    assert out.read_text() == open(build.KERNEL_HEADER).read()
    text = out.read_text()
    for k, path in enumerate(build.kernel_files()):
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        assert f'"{digest}",  // pk_{os.path.basename(path)[:-3]}' in text
    assert f"#define PACUDA_KERNEL_COUNT {len(build.kernel_files())}" in text


def test_emitter_rejects_malformed_kernels(tmp_path):
    bad = {"undefined.pk": "in 1\nout 2\nout0 = add in0 t\nout1 = copy in0\n",
           "twice.pk": "in 1\nout 2\nout0 = copy in0\nout0 = copy in0\nout1 = copy in0\n",
           "missing.pk": "in 1\nout 2\nout0 = copy in0\n",
           "beyond.pk": "in 1\nout 2\nout0 = copy in1\nout1 = copy in0\n",
           "op.pk": "in 1\nout 2\nout0 = pow in0 in0\nout1 = copy in0\n"}
    for name, text in bad.items():
        f = tmp_path / name
        f.write_text(text)
        p = subprocess.run([build.GEN_TOOL, "hip", str(f)], capture_output=True, text=True)
        assert p.returncode != 0, name
        assert "pacuda-gen:" in p.stderr


def _random_momenta(rng, legs, points):
    """massless momenta with E > 0, p_T > 0: ext[(leg*4+k)][point]"""
    ext = np.zeros((4 * legs, points))
    for leg in range(legs):
        p3 = rng.uniform(-300, 300, size=(3, points))
        ext[4 * leg] = np.sqrt((p3 ** 2).sum(axis=0))
        ext[4 * leg + 1:4 * leg + 4] = p3
    return ext


def probe_inputs(points=29, seed=3):
    plan = symjit_cases.probe_plan()
    ext = _random_momenta(np.random.default_rng(seed), 3, points)
    return plan, ext


def sources_of(plan, ext):
    """the source currents of the probe plan from the oracle's wavefunctions"""
    out = {}
    for r in plan.rows:
        if r[0] != "source":
            continue
        _, dst, form, _, _, code = r
        leg = plan.forms[form][0][0]
        n = ext.shape[1]
        mom = np.ascontiguousarray(ext[4 * leg:4 * leg + 4].T).reshape(-1)
        got, _, _ = refcpu.test_sources([code] * n, [0.0] * n, mom)
        name = [k for k, v in plan.layout.items() if v == dst][0]
        comps = 2 if (code & 7) == plans.FAMILY_WEYL else 4
        out[name] = got[:, :comps].T.copy()
    return out


def test_oracle_symjit_rows_follow_the_binding_rules():
    plan, ext = probe_inputs()
    points = ext.shape[1]
    lib = refcpu.load()
    lib.refcpu_reset()
    v1 = harness.V1(lib, prefix="refcpu_v1_")
    assert v1.configure_kernel(abi.KERNEL_F64) == 0
    cur = symjit_cases.run_currents(v1, plan, ext, points)
    forms = refcpu.fill_forms(plan.arrays(), ext, points)
    want = symjit_cases.expected(sources_of(plan, ext), forms)
    for name, value in want.items():
        base = plan.layout[name]
        got = cur[base:base + value.shape[0]]
        assert np.array_equal(got.view(np.uint64), value.view(np.uint64)), (name, np.max(np.abs(got - value)))
    lib.refcpu_reset()


def _program(lib, plan):
    from pacuda_amd.runtime import plan_program
    return plan_program(lib, plan, 0)


def test_library_compiles_symjit_rows():
    lib = abi.load()
    plan = symjit_cases.probe_plan()
    ops, sources, planes, lds = _program(lib, plan)
    symjit = [o for o in ops if o[0] & 0xff == 7]
    assert sources == 3 and len(symjit) == 6
    # w[1] parent 0, w[2] parent 1, w[3] destination, w[4] forms, w[5] factor, w[6] binding
    L = plan.layout
    first = symjit[0]
    assert first[0] >> 16 & 1 == 1 and (first[1], first[2], first[3]) == (L["W"], L["V"], L["D0"])
    assert first[4] == 0 | 1 << 16 and first[6] == 0
    # the second binding starts after the first's 4 + 8 + 4 words
    assert symjit[1][6] == 16 and symjit[1][0] >> 16 & 1 == 0
    # plans with symjit rows keep every current plane in the slot
    assert planes == plan.current_planes


def _status(lib, plan):
    a = plan.arrays()
    d = plan_desc(plan, a)
    n = C.c_uint32()
    return lib.pacuda_plan_program(C.byref(d), 0, None, 0, C.byref(n), None, None, None)


@pytest.mark.parametrize("case", ["no binding", "kernel", "inputs", "outputs"])
def test_invalid_bindings_are_refused(case):
    lib = abi.load()
    plan = symjit_cases.probe_plan()
    rows = [r for r in plan.rows if r[0] == "source"]
    codes = plans.weyl_pair_codes()
    if case == "no binding":
        rows += [("symjit", 1, 0, 2, 12, 0, 0, 0)]
    elif case == "kernel":
        rows += [("binding", 50, codes, [0, 1], 0), ("symjit", 1, 0, 2, 12, 0, 0, 0)]
    elif case == "inputs":
        rows += [("binding", 1, codes + [0], [0], 0), ("symjit", 1, 0, 2, 12, 0, 0, 0)]
    else:
        rows += [("binding", 1, codes, [0, 1, 2, 3, 0], 0), ("symjit", 1, 0, 2, 12, 0, 0, 0)]
    plan.rows = rows
    assert _status(lib, plan) == abi.INVALID
    # the oracle refuses the same programs
    a = plan.arrays()
    with pytest.raises(RuntimeError):
        refcpu.evaluate(a, harness.plan_counts(plan), np.zeros((12, 1)), 1, kernel=1)


def test_build_refuses_a_cuda_kernel_header(tmp_path):
    """pacuda_amd/build.sh's PACUDA_KERNELS takes pacuda-gen's HIP headers only; the reference's
    generator writes CUDA device code, which does not build for gfx950"""
    h = tmp_path / "kernels.cuh"
    h.write_text("// Generated by pacuda-gen from pyAmpliCol's symjit plane kernels.\n")
    with pytest.raises(ValueError):
        build.build_library(kernels_header=str(h), out=str(tmp_path / "lib.so"))
