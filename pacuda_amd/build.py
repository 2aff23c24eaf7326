# SPDX-License-Identifier: 0BSD
"""Builds librecurrence.so for gfx950 in-tree (pacuda_amd/lib/) with hipcc, and the CPU oracle
(oracle/, with make) that the tests and the benchmark's CPU baseline use as the checker.

Flags: -ffp-contract=off (no fused multiply-add unless written: the f64 kernel is bit-exact to
the CPU), denormals kept, packed FP32 instructions off."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIBRARY = os.path.join(LIB_DIR, "librecurrence.so")
SOURCES = ["runner.hip", "plan_compiler.hip", "fused.hip", "surrogate.hip", "phase.hip", "unweight.hip"]
KERNEL_DIR = os.path.join(HERE, "kernels")
KERNEL_HEADER = os.path.join(HERE, "csrc", "plane_kernels.hpp")
GEN_TOOL = os.path.join(LIB_DIR, "pacuda-gen")
ARCH = os.environ.get("PACUDA_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no sinking of code common to the arms of a branch into their join: it turns the register
    # tier's per-unit copies into one access through a chosen address, which puts the tier in
    # scratch memory
    "-mllvm", "-simplifycfg-sink-common=false",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


GEN_MARK = "// Generated by pacuda-gen from plane-kernel files"


def build_library(verbose=False, kernels_header=None, out=None):
    """librecurrence.so; `kernels_header`: a pacuda-gen header to build in instead of the one the
    build generates from pacuda_amd/kernels (a directory holding it as plane_kernels.hpp goes
    first on the include path)"""
    import shutil
    import tempfile
    os.makedirs(LIB_DIR, exist_ok=True)
    target = out or LIBRARY
    tmp = target + ".tmp"
    include = []
    if kernels_header:
        with open(kernels_header) as f:
            if not f.readline().startswith(GEN_MARK):
                raise ValueError(f"{kernels_header}: not a pacuda-gen HIP header (a CUDA one cannot be built for gfx950)")
        kdir = tempfile.mkdtemp(prefix="pacuda_kernels_")
        shutil.copy(kernels_header, os.path.join(kdir, "plane_kernels.hpp"))
        include = ["-I", kdir]
    cmd = [hipcc(), *HIPCC_FLAGS, *include, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           *[os.path.join(HERE, "csrc", s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, target)
    return target


CHECK_TOOL = os.path.join(LIB_DIR, "pacuda-check")


def build_check_tool(verbose=False):
    """pacuda-check: the replay of .pacuda recordings through a runner library (host C++)"""
    os.makedirs(LIB_DIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "csrc", "pacuda_check.cpp"), "-o", CHECK_TOOL + ".tmp", "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CHECK_TOOL + ".tmp", CHECK_TOOL)


def kernel_files():
    """the plane-kernel files the runner carries, in kernel-index order (by file name)"""
    return sorted(os.path.join(KERNEL_DIR, f) for f in os.listdir(KERNEL_DIR) if f.endswith(".pk"))


def build_kernels(verbose=False):
    """pacuda-gen (host C++), then the generated kernels' header from pacuda_amd/kernels/*.pk"""
    os.makedirs(LIB_DIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(HERE, "csrc", "pacuda_gen.cpp"), "-o", GEN_TOOL + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(GEN_TOOL + ".tmp", GEN_TOOL)
    cmd = [GEN_TOOL, "hip", *kernel_files(), "-o", KERNEL_HEADER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(KERNEL_HEADER + ".tmp", KERNEL_HEADER)


def build_oracle(verbose=False):
    targets = ["all"]
    if os.path.exists("/root/reference/runner/rectest.c"):
        targets.append("ref")
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "oracle"), *targets]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False):
    build_kernels(verbose)
    build_oracle(verbose)
    build_check_tool(verbose)
    return build_library(verbose)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
