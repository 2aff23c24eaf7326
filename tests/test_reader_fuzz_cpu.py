# SPDX-License-Identifier: 0BSD
"""The recording and examples.json readers of pacuda-check (pacuda_amd/csrc/pacuda_check.cpp) on
corrupted inputs, built with AddressSanitizer and UndefinedBehaviorSanitizer: truncations, bit
flips, overwritten lengths and deeply nested JSON are refused or read, never a memory error, an
undefined operation or a crash. The inputs are the golden recording and the examples.json the
tool writes from it, corrupted by a seeded generator."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden", "gg_ggg_qqbar.pacuda")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    out = str(tmp_path_factory.mktemp("fuzz") / "pacuda-check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fno-omit-frame-pointer", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "pacuda_amd", "csrc", "pacuda_check.cpp"), "-o", out, "-ldl"], check=True,
                   timeout=300)
    return out


def _run(tool, args):
    p = subprocess.run([tool, *args], capture_output=True, timeout=60,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0"))
    err = p.stderr.decode("utf-8", "replace")
    assert "AddressSanitizer" not in err and "runtime error" not in err and p.returncode in (0, 1, 2), err[-3000:]
    return p.returncode


def test_corrupted_recordings(tool, tmp_path):
    src = open(GOLDEN, "rb").read()
    assert _run(tool, [GOLDEN, "--dump"]) == 0
    rng = np.random.default_rng(11)
    path = str(tmp_path / "f.pacuda")
    codes = []
    for t in range(120):
        b = bytearray(src)
        if t % 3 == 0:
            b = b[:rng.integers(0, len(b))]
        elif t % 3 == 1:
            for _ in range(rng.integers(1, 4)):
                k = rng.integers(0, len(b))
                b[k] ^= 1 << int(rng.integers(0, 8))
        else:  # a length or count field made huge
            k = rng.integers(0, len(b) - 4)
            b[k:k + 4] = bytes([0xff, 0xff, 0xff, 0x7f])
        open(path, "wb").write(b)
        codes.append(_run(tool, [path, "--dump"]))
    assert codes.count(1) > 0  # (some are refused)


def test_corrupted_examples(tool, tmp_path):
    examples = str(tmp_path / "ex.json")
    assert _run(tool, [GOLDEN, "--write-examples", examples]) == 0
    src = open(examples, "rb").read()
    rng = np.random.default_rng(12)
    path = str(tmp_path / "fx.json")
    for t in range(120):
        b = bytearray(src)
        if t % 3 == 0:
            b = b[:rng.integers(0, len(b))]
        elif t % 3 == 1:
            for _ in range(rng.integers(1, 4)):
                k = rng.integers(0, len(b))
                b[k] = int(rng.choice(list(b'{}[],:"0123456789.eE-+ntrufals')))
        else:  # deep nesting
            k = rng.integers(0, len(b))
            b[k:k] = b"[" * int(rng.integers(1, 3000))
        open(path, "wb").write(b)
        _run(tool, [GOLDEN, "--examples", path, "--dump"])
