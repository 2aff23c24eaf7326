# SPDX-License-Identifier: 0BSD
"""The .pacuda v2 recording and the replay tool, on the CPU.

tests/golden/gg_ggg_qqbar.pacuda (written by tests/golden/make_recording.py: two generated
processes, 8 reference points each, values from the CPU oracle) must parse with both readers
(pacuda_amd/pacuda_file.py and the C++ pacuda-check), write back byte for byte, and replay
through the oracle's v1 emulation with every f64 value equal (device and host reductions, odd
tiles). A recording with one value changed in its last bit, a truncated one and one with a
trailing byte are refused."""
import os
import subprocess

import numpy as np
import pytest

from pacuda_amd import build, pacuda_file
import refcpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "gg_ggg_qqbar.pacuda")


def check(path, runner, *args):
    p = subprocess.run([build.CHECK_TOOL, path, "--runner", runner, *args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout + p.stderr


def test_round_trip():
    b = open(GOLDEN, "rb").read()
    f = pacuda_file.read(b)
    assert [p["key"] for p in f["processes"]] == ["g g > g g g", "u u~ > g g g"]
    assert f["processes"][0]["numerator"] == 1 and f["processes"][0]["denominator"] == 4
    assert pacuda_file.write(f["processes"], kernel=f["kernel"], batch=f["batch"], sha256=f["sha256"]) == b


@pytest.mark.parametrize("args", [(), ("--host-reduce",), ("--tile", "3"), ("--host-reduce", "--tile", "5")])
def test_replay_on_oracle(args):
    code, out = check(GOLDEN, refcpu.ABI_LIBRARY, *args)
    assert code == 0, out
    assert out.count("0 of 256 values differ") == 1 and out.count("0 of 128 values differ") == 1


def test_replay_refuses_a_changed_value(tmp_path):
    f = pacuda_file.read(open(GOLDEN, "rb").read())
    v = f["processes"][1]["colours"][0]["values"]
    k = int(np.argmax(v))
    v[k] = np.nextafter(v[k], np.inf)
    bad = tmp_path / "bad.pacuda"
    bad.write_bytes(pacuda_file.write(f["processes"], kernel=1))
    code, out = check(str(bad), refcpu.ABI_LIBRARY)
    assert code == 1 and "1 of 128 values differ" in out


def test_malformed_files(tmp_path):
    b = open(GOLDEN, "rb").read()
    for name, data in [("truncated", b[:-5]), ("trailing", b + b"\0"), ("magic", b"X" + b[1:])]:
        p = tmp_path / f"{name}.pacuda"
        p.write_bytes(data)
        code, out = check(str(p), refcpu.ABI_LIBRARY)
        assert code == 1, name
        with pytest.raises(ValueError):
            pacuda_file.read(data)
