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


def test_examples_json_round_trip(tmp_path):
    """pacuda-check writes the recording's reference points as pac2json's examples.json
    (convert/src/pac2json.rs:774-804) and replays from such a file: every number reads back to
    its bits, the replay through the oracle agrees bit for bit, a value changed in its last bit is
    found, and a file whose programs do not match the recording's is refused"""
    import json
    ex = tmp_path / "examples.json"
    p = subprocess.run([build.CHECK_TOOL, GOLDEN, "--write-examples", str(ex)], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0, p.stderr
    doc = json.loads(ex.read_text())
    f = pacuda_file.read(open(GOLDEN, "rb").read())
    programs = doc["programs"]
    assert [(e["process"], e["colour"]) for e in programs] == \
        [(pr["key"], c["id"]) for pr in f["processes"] for c in pr["colours"]]
    n = 0
    for pr in f["processes"]:
        for c in pr["colours"]:
            e = programs[n]
            n += 1
            mom = np.array([pt["momenta"] for pt in e["points"]], dtype=np.float64).reshape(-1)
            val = np.array([pt["outputs"] for pt in e["points"]], dtype=np.float64).reshape(-1)
            assert np.array_equal(mom.view(np.uint64), np.asarray(c["momenta"], dtype=np.float64).view(np.uint64))
            assert np.array_equal(val.view(np.uint64), np.asarray(c["values"], dtype=np.float64).view(np.uint64))
            for pt in e["points"]:
                total = 0.0
                for v in pt["outputs"]:
                    total = total + v
                assert pt["total"] == total
    code, out = check(GOLDEN, refcpu.ABI_LIBRARY, "--examples", str(ex))
    assert code == 0 and out.count("0 of 256 values differ") == 1 and out.count("0 of 128 values differ") == 1, out
    # one output in its last bit
    doc["programs"][0]["points"][3]["outputs"][5] = float(np.nextafter(doc["programs"][0]["points"][3]["outputs"][5],
                                                                     np.inf))
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(doc))
    code, out = check(GOLDEN, refcpu.ABI_LIBRARY, "--examples", str(bad))
    assert code == 1 and "1 of 256 values differ" in out
    # the programs must be the recording's, in order
    doc["programs"][0]["colour"] = "2"
    bad.write_text(json.dumps(doc))
    code, out = check(GOLDEN, refcpu.ABI_LIBRARY, "--examples", str(bad))
    assert code == 1 and "is not g g > g g g 1" in out
