# SPDX-License-Identifier: 0BSD
"""The v1 lock contract on the HIP library (tests/lock_cases.py: concurrent KEEP + reduce pairs
equal a single thread bit for bit; a kept run blocks other threads until its reduction; a
holder may run again), and the statistics line a configuration with `stats` prints at exit."""
import os
import re
import subprocess
import sys

import pytest

from pacuda_amd import abi
import harness
import lock_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def v1():
    v = harness.V1(abi.load())
    assert v.configure_kernel(abi.KERNEL_F64) == 0
    return v


def test_concurrent_keep_reduce_pairs_equal_single_thread(v1):
    print("lock: 4 threads x 5 pairs", flush=True)
    single, concurrent = lock_cases.pairs(v1)
    for k in single:
        assert harness.bits_equal(single[k], concurrent[k]), k


def test_concurrent_keyed_pairs_on_the_fused_kernels(v1):
    """the same pairs with keyed calls, which run on the fused kernels: every run does, and the
    results equal the unkeyed (v1 kernels') ones bit for bit"""
    print("lock: 4 threads x 5 keyed pairs", flush=True)
    lib = v1.lib
    before = lib.pacuda_v1_fused_runs()
    single, concurrent = lock_cases.pairs(v1, key=0x10C4)
    assert lib.pacuda_v1_fused_runs() - before == 2 * len(single)
    plain, _ = lock_cases.pairs(v1, threads=4, per_thread=5)
    for k in single:
        assert harness.bits_equal(single[k], concurrent[k]), k
        assert harness.bits_equal(single[k], plain[k]), k


def test_a_kept_run_blocks_other_threads(v1):
    print("lock: blocking", flush=True)
    reduce_at, run_at, unlocked = lock_cases.blocking(v1)
    assert run_at >= reduce_at, "the other thread's run finished before the holder's reduction"
    assert unlocked < 0.2


def test_a_thread_may_run_again_while_holding(v1):
    print("lock: rerun", flush=True)
    a, b = lock_cases.rerun(v1)
    assert harness.bits_equal(a, b)


def test_statistics_line_at_exit():
    p = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "stats_driver.py")],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    m = re.search(r"pacman recurrence runner: (\d+) calls, (\d+) points, (\d+) reductions, (\d+) copies", p.stderr)
    assert m, p.stderr
    calls, points, reductions, copies = map(int, m.groups())
    assert (calls, points, reductions) == (2, 128, 2)
    assert copies > 0
