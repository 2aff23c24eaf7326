# SPDX-License-Identifier: 0BSD
"""The lock scenarios of tests/lock_cases.py against the oracle's v1 emulation (its lock follows
runner/recurrence_runner.cu:1725-1755): they pin the scenarios themselves; tests/test_gpu_lock.py
runs them against the HIP library."""
import numpy as np
import pytest

from pacuda_amd import abi
import harness
import lock_cases
import refcpu


@pytest.fixture
def oracle_v1():
    lib = refcpu.load()
    lib.refcpu_reset()
    v1 = harness.V1(lib, prefix="refcpu_v1_")
    assert v1.configure_kernel(abi.KERNEL_F64) == 0
    yield v1
    lib.refcpu_reset()


def test_concurrent_keep_reduce_pairs_equal_single_thread(oracle_v1):
    single, concurrent = lock_cases.pairs(oracle_v1)
    assert single.keys() == concurrent.keys()
    for k in single:
        assert harness.bits_equal(single[k], concurrent[k]), k


def test_a_kept_run_blocks_other_threads(oracle_v1):
    reduce_at, run_at, unlocked = lock_cases.blocking(oracle_v1)
    assert run_at >= reduce_at, "the other thread's run finished before the holder's reduction"
    assert unlocked < 0.2


def test_a_thread_may_run_again_while_holding(oracle_v1):
    a, b = lock_cases.rerun(oracle_v1)
    assert harness.bits_equal(a, b)
