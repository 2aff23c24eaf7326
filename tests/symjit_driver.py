# SPDX-License-Identifier: 0BSD
"""One v1 configuration per process: runs the symjit probe plan (tests/symjit_cases.py) through
pacman_recurrence_v1_run with the given kernel and saves every current plane (complex).
usage: symjit_driver.py KERNEL EXT.npy OUT.npy"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

from pacuda_amd import abi  # noqa: E402
from harness import V1  # noqa: E402
import symjit_cases  # noqa: E402


def main():
    kernel, ext_path, out_path = sys.argv[1], sys.argv[2], sys.argv[3]
    ext = np.load(ext_path)
    v1 = V1(abi.load())
    assert v1.configure_kernel(abi.KERNELS[kernel]) == 0
    np.save(out_path, symjit_cases.run_currents(v1, symjit_cases.probe_plan(), ext, ext.shape[1]))


if __name__ == "__main__":
    main()
