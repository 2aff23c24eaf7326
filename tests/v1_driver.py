# SPDX-License-Identifier: 0BSD
"""One v1 configuration per process: runs an n-gluon plan through pacman_recurrence_v1_run /
reduce_resolved with the given kernel and saves [points][outputs].
usage: v1_driver.py KERNEL EXT.npy OUT.npy N"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

from pacuda_amd import abi, plans  # noqa: E402
from harness import V1  # noqa: E402


def main():
    kernel, ext_path, out_path, n = sys.argv[1], sys.argv[2], sys.argv[3], int(sys.argv[4])
    ext = np.load(ext_path)
    plan = plans.gluon_plan(n)
    v1 = V1(abi.load())
    assert v1.configure_kernel(abi.KERNELS[kernel]) == 0
    np.save(out_path, v1.run_plan(plan, ext, ext.shape[1]))


if __name__ == "__main__":
    main()
