# SPDX-License-Identifier: 0BSD
"""pacman_recurrence_v1_run on the fused kernels: for every v1 kernel (one configuration per
process, tests/v1_fused_driver.py), a sequence of runs and reductions (g g > g g g with a stride
beyond the points and cleared planes no closure writes, g g > g g g g with helicity sampling, a
quark line with the amplitudes downloaded, a four-quark subprocess with symjit rows, a key whose
scales change, a run of 2^18 + 4099 points uploaded in chunks, and the calls that must stay on
the v1 kernels: a written plane not cleared, no key, an uploaded current run, PACUDA_V1_FUSED=0)
equals the oracle's emulation of the v1 runner bit for bit, and exactly the qualifying calls take
the fused kernels."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", ["f64", "df64", "mixed", "mixed64"])
def test_v1_runs_on_the_fused_kernels(kernel):
    print("v1 fused driver", kernel, flush=True)
    env = {k: v for k, v in os.environ.items() if k != "PACUDA_V1_FUSED"}
    p = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "v1_fused_driver.py"), kernel],
                       capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert f"{kernel}: all equal" in p.stdout
