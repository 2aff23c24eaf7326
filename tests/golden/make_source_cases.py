# SPDX-License-Identifier: 0BSD
"""Writes tests/golden/source_cases.npz: the source-wavefunction (op 12) test vectors.

Inputs are the cases of the reference's two generators (tests/check_sources.py:75-90, seed 3, 300
per family; rusticol/patches/0009-...patch:64-117, xorshift64 from 0x9e3779b97f4a7c15) as
restated in tests/source_cases.py; the expected values are the closed forms of
convert/CONVENTIONS.md:52-61 evaluated by tests/source_cases.legend in Python complex arithmetic.
The GPU test loads this data file; the restated generators and legend stay on the CPU.

Run from the repository root: python tests/golden/make_source_cases.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import source_cases  # noqa: E402

OUT = os.path.join(HERE, "source_cases.npz")


def build():
    cases = source_cases.legend_cases(seed=3, per_family=300) + source_cases.patch_cases()
    n = len(cases)
    codes = np.array([c[0] for c in cases], dtype=np.uint32)
    masses = np.array([c[1] for c in cases], dtype=np.float64)
    momenta = np.array([c[2] for c in cases], dtype=np.float64).reshape(n, 4)
    want = np.zeros((n, 4), dtype=np.complex128)
    for i, (_, m, k, name, h, ch) in enumerate(cases):
        v = [complex(x) for x in source_cases.legend(name, k, h, ch, m)]
        want[i, :len(v)] = v
    origin = np.array([0] * len(source_cases.legend_cases(seed=3, per_family=300)) +
                      [1] * len(source_cases.patch_cases()), dtype=np.uint8)
    return dict(codes=codes, masses=masses, momenta=momenta, legend_re=want.real.copy(),
                legend_im=want.imag.copy(), origin=origin)


if __name__ == "__main__":
    np.savez_compressed(OUT, **build())
    print(OUT)
