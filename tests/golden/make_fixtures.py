"""Writes tests/golden/pade_fixtures.npz: seeded small inputs and the oracle's
outputs for every operator, periodicity and boundary combination.

The reference ships no stored vectors for this path (its tests print errors
against analytic answers), so these fixtures come from the CPU oracle
(oracle/pade_dense.py), itself pinned by tests/test_oracle.py.  They freeze
expected outputs so GPU results and future oracle edits are checked against
the same numbers.  Run from the repository root:
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pade_dense as pd  # noqa: E402

CASES = []


def add(kind, which, n, periodic, bc, axis, seed):
    CASES.append((kind, which, n, periodic, bc, axis, seed))


for kind, which in [("cd06", "d1"), ("cd06", "d2"), ("cd10", "d1"), ("cd10", "d2"), ("cf90", "filter")]:
    for n in (16, 33):
        add(kind, which, n, True, (0, 0), 1 + n % 3, n)
    if kind == "cd06":
        if which == "d1":
            for bc in ((0, 0), (1, 1)):
                add(kind, which, 24, False, bc, 2, 3)
        continue
    for bc in [(a, b) for a in (0, 1, -1) for b in (0, 1, -1)]:
        add(kind, which, 24, False, bc, 1 + (bc[0] + 1) % 3, 11 + 3 * bc[0] + bc[1])


def main():
    out = {}
    for idx, (kind, which, n, periodic, bc, axis, seed) in enumerate(CASES):
        rng = np.random.default_rng(seed)
        shape = [3, 2]
        shape.insert(axis - 1, n)
        f = rng.standard_normal(shape)
        dx = 0.1 + 0.01 * n
        y = pd.apply(kind, which, f, axis - 1, dx, periodic, *bc)
        key = f"c{idx:03d}"
        out[key + "_f"] = f
        out[key + "_y"] = y
        out[key + "_meta"] = np.array([["cd06", "cd10", "cf90"].index(kind), ["d1", "d2", "filter"].index(which),
                                       n, int(periodic), bc[0], bc[1], axis], dtype=np.int64)
        out[key + "_dx"] = np.array([dx])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "pade_fixtures.npz"), **out)
    print(f"wrote {len(CASES)} cases")


if __name__ == "__main__":
    main()
