"""CPU checks of libpadeops_hip's host planner and C ABI (no GPU calls).

The NumPy emulator (tests/spike_emulator.py) runs the device algorithm on the
tables the library exports, so the planner (coefficients, closures, band LU
of each partition, coupling columns, reduced-system inverse, transposed
closures of the adjoint) is checked against the dense oracle on the CPU.
"""
import ctypes
import re

import numpy as np
import pytest

import spike_emulator as se
from oracle import pade_dense as pd
from padeops_amd import _lib

OPS = [("cd06", "d1"), ("cd06", "d2"), ("cd10", "d1"), ("cd10", "d2"), ("cf90", "filter")]


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    hdr = open(_lib.LIB_PATH.replace("padeops_amd/lib/libpadeops_hip.so", "include/padeops_hip.h")).read()
    declared = set(re.findall(r"\b(pade_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.pade_version() >= 100


@pytest.mark.parametrize("kind,which", OPS)
@pytest.mark.parametrize("periodic", [True, False])
def test_planner_matches_dense_oracle(kind, which, periodic):
    if kind == "cd06" and which == "d2" and not periodic:
        pytest.skip("bounded cd06 d2 is composed from d1")
    rng = np.random.default_rng(7)
    if periodic:
        bcs = [(0, 0)]
    elif kind == "cd06":
        bcs = [(0, 0), (1, 1), (0, 1)]
    else:
        bcs = [(a, b) for a in (0, 1, -1) for b in (0, 1, -1)]
    for bc in bcs:
        for n, M in [(64, 16), (64, 32), (64, 64), (96, 32), (48, 16), (40, 40), (512, 32), (256, 16)]:
            v = rng.standard_normal(n)
            D = pd.dense_operator(kind, which, n, 0.3, periodic, *bc)
            for adj in (False, True):
                ref = (D.T if adj else D) @ v
                out = se.apply_pencil(kind, which, v, 0.3, periodic, *bc, adjoint=adj, M=M)
                err = np.abs(out - ref).max() / np.abs(ref).max()
                assert err < 1e-12, (kind, which, periodic, bc, n, M, adj, err)


def test_periodic_coefficients_collapse_to_one_class():
    tab, m = se.export("cd10", "d1", 512, 0.1, True, 0, 0, False, 32)
    assert m["P"] == 16 and m["ncls"] == 1
    tab, m = se.export("cf90", "filter", 512, 0.1, False, 0, 0, False, 32)
    assert m["ncls"] == 3 and m["cls"][0] != m["cls"][1] == m["cls"][14] != m["cls"][15]


def test_reduced_system_window():
    # cd10 d1 couplings decay as ~0.56^M per partition; cf90's as ~0.8^M
    _, m = se.export("cd10", "d1", 512, 0.1, True, 0, 0, False, 64)
    assert m["kq"] <= 2
    _, m = se.export("cf90", "filter", 512, 0.1, True, 0, 0, False, 32)
    assert m["kq"] >= 2


def test_export_error_codes():
    lib = _lib.load()
    meta = (ctypes.c_int * 32)()
    assert lib.pade_plan_export(2, 1, 5, 0.1, 1, 0, 0, 0, 0, None, 0, meta) == 2     # cd10 n < 8
    assert lib.pade_plan_export(1, 1, 5, 0.1, 1, 0, 0, 0, 0, None, 0, meta) == 3     # cd06 periodic n < 6
    assert lib.pade_plan_export(3, 3, 9, 0.1, 0, 0, 0, 0, 0, None, 0, meta) == 7     # cf90 n < 10
    assert lib.pade_plan_export(2, 1, 32, 0.1, 0, 2, 0, 0, 0, None, 0, meta) == 324  # bad bc
    assert lib.pade_plan_export(3, 1, 32, 0.1, 1, 0, 0, 0, 0, None, 0, meta) == 21   # cf90 has no d1
    assert b"small" in lib.pade_last_error() or lib.pade_last_error()
