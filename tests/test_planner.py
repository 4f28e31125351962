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


STAGG_CODES = {"d1_e2c": 11, "d1_c2e": 12, "d2_c2c": 13, "d2_e2e": 14, "interp_e2c": 15, "interp_c2e": 16,
               "d1_c2c": 17, "d1_e2e": 18}
STAGG_ENDS = {1: (True, False), -1: (False, False), 2: (True, True), -2: (False, True)}


@pytest.mark.parametrize("op", sorted(STAGG_CODES))
def test_planner_stagg_bounded_matches_oracle(op):
    """Bounded staggered operators: the planner's mirror construction and
    one-sided tables, run through the exported tables the way the kernels run
    them, equal the oracle's statement-by-statement restatement of the
    reference, for every end combination; at n = 64 also through the
    partitioned solve (the system form for operators with an edge array)."""
    rng = np.random.default_rng(3)
    for bc1 in STAGG_ENDS:
        for bcn in STAGG_ENDS:
            for n in (5, 6, 7, 12, 37, 64):
                nin = n + (1 if pd.STAGG_EDGE_IN[op] else 0)
                v = rng.standard_normal(nin)
                D = pd.stagg_bounded_operator(op, n, 0.3, STAGG_ENDS[bc1], STAGG_ENDS[bcn])
                ref = D @ v
                # partitioned forms: square cell operators as they are, the
                # others in system form (n rows solved, plus the tail row)
                Ms = [0] + ([16, 32, 64] if n == 64 else [])
                for M in Ms:
                    out, m = se.apply_stagg(STAGG_CODES[op], v, 0.3, bc1, bcn, M)
                    assert m["n"] + m["tail"] == D.shape[0] and m["n_in"] == D.shape[1]
                    assert m["sys"] == (M > 0 and (pd.STAGG_EDGE_IN[op] or pd.STAGG_EDGE_OUT[op]))
                    err = np.abs(out - ref).max() / np.abs(ref).max()
                    assert err < 1e-12, (op, bc1, bcn, n, M, err)


def test_planner_stagg_bounded_codes():
    lib = _lib.load()
    meta = (ctypes.c_int * 32)()
    for bc in (0, 3, -3):
        rc = lib.pade_plan_export(4, 17, 16, 0.1, 0, bc, 1, 0, 0, None, 0, meta)
        assert rc == 324
    # ddz_C2C / ddz_E2E have no periodic form
    assert lib.pade_plan_export(4, 17, 16, 0.1, 1, 0, 0, 0, 0, None, 0, meta) == 21


@pytest.mark.parametrize("n,M", [(64, 16), (64, 32), (512, 64), (37, 37), (9, 9)])
def test_planner_gaussian(n, M):
    """The Gaussian filter (identity A, 9-point B), periodic and bounded with
    every bc pair, through the exported tables, forward and adjoint."""
    v = np.random.default_rng(n).standard_normal(n)
    cases = [(True, 0, 0)] + [(False, a, b) for a in (0, 1, -1) for b in (0, 1, -1)]
    for periodic, b1, bn in cases:
        G = pd.dense_operator("gaussian", "filter", n, 1.0, periodic, b1, bn)
        for adj in (False, True):
            if adj and not periodic and n < 16:
                continue                                     # bounded adjoints need n >= 16
            out = se.apply_pencil("gaussian", "filter", v, 1.0, periodic, b1, bn, adjoint=adj, M=M)
            ref = (G.T if adj else G) @ v
            assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-13, (periodic, b1, bn, adj)
    meta = (ctypes.c_int * 32)()
    lib = _lib.load()
    assert lib.pade_plan_export(5, 3, 8, 1.0, 1, 0, 0, 0, 0, None, 0, meta) == 21   # shorter than the stencil


def test_compressible_kernels_reject_bad_arguments():
    """The compressible-step entry points check their arguments on the host
    (no GPU call): empty fields or null arrays give 400."""
    lib = _lib.load()
    nul = None
    assert lib.pade_cg_conserved(0, nul, nul, nul, nul, nul, nul, nul, nul) == 400
    assert lib.pade_cg_primitive(8, nul, 1.4, 1.0, nul, nul, nul, nul, nul, nul, nul, nul) == 400
    assert lib.pade_cg_strain(8, nul, nul, nul, nul) == 400
    assert lib.pade_cg_lad_mu(-1, 0.002, nul, nul, nul) == 400
    assert lib.pade_cg_lad_bulk(8, 1.75, 0.1, 0.1, 0.1, nul, nul, nul, nul, nul, nul, nul) == 400
    assert lib.pade_cg_lad_kap(8, 0.01, 1.4, 0.1, 0.1, 0.1, nul, nul, nul, nul, nul, nul, nul) == 400
    assert lib.pade_cg_fluxes(8, nul, nul, nul, nul, nul, nul, nul, nul, nul, nul, nul, nul) == 400
    assert lib.pade_cg_rk_stage(8, 1e-3, 0.0, 0.1, nul, nul, nul, nul) == 400
