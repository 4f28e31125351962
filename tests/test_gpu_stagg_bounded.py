"""Bounded staggered z operators (cd06stagg.F90 init_nonperiodic :197-232)
through the C ABI against the oracle's restatement of the reference's rows:
every operation and every combination of even, odd and one-sided ends, FP64
to 1e-12 (partitioned kernels where the operator is square and its length a
multiple of 16, the one-lane-per-pencil kernel otherwise), FP32 to 1e-4,
host arrays, complex data, and the error codes."""
import ctypes
import itertools

import numpy as np
import pytest
import torch

from oracle import pade_dense as pd
import padeops_amd as po
from padeops_amd import _lib

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
OPS = {"d1_e2c": "ddz_E2C", "d1_c2e": "ddz_C2E", "d1_c2c": "ddz_C2C", "d1_e2e": "ddz_E2E",
       "d2_c2c": "d2dz2_C2C", "d2_e2e": "d2dz2_E2E", "interp_e2c": "InterpZ_E2C", "interp_c2e": "InterpZ_C2E"}
# (even, sided) per end
ENDS = {"E": (True, False), "O": (False, False), "SE": (True, True), "SO": (False, True)}


def _make(nz, dz, bot, top):
    s = po.cd06stagg()
    s.init(nz, dz, isTopEven=ENDS[top][0], isBotEven=ENDS[bot][0],
           isTopSided=ENDS[top][1], isBotSided=ENDS[bot][1])
    return s


def _field(n1, n2, m, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n1, n2, m))
    if dtype == np.complex128:
        f = f + 1j * rng.standard_normal((n1, n2, m))
    return np.asfortranarray(f)


def _ref(op, f, dz, bot, top):
    if np.iscomplexobj(f):
        return (pd.stagg_bounded_apply(op, np.asfortranarray(f.real), dz, ENDS[bot], ENDS[top])
                + 1j * pd.stagg_bounded_apply(op, np.asfortranarray(f.imag), dz, ENDS[bot], ENDS[top]))
    return pd.stagg_bounded_apply(op, f, dz, ENDS[bot], ENDS[top])


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _m(op, nz, edge):
    return nz + (1 if edge else 0)


@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("bot,top", list(itertools.product(sorted(ENDS), sorted(ENDS))))
@pytest.mark.parametrize("nz,n1,n2", [(64, 12, 5), (37, 9, 4)])
def test_stagg_bounded_fp64(op, bot, top, nz, n1, n2):
    dz = 1.0 / nz
    s = _make(nz, dz, bot, top)
    f = _field(n1, n2, _m(op, nz, pd.STAGG_EDGE_IN[op]), nz + len(op) + len(bot) * 3 + len(top))
    ref = _ref(op, f, dz, bot, top)
    ft = po.as_fortran(torch.tensor(f, device=DEV))
    out = po.fortran_empty(ref.shape, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-12, (op, bot, top, nz, relerr(out.cpu().numpy(), ref))


@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("bot,top", [("SE", "E"), ("O", "SO"), ("SE", "SE")])
def test_stagg_bounded_512(op, bot, top):
    # the length of the reference's igrid z pencils; square cell operators run
    # the partitioned kernels (8 partitions of 64), the others the generic one
    nz, n1, n2 = 512, 16, 8
    dz = 1.0 / nz
    s = _make(nz, dz, bot, top)
    f = _field(n1, n2, _m(op, nz, pd.STAGG_EDGE_IN[op]), 5)
    ref = _ref(op, f, dz, bot, top)
    ft = po.as_fortran(torch.tensor(f, device=DEV))
    out = po.fortran_empty(ref.shape, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("op", sorted(OPS))
def test_stagg_bounded_512_many_pencils(op):
    # enough pencils (256 x 128) for the library to pick 64-point partitions,
    # as at 512^3; operators with an edge array run their system form
    nz, n1, n2 = 512, 256, 128
    dz = 1.0 / nz
    bot, top = "SO", "SE"
    s = _make(nz, dz, bot, top)
    f = _field(n1, n2, _m(op, nz, pd.STAGG_EDGE_IN[op]), 9)
    D = pd.stagg_bounded_operator(op, nz, dz, ENDS[bot], ENDS[top])
    ref = np.asfortranarray((f.reshape(n1 * n2, -1, order="F") @ D.T).reshape(n1, n2, -1, order="F"))
    ft = po.as_fortran(torch.tensor(f, device=DEV))
    out = po.fortran_empty(ref.shape, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-12
    # the accumulating form on the same launch shape: out = 2 op(f) - out
    prev = out.clone()
    import ctypes
    lib = _lib.load()
    code = {"d1_e2c": 11, "d1_c2e": 12, "d2_c2c": 13, "d2_e2e": 14, "interp_e2c": 15, "interp_c2e": 16,
            "d1_c2c": 17, "d1_e2e": 18}[op]
    _lib.check(lib.pade_op_apply_axpby(s.handle, code, 3, ctypes.c_double(2.0), ctypes.c_void_p(ft.data_ptr()),
                                       ctypes.c_double(-1.0), ctypes.c_void_p(out.data_ptr()), n1, n2, 0, 0,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), 2 * ref - prev.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("op", sorted(OPS))
def test_stagg_bounded_fp32_host_complex(op):
    nz, n1, n2 = 48, 10, 3
    dz = 1.0 / nz
    bot, top = "SO", "E"
    s = _make(nz, dz, bot, top)
    mi = _m(op, nz, pd.STAGG_EDGE_IN[op])
    f = _field(n1, n2, mi, 11)
    ref = _ref(op, f, dz, bot, top)
    # FP32 device arrays
    f32 = f.astype(np.float32)
    ref32 = _ref(op, f32.astype(np.float64), dz, bot, top)
    ft = po.as_fortran(torch.tensor(f32, device=DEV))
    out = po.fortran_empty(ref.shape, dtype=torch.float32, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    assert relerr(out.cpu().numpy().astype(np.float64), ref32) < 1e-4
    # host arrays
    host = np.zeros(ref.shape, order="F")
    getattr(s, OPS[op])(f, host, n1, n2)
    assert relerr(host, ref) < 1e-12
    # complex data, device and host
    fc = _field(n1, n2, mi, 12, np.complex128)
    refc = _ref(op, fc, dz, bot, top)
    fct = po.as_fortran(torch.tensor(fc, device=DEV))
    outc = po.fortran_empty(refc.shape, dtype=torch.complex128, device=DEV)
    getattr(s, OPS[op])(fct, outc, n1, n2)
    assert relerr(outc.cpu().numpy(), refc) < 1e-12
    hostc = np.zeros(refc.shape, dtype=np.complex128, order="F")
    getattr(s, OPS[op])(fc, hostc, n1, n2)
    assert relerr(hostc, refc) < 1e-12


def test_stagg_bounded_errors():
    lib = _lib.load()
    h = ctypes.c_void_p()
    # end codes: +-1 (even / odd) or +-2 (one-sided); 0 and 3 are rejected
    for bc1, bcn in [(0, 1), (1, 0), (3, 1), (1, -3)]:
        assert lib.pade_op_create(_lib.PADE_CD06STAGG, 16, 0.1, 0, bc1, bcn, ctypes.byref(h)) == 400
    assert lib.pade_op_create(_lib.PADE_CD06STAGG, 4, 0.1, 0, 1, 1, ctypes.byref(h)) == 21
    # ddz_C2C and ddz_E2E belong to the bounded form only
    s = po.cd06stagg()
    s.init(16, 0.1)
    f = torch.zeros((4, 4, 17), dtype=torch.float64, device=DEV)
    g = torch.zeros_like(f)
    for oper in (_lib.PADE_D1_C2C, _lib.PADE_D1_E2E):
        rc = lib.pade_op_apply(s.handle, oper, 3, ctypes.c_void_p(f.data_ptr()),
                               ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
        assert rc == 21
    b = _make(16, 0.1, "SE", "O")
    rc = lib.pade_op_apply(b.handle, _lib.PADE_D1_C2C, 1, ctypes.c_void_p(f.data_ptr()),
                           ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
    assert rc == 400
    rc = lib.pade_op_apply_adjoint(b.handle, _lib.PADE_D1_C2C, 3, ctypes.c_void_p(f.data_ptr()),
                                   ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
    assert rc == 21


def test_stagg_bounded_n5():
    # the shortest grid the reference accepts: closures from both ends meet
    nz, n1, n2 = 5, 3, 2
    dz = 0.2
    for op in sorted(OPS):
        for bot, top in [("SE", "O"), ("E", "SO"), ("SO", "SE")]:
            s = _make(nz, dz, bot, top)
            f = _field(n1, n2, _m(op, nz, pd.STAGG_EDGE_IN[op]), 3)
            ref = _ref(op, f, dz, bot, top)
            ft = po.as_fortran(torch.tensor(f, device=DEV))
            out = po.fortran_empty(ref.shape, device=DEV)
            getattr(s, OPS[op])(ft, out, n1, n2)
            torch.cuda.synchronize()
            assert relerr(out.cpu().numpy(), ref) < 1e-12, (op, bot, top)
