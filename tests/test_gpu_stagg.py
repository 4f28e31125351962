"""Periodic staggered z operators (cd06stagg.F90, periodic form) through the C
ABI against the oracle: every operation, FP64 to 1e-12, FP32 to 1e-4, the
repeated last edge plane, host arrays, and the error codes."""
import numpy as np
import pytest
import torch

from oracle import pade_dense as pd
import padeops_amd as po
from padeops_amd import _lib

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
OPS = {"d1_e2c": "ddz_E2C", "d1_c2e": "ddz_C2E", "d2_c2c": "d2dz2_C2C", "d2_e2e": "d2dz2_E2E",
       "interp_e2c": "InterpZ_E2C", "interp_c2e": "InterpZ_C2E"}


def _field(n1, n2, nz, seed, edges):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n1, n2, nz))
    if edges:
        f = np.concatenate([f, f[:, :, :1]], axis=2)
    return np.asfortranarray(f)


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("nz,n1,n2", [(64, 20, 7), (96, 40, 40), (200, 9, 5), (16, 64, 64)])
def test_stagg_fp64(op, nz, n1, n2):
    dz = 2 * np.pi / nz
    s = po.cd06stagg()
    s.init(nz, dz)
    f = _field(n1, n2, nz, nz + len(op), pd.STAGG_EDGE_IN[op])
    ref = pd.stagg_apply(op, f, dz)
    ft = po.as_fortran(torch.tensor(f, device=DEV))
    out = po.fortran_empty(ref.shape, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert relerr(got, ref) < 1e-12, (op, nz, relerr(got, ref))
    if pd.STAGG_EDGE_OUT[op]:
        assert np.array_equal(got[:, :, nz], got[:, :, 0])
    host = np.zeros(ref.shape, order="F")
    getattr(s, OPS[op])(f, host, n1, n2)
    assert relerr(host, ref) < 1e-12


@pytest.mark.parametrize("op", sorted(OPS))
def test_stagg_fp32(op):
    nz, n1, n2 = 128, 32, 8
    dz = 2 * np.pi / nz
    s = po.cd06stagg()
    s.init(nz, dz)
    f = _field(n1, n2, nz, 7, pd.STAGG_EDGE_IN[op])
    ref = pd.stagg_apply(op, f.astype(np.float32).astype(np.float64), dz)
    ft = po.as_fortran(torch.tensor(f, device=DEV, dtype=torch.float32))
    out = po.fortran_empty(ref.shape, dtype=torch.float32, device=DEV)
    getattr(s, OPS[op])(ft, out, n1, n2)
    assert relerr(out.cpu().numpy().astype(np.float64), ref) < 1e-4


def test_stagg_errors():
    s = po.cd06stagg()
    with pytest.raises(po.PadeError) as e:
        s.init(4, 0.1)
    assert e.value.code == 21
    with pytest.raises(po.PadeError) as e:
        s.init(4, 0.1, True, True)
    assert e.value.code == 21
    with pytest.raises(ValueError):
        s.init(16, 0.1, isTopEven=True)
    s.init(16, 0.1)
    f = torch.zeros((16, 4, 4), dtype=torch.float64, device=DEV)
    g = torch.zeros_like(f)
    import ctypes
    lib = _lib.load()
    # the staggered operators act along z only, and have no adjoint or dense form
    rc = lib.pade_op_apply(s.handle, _lib.PADE_D2_C2C, 1, ctypes.c_void_p(f.data_ptr()),
                           ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
    assert rc == 400
    rc = lib.pade_op_apply_adjoint(s.handle, _lib.PADE_D2_C2C, 3, ctypes.c_void_p(f.data_ptr()),
                                   ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
    assert rc == 21
    rc = lib.pade_op_apply(s.handle, _lib.PADE_D1, 3, ctypes.c_void_p(f.data_ptr()),
                           ctypes.c_void_p(g.data_ptr()), 4, 4, 0, 0, None)
    assert rc == 21
