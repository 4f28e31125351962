"""NumPy emulation of the device kernels' partitioned solve, driven by the
tables the library exports (pade_plan_export).  Test helper: lets the CPU
suite check the planner (factors, coupling columns, reduced system, boundary
rows of B and B^T) without a GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from padeops_amd import _lib

KIND = {"cd06": 1, "cd10": 2, "cf90": 3}
OPER = {"d1": 1, "d2": 2, "filter": 3}


def export(kind, which, n, dx, periodic, bc1, bcn, adjoint, M=0):
    lib = _lib.load()
    meta = (ctypes.c_int * 12)()
    _lib.check(lib.pade_plan_export(KIND[kind], OPER[which], n, dx, int(periodic), bc1, bcn, int(adjoint), M,
                                    None, 0, meta))
    tab = np.zeros(meta[11])
    _lib.check(lib.pade_plan_export(KIND[kind], OPER[which], n, dx, int(periodic), bc1, bcn, int(adjoint), M,
                                    tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), tab.size, meta))
    keys = ["n", "M", "P", "periodic", "nbl", "nbr", "off_bl", "off_br", "off_lu", "off_S", "off_Q", "len"]
    return tab, dict(zip(keys, list(meta)))


def _stencil(tab, m, v, pad_mode):
    n = m["n"]
    st = tab[:9]
    if pad_mode == "wrap":
        ext = np.concatenate([v[-4:], v, v[:4]])
    else:
        ext = np.concatenate([np.full(4, v[0]), v, np.full(4, v[-1])])
    out = np.zeros(n)
    for k in range(9):
        out += st[k] * ext[k:k + n]
    if not m["periodic"]:
        bl = tab[m["off_bl"]:m["off_bl"] + 128].reshape(8, 16)
        br = tab[m["off_br"]:m["off_br"] + 128].reshape(8, 16)
        kk = min(16, n)
        for r in range(m["nbl"]):
            out[r] = bl[r, :kk] @ v[:kk]
        for r in range(m["nbr"]):
            out[n - 1 - r] = br[r, :kk] @ v[::-1][:kk]
    return out


def _solve(tab, m, r):
    n, M, P = m["n"], m["M"], m["P"]
    lu = tab[m["off_lu"]:m["off_lu"] + P * M * 5].reshape(P, M, 5)
    S = tab[m["off_S"]:m["off_S"] + P * M * 4].reshape(P, M, 4)
    Q = tab[m["off_Q"]:m["off_Q"] + P * 4 * 4 * P].reshape(P, 4, 4 * P)
    g = r.reshape(P, M).copy()
    for p in range(P):
        w = g[p]
        for t in range(1, M):
            w[t] -= lu[p, t, 0] * w[t - 1] + (lu[p, t, 1] * w[t - 2] if t >= 2 else 0.0)
        w[M - 1] *= lu[p, M - 1, 2]
        w[M - 2] = w[M - 2] * lu[p, M - 2, 2] - lu[p, M - 2, 3] * w[M - 1]
        for t in range(M - 3, -1, -1):
            w[t] = w[t] * lu[p, t, 2] - lu[p, t, 3] * w[t + 1] - lu[p, t, 4] * w[t + 2]
    if m["periodic"] or P > 1:
        G = np.stack([g[:, 0], g[:, 1], g[:, M - 2], g[:, M - 1]], axis=1).reshape(-1)
        for p in range(P):
            u = Q[p] @ G
            g[p] -= S[p] @ u
    return g.reshape(-1)


def apply_pencil(kind, which, v, dx, periodic, bc1=0, bcn=0, adjoint=False, M=0):
    n = v.size
    tab, m = export(kind, which, n, dx, periodic, bc1, bcn, adjoint, M)
    pad = "wrap" if periodic else "clamp"
    if not adjoint:
        return _solve(tab, m, _stencil(tab, m, v, pad))
    u = _solve(tab, m, v)
    return _stencil(tab, m, u, pad)
