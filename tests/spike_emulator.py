"""NumPy emulation of the device kernels' partitioned solve, driven by the
tables the library exports (pade_plan_export).  Test helper: lets the CPU
suite check the planner (factors, coupling columns, reduced system, boundary
rows of B and B^T) without a GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from padeops_amd import _lib

KIND = {"cd06": 1, "cd10": 2, "cf90": 3, "gaussian": 5}
OPER = {"d1": 1, "d2": 2, "filter": 3}


def export(kind, which, n, dx, periodic, bc1, bcn, adjoint, M=0):
    lib = _lib.load()
    meta = (ctypes.c_int * 32)()
    _lib.check(lib.pade_plan_export(KIND[kind], OPER[which], n, dx, int(periodic), bc1, bcn, int(adjoint), M,
                                    None, 0, meta))
    tab = np.zeros(meta[12])
    _lib.check(lib.pade_plan_export(KIND[kind], OPER[which], n, dx, int(periodic), bc1, bcn, int(adjoint), M,
                                    tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), tab.size, meta))
    keys = ["n", "M", "P", "periodic", "nbl", "nbr", "off_bl", "off_br", "off_coef", "coef_len", "ncls", "kq",
            "len", "n_in"]
    m = dict(zip(keys, list(meta)[:14]))
    m["cls"] = list(meta)[16:32]
    return tab, m


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
    ca = tab[m["off_coef"]:m["off_coef"] + m["coef_len"]]
    ncls = m["ncls"]
    blocks = ca[:ncls * 10 * M].reshape(ncls, 10 * M)
    cls = [m["cls"][p] if not m["periodic"] else 0 for p in range(P)]
    lu = np.stack([blocks[cls[p], :6 * M].reshape(M, 6) for p in range(P)])
    S = np.stack([blocks[cls[p], 6 * M:].reshape(M, 4) for p in range(P)])
    qa = ca[ncls * 10 * M:]
    if m["periodic"]:
        Q0 = qa[:16 * P].reshape(4, 4 * P)
        # partition p's rows are partition 0's rotated by p partitions
        Q = np.stack([np.roll(Q0.reshape(4, P, 4), p, axis=1).reshape(4, 4 * P) for p in range(P)])
    else:
        Q = qa[:P * 16 * P].reshape(P, 4, 4 * P)
    kq = m["kq"]
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
            # only partitions within the window kq contribute (as on the device)
            mask = np.zeros(4 * P)
            for qq in range(P):
                d = abs(qq - p)
                if m["periodic"]:
                    d = min(d, P - d)
                if 2 * kq + 1 >= P or d <= kq:
                    mask[4 * qq:4 * qq + 4] = 1.0
            u = Q[p] @ (G * mask)
            g[p] -= S[p] @ u
    return g.reshape(-1)


def apply_pencil(kind, which, v, dx, periodic, bc1=0, bcn=0, adjoint=False, M=0):
    n = v.size
    tab, m = export(kind, which, n, dx, periodic, bc1, bcn, adjoint, M)
    pad = "wrap" if periodic else "clamp"
    if not adjoint or periodic:   # periodic adjoint: transposed tables, forward order
        return _solve(tab, m, _stencil(tab, m, v, pad))
    u = _solve(tab, m, v)
    return _stencil(tab, m, u, pad)


def _sys_right(row, v, n):
    """System form: column c of a right row multiplies input point n - c
    (the kernel's window clamps indices past the input's end)."""
    idx = np.clip(n - np.arange(16), 0, v.size - 1)
    return row @ v[idx]


def stagg_rhs(tab, m, v):
    """B of a bounded staggered operator (rows m["n"], input m["n_in"]) on v,
    as the kernels form it: explicit left / right rows, the interior stencil
    in between (every tap must fall inside v); system form when m["sys"]."""
    n, nin = m["n"], m["n_in"]
    assert v.size == nin
    st = tab[:9]
    bl = tab[m["off_bl"]:m["off_bl"] + 128].reshape(8, 16)
    br = tab[m["off_br"]:m["off_br"] + 128].reshape(8, 16)
    kk = min(16, nin)
    out = np.zeros(n)
    for r in range(n):
        if r < m["nbl"]:
            out[r] = bl[r, :kk] @ v[:kk]
        elif r >= n - m["nbr"]:
            if m.get("sys"):
                out[r] = _sys_right(br[n - 1 - r], v, n)
            else:
                out[r] = br[n - 1 - r, :kk] @ v[::-1][:kk]
        else:
            for k in range(9):
                if st[k] != 0.0:
                    j = r - 4 + k
                    assert 0 <= j < nin, (r, k)
                    out[r] += st[k] * v[j]
    return out


def apply_stagg(oper, v, dx, bc1, bcn, M=0):
    """A bounded staggered operator (cd06stagg kind, oper code) through the
    exported tables."""
    lib = _lib.load()
    meta = (ctypes.c_int * 32)()
    _lib.check(lib.pade_plan_export(4, oper, 0 + _stagg_n(oper, v.size), dx, 0, bc1, bcn, 0, M, None, 0, meta))
    tab = np.zeros(meta[12])
    _lib.check(lib.pade_plan_export(4, oper, _stagg_n(oper, v.size), dx, 0, bc1, bcn, 0, M,
                                    tab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), tab.size, meta))
    keys = ["n", "M", "P", "periodic", "nbl", "nbr", "off_bl", "off_br", "off_coef", "coef_len", "ncls", "kq",
            "len", "n_in", "sys", "tail"]
    m = dict(zip(keys, list(meta)[:16]))
    m["cls"] = list(meta)[16:32]
    y = _solve(tab, m, stagg_rhs(tab, m, v))
    if m["tail"]:
        # the row left out of the solve (cell-to-edge operators' last output)
        br = tab[m["off_br"]:m["off_br"] + 128].reshape(8, 16)
        r = _sys_right(br[7], v, m["n"])
        y = np.append(y, (r - tab[9] * y[-1]) * tab[10])
    return y, m


def _stagg_n(oper, nin):
    # cells n from the input length: edge inputs (E2C, E2E d2, E2C interp, E2E d1) hold n + 1
    return nin - 1 if oper in (11, 14, 15, 18) else nin
