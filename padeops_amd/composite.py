"""Composite operators on whole fields: gradient, curl, divergence, filter3D
and the cross product, mirroring src/utilities/operators.F90:18-223.

The reference assembles them from the derivative / filter facades with
2DECOMP transposes between the y, x and z pencil orientations and separate
array additions.  On one GPU every orientation is the same Fortran-ordered
array, no transposes are needed, and each sum is accumulated by the
producing kernel itself (``df = alpha * op(f) + beta * df``), so e.g. a
divergence costs 64 bytes of HBM traffic per point (FP64) instead of 96.

``decomp`` is accepted for signature compatibility and may be None; the
arrays are the facade's (nx, ny, nz) fields (torch tensors on the GPU in
Fortran order).  Boundary arguments are pairs (bc1, bcn) as in the reference.
"""
from __future__ import annotations

from ._lib import PADE_D1, PadeError

_AX = {"x": 1, "y": 2, "z": 3}


def _d1(der, axis, f, out, bc, alpha=None, beta=None):
    op = der.operator(axis)
    sz = {"x": der.xsz, "y": der.ysz, "z": der.zsz}[axis]
    k = _AX[axis]
    na, nb = [s for i, s in enumerate(sz) if i != k - 1]
    method = {"x": der.getMethodx, "y": der.getMethody, "z": der.getMethodz}[axis]()
    # cd06 takes no apply-time bc (src/derivatives/cd06.F90:775-839); cd10
    # takes the pair (derivatives.F90:447-513)
    b1, bn = (None, None) if method == "cd06" else bc
    return op._apply(PADE_D1, k, f, out, na, nb, b1, bn, alpha=alpha, beta=beta)


def _bc(bc):
    return (0, 0) if bc is None else (int(bc[0]), int(bc[1]))


def _check_shape(der, *arrays):
    count = der.ysz[0] * der.ysz[1] * der.ysz[2]
    for a in arrays:
        if a.numel() != count:   # operators.F90:32-35 (GracefulExit 234)
            raise PadeError(234, "Either size of input array to the operator is inconsistent with decomp")


def gradient(decomp, der, f, dfdx, dfdy, dfdz, x_bc=None, y_bc=None, z_bc=None):
    """operators.F90:18-53"""
    _check_shape(der, f, dfdx, dfdy, dfdz)
    _d1(der, "y", f, dfdy, _bc(y_bc))
    _d1(der, "x", f, dfdx, _bc(x_bc))
    _d1(der, "z", f, dfdz, _bc(z_bc))


def curl(decomp, der, u, v, w, curlu, x_bc=None, y_bc=None, z_bc=None):
    """operators.F90:55-116; curlu is a list/tuple of three fields or a
    tensor whose last (Fortran: 4th) index selects the component."""
    c = [curlu[..., i] for i in range(3)] if hasattr(curlu, "dim") and curlu.dim() == 4 else list(curlu)
    _check_shape(der, u, v, w, *c)
    xb, yb, zb = _bc(x_bc), _bc(y_bc), _bc(z_bc)
    _d1(der, "y", w, c[0], yb)                          # dw/dy
    _d1(der, "z", v, c[0], zb, alpha=-1.0, beta=1.0)    # - dv/dz
    _d1(der, "z", u, c[1], zb)                          # du/dz
    _d1(der, "x", w, c[1], xb, alpha=-1.0, beta=1.0)    # - dw/dx
    _d1(der, "x", v, c[2], xb)                          # dv/dx
    _d1(der, "y", u, c[2], yb, alpha=-1.0, beta=1.0)    # - du/dy


def divergence(decomp, der, u, v, w, div, x_bc=None, y_bc=None, z_bc=None):
    """operators.F90:118-156 (the boundary codes enter negated, as there)."""
    _check_shape(der, u, v, w, div)
    xb, yb, zb = _bc(x_bc), _bc(y_bc), _bc(z_bc)
    _d1(der, "y", v, div, (-yb[0], -yb[1]))
    _d1(der, "x", u, div, (-xb[0], -xb[1]), alpha=1.0, beta=1.0)
    _d1(der, "z", w, div, (-zb[0], -zb[1]), alpha=1.0, beta=1.0)


def filter3D(decomp, fil, arr, numtimes=1, x_bc=None, y_bc=None, z_bc=None, work=None):
    """operators.F90:158-223: filter in y, then x, then z, each ``numtimes``
    times, in place on ``arr``.  ``work``: optional scratch field of the
    same shape (allocated when absent)."""
    from .operators import fortran_empty

    count = fil.ysz[0] * fil.ysz[1] * fil.ysz[2]
    if arr.numel() != count:
        raise PadeError(234, "Either size of input array to the operator is inconsistent with decomp")
    tmp = work if work is not None else fortran_empty(tuple(fil.ysz), dtype=arr.dtype, device=arr.device)
    src, dst = arr, tmp
    for axis, bc in (("y", _bc(y_bc)), ("x", _bc(x_bc)), ("z", _bc(z_bc))):
        for _ in range(max(1, int(numtimes))):
            getattr(fil, "filter" + axis)(src, dst, bc[0], bc[1])
            src, dst = dst, src
    if src.data_ptr() != arr.data_ptr():
        arr.copy_(src)


def crossprod(r, u, v):
    """operators.F90:225-243: r = u x v for 3-component fields given as
    sequences (or tensors with the component index last)."""
    comp = lambda a, i: a[..., i] if hasattr(a, "dim") and a.dim() == 4 else a[i]  # noqa: E731
    ux, uy, uz = (comp(u, i) for i in range(3))
    vx, vy, vz = (comp(v, i) for i in range(3))
    comp(r, 0).copy_(uy * vz - uz * vy)
    comp(r, 1).copy_(uz * vx - ux * vz)
    comp(r, 2).copy_(ux * vy - uy * vx)
