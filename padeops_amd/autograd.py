"""Differentiable entry point: compact operators as ``torch.autograd.Function``.

Forward runs ``pade_op_apply`` (``out = A^{-1} B f``); backward runs the
library's adjoint kernel ``pade_op_apply_adjoint`` (``g_f = B^T A^{-T} g``,
transposed boundary closures included).  Tensors are logical 3-D arrays
``f[i, j, k]`` of shape ``(n1, n2, n3)``; inputs that are not laid out in
Fortran order are copied into it, and outputs come back Fortran-ordered.

The operators are those of the reference's modules (cd06 ``dd*``, cd10
``dd*``/``d2d*``, cf90 ``filter*``; src/derivatives/cd06.F90:775-839,
src/derivatives/cd10.F90:2029-2447, src/filters/cf90.F90:1020-1228); the
adjoint has no reference counterpart.
"""
from __future__ import annotations

import torch
import torch.utils.checkpoint

from ._lib import PADE_D1, PADE_D2, PADE_FILTER
from .operators import fortran_empty

__all__ = ["compact_apply", "ddx", "ddy", "ddz", "d2dx2", "d2dy2", "d2dz2", "filterx", "filtery", "filterz",
           "checkpointed_steps"]


def _fortran(t):
    if t.dim() == 3 and t.permute(2, 1, 0).is_contiguous():
        return t
    out = fortran_empty(tuple(t.shape), dtype=t.dtype, device=t.device)
    out.copy_(t)
    return out


def _sizes(shape, axis):
    n = shape[axis - 1]
    na, nb = [s for i, s in enumerate(shape) if i != axis - 1]
    return n, na, nb


class _CompactApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, op, oper, axis, bc1, bcn):
        if f.dim() != 3:
            raise ValueError("expected a 3-D tensor (n1, n2, n3)")
        n, na, nb = _sizes(tuple(f.shape), axis)
        if n != op.GetSize():
            raise ValueError(f"axis {axis} has {n} points, the operator {op.GetSize()}")
        fin = _fortran(f.detach())
        out = fortran_empty(tuple(f.shape), dtype=f.dtype, device=f.device)
        op._apply(oper, axis, fin, out, na, nb, bc1, bcn)
        ctx.op, ctx.oper, ctx.axis, ctx.bc = op, oper, axis, (bc1, bcn)
        return out

    @staticmethod
    def backward(ctx, g):
        if g is None:
            return (None,) * 6
        n, na, nb = _sizes(tuple(g.shape), ctx.axis)
        gin = _fortran(g.detach())
        out = fortran_empty(tuple(g.shape), dtype=g.dtype, device=g.device)
        ctx.op._apply(ctx.oper, ctx.axis, gin, out, na, nb, ctx.bc[0], ctx.bc[1], adjoint=True)
        return out, None, None, None, None, None


def compact_apply(op, oper, axis, f, bc1=None, bcn=None):
    """Differentiable ``op`` (cd06 / cd10 / cf90 instance) of kind ``oper``
    (PADE_D1, PADE_D2, PADE_FILTER) along ``axis`` (1, 2, 3) of the 3-D tensor f."""
    return _CompactApply.apply(f, op, oper, axis, bc1, bcn)


def ddx(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D1, 1, f, bc1, bcn)


def ddy(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D1, 2, f, bc1, bcn)


def ddz(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D1, 3, f, bc1, bcn)


def d2dx2(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D2, 1, f, bc1, bcn)


def d2dy2(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D2, 2, f, bc1, bcn)


def d2dz2(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_D2, 3, f, bc1, bcn)


def filterx(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_FILTER, 1, f, bc1, bcn)


def filtery(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_FILTER, 2, f, bc1, bcn)


def filterz(op, f, bc1=None, bcn=None):
    return compact_apply(op, PADE_FILTER, 3, f, bc1, bcn)


def checkpointed_steps(step, u, nsteps):
    """Apply ``u = step(u)`` ``nsteps`` times keeping only each step's input
    for the backward pass (per-step checkpointing): the backward recomputes
    one step's intermediates at a time, so memory is O(nsteps) fields instead
    of O(nsteps * stages)."""
    for _ in range(nsteps):
        u = torch.utils.checkpoint.checkpoint(step, u, use_reentrant=False)
    return u
