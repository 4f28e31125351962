"""Host-side mirror of the reference's ``derivatives`` and ``filters`` facades.

``derivatives`` follows ``src/derivatives/derivatives.F90:14-569`` and
``filters`` follows ``src/filters/filters.F90:13-322``: per-axis method
selection at ``init``, then ``ddx/ddy/ddz``, ``d2dx2/d2dy2/d2dz2`` and
``filterx/filtery/filterz`` on whole 3-D arrays.  On one GPU the local
pencil sizes of the three orientations are the global sizes
(``xsz = ysz = zsz = [nx, ny, nz]``); a ``decomp_info``-like record with
``xsz``/``ysz``/``zsz`` attributes is accepted in place of the sizes, as the
reference's parallel ``init`` accepts ``gp``.

Where the reference calls ``GracefulExit(message, code)`` (which ends the
MPI job) these classes raise :class:`PadeError` with the same code:
11 / 12 cd10 / cd06 init failed, 21 cd06 second derivative, 4 / 5 / 6 metric
or curvilinear terms requested, 1 invalid derivative method, 51 cf90 init
failed, 52 invalid filter method.  The spectral ("four", "cheb") derivative
methods and the least-squares filter are outside this library and raise
13 / 14 / 52; the Gaussian filter is provided.
"""
from __future__ import annotations

from ._lib import PadeError
from .operators import cd06, cd10, cf90, gaussian

_AXES = ("x", "y", "z")


def _sizes(args):
    """(nx, ny, nz, xsz, ysz, zsz, rest) from either (nx, ny, nz, ...) or (gp, ...)."""
    first = args[0]
    if hasattr(first, "xsz"):
        xsz, ysz, zsz = (tuple(int(v) for v in getattr(first, a)) for a in ("xsz", "ysz", "zsz"))
        return xsz[0], ysz[1], zsz[2], xsz, ysz, zsz, args[1:]
    nx, ny, nz = (int(v) for v in args[:3])
    s = (nx, ny, nz)
    return nx, ny, nz, s, s, s, args[3:]


def _check_shape(arr, sz):
    count = sz[0] * sz[1] * sz[2]
    size = arr.size if hasattr(arr, "size") and not callable(arr.size) else arr.numel()
    if size != count:
        raise ValueError(f"array of {size} values where the pencil layout {list(sz)} has {count}")


class derivatives:
    """Per-axis compact derivatives (derivatives.F90:14-569)."""

    def __init__(self):
        self._ops = {}
        self._method = {}
        self.xsz = self.ysz = self.zsz = None
        self.initialized = False

    def init(self, *args, xmetric=None, ymetric=None, zmetric=None, curvilinear=None):
        """``init(nx, ny, nz, dx, dy, dz, periodic_x, periodic_y, periodic_z,
        method_x, method_y, method_z [, xmetric, ymetric, zmetric, curvilinear])``
        (derivatives.F90:140-186) or ``init(gp, dx, dy, dz, ...)``
        (derivatives.F90:189-232)."""
        nx, ny, nz, xsz, ysz, zsz, rest = _sizes(args)
        rest = list(rest)
        if len(rest) > 9:
            xmetric, ymetric, zmetric, *tail = rest[9:]
            curvilinear = tail[0] if tail else curvilinear
            rest = rest[:9]
        if len(rest) != 9:
            raise TypeError("init takes sizes, (dx, dy, dz), (periodic_x, periodic_y, periodic_z) and "
                            "(method_x, method_y, method_z)")
        dxs, pers, methods = rest[0:3], rest[3:6], rest[6:9]
        for code, flag, axis in ((4, xmetric, "x"), (5, ymetric, "y"), (6, zmetric, "z")):
            if flag:   # derivatives.F90:239-257
                raise PadeError(code, f"Code INCOMPLETE! Metric terms are not supported in {axis} direction!")
        if curvilinear:  # derivatives.F90:259-272
            raise PadeError(6, "Code INCOMPLETE! Curvilinear Formulation is not supported!")
        for axis, m in zip(_AXES, methods):
            if m in ("four", "cheb"):
                raise PadeError(13 if m == "four" else 14,
                                f"{m} collocation in {axis} is not provided by this compact-operator library")
            if m not in ("cd10", "cd06"):
                raise PadeError(1, f"Invalid method selected in {axis} direction")
        if self.initialized:
            self.destroy()
        self.xsz, self.ysz, self.zsz = xsz, ysz, zsz
        for axis, n, dx, per, m in zip(_AXES, (nx, ny, nz), dxs, pers, methods):
            op = cd10() if m == "cd10" else cd06()
            if op.init(n, float(dx), bool(per), 0, 0) != 0:   # derivatives.F90:276-389
                raise PadeError(11 if m == "cd10" else 12, f"Initializing {m} failed in {axis.upper()}")
            self._ops[axis] = op
            self._method[axis] = m
        self.initialized = True

    def destroy(self):
        for op in self._ops.values():
            op.destroy()
        self._ops.clear()
        self._method.clear()
        self.initialized = False

    def getMethodx(self):
        return self._method.get("x")

    def getMethody(self):
        return self._method.get("y")

    def getMethodz(self):
        return self._method.get("z")

    def _d1(self, axis, f, df, bc1, bcn):
        sz = {"x": self.xsz, "y": self.ysz, "z": self.zsz}[axis]
        _check_shape(f, sz)
        _check_shape(df, sz)
        k = _AXES.index(axis)
        na, nb = [s for i, s in enumerate(sz) if i != k]
        op = self._ops[axis]
        fn = getattr(op, f"dd{k + 1}")
        if self._method[axis] == "cd10" and bc1 is not None and bcn is not None:
            return fn(f, df, na, nb, bc1, bcn)          # derivatives.F90:455-458
        return fn(f, df, na, nb)

    def _d2(self, axis, f, df, bc1, bcn):
        if self._method[axis] == "cd06":                 # derivatives.F90:524-525
            raise PadeError(21, "CD06 is incomplete right now")
        sz = {"x": self.xsz, "y": self.ysz, "z": self.zsz}[axis]
        _check_shape(f, sz)
        _check_shape(df, sz)
        k = _AXES.index(axis)
        na, nb = [s for i, s in enumerate(sz) if i != k]
        return getattr(self._ops[axis], f"d2d{k + 1}")(f, df, na, nb, bc1, bcn)

    def ddx(self, f, dfdx, bc1=None, bcn=None):
        """derivatives.F90:447-468"""
        return self._d1("x", f, dfdx, bc1, bcn)

    def ddy(self, f, dfdy, bc1=None, bcn=None):
        """derivatives.F90:470-491"""
        return self._d1("y", f, dfdy, bc1, bcn)

    def ddz(self, f, dfdz, bc1=None, bcn=None):
        """derivatives.F90:493-513"""
        return self._d1("z", f, dfdz, bc1, bcn)

    def d2dx2(self, f, d2fdx2, bc1=None, bcn=None):
        """derivatives.F90:515-532"""
        return self._d2("x", f, d2fdx2, bc1, bcn)

    def d2dy2(self, f, d2fdy2, bc1=None, bcn=None):
        """derivatives.F90:534-551"""
        return self._d2("y", f, d2fdy2, bc1, bcn)

    def d2dz2(self, f, d2fdz2, bc1=None, bcn=None):
        """derivatives.F90:553-569"""
        return self._d2("z", f, d2fdz2, bc1, bcn)

    def operator(self, axis):
        """The underlying cd06 / cd10 operator of one axis ("x", "y" or "z")."""
        return self._ops[axis]


class filters:
    """Per-axis compact filters (filters.F90:13-322)."""

    def __init__(self):
        self._ops = {}
        self._method = {}
        self.xsz = self.ysz = self.zsz = None
        self.initialized = False

    def init(self, *args):
        """``init(nx, ny, nz, periodicx, periodicy, periodicz, methodx, methody,
        methodz)`` (filters.F90:299-322) or ``init(gp, periodicx, ...)``
        (filters.F90:274-297)."""
        nx, ny, nz, xsz, ysz, zsz, rest = _sizes(args)
        if len(rest) != 6:
            raise TypeError("init takes sizes, (periodicx, periodicy, periodicz) and (methodx, methody, methodz)")
        pers, methods = rest[0:3], rest[3:6]
        for axis, m, per in zip(_AXES, methods, pers):
            if m == "lstsq":
                raise PadeError(52, f"the {m} filter in {axis} is not provided by this compact-operator library")
            if m not in ("cf90", "gaussian"):
                raise PadeError(52, f"Incorrect method select in direction {axis.upper()}")
        if self.initialized:
            self.destroy()
        self.xsz, self.ysz, self.zsz = xsz, ysz, zsz
        for axis, n, per, m in zip(_AXES, (nx, ny, nz), pers, methods):
            op = cf90() if m == "cf90" else gaussian()
            if op.init(n, bool(per)) != 0:          # filters.F90:82-184
                raise PadeError(51, f"Initializing {m} {'failed' if m == 'cf90' else 'filter failed'} in "
                                    f"{axis.upper()} ")
            self._ops[axis] = op
            self._method[axis] = m
        self.initialized = True

    def destroy(self):
        # The reference's destroy leaves the components allocated
        # (filters.F90:186-217), so its re-init after destroy fails; here the
        # handles are released and re-init works.
        for op in self._ops.values():
            op.destroy()
        self._ops.clear()
        self._method.clear()
        self.initialized = False

    def getMethodx(self):
        return self._method.get("x")

    def getMethody(self):
        return self._method.get("y")

    def getMethodz(self):
        return self._method.get("z")

    def _filter(self, axis, f, ff, bc1, bcn):
        sz = {"x": self.xsz, "y": self.ysz, "z": self.zsz}[axis]
        _check_shape(f, sz)
        _check_shape(ff, sz)
        k = _AXES.index(axis)
        na, nb = [s for i, s in enumerate(sz) if i != k]
        return getattr(self._ops[axis], f"filter{k + 1}")(f, ff, na, nb, bc1, bcn)

    def filterx(self, f, ff, bc1=None, bcn=None):
        """filters.F90:220-235"""
        return self._filter("x", f, ff, bc1, bcn)

    def filtery(self, f, ff, bc1=None, bcn=None):
        """filters.F90:237-252"""
        return self._filter("y", f, ff, bc1, bcn)

    def filterz(self, f, ff, bc1=None, bcn=None):
        """filters.F90:254-269"""
        return self._filter("z", f, ff, bc1, bcn)

    def operator(self, axis):
        """The underlying cf90 / gaussian operator of one axis ("x", "y" or "z")."""
        return self._ops[axis]
