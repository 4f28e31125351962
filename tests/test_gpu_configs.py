"""The five workload configurations end to end on the GPU, through the
drop-in facades (derivatives / filters) and the autograd entry point.

C1  64^3 CD06 periodic ddx/ddy/ddz of a 32-mode Fourier sum (seed 1): the
    discrete closed form (the reference's modified wavenumber,
    src/utilities/spectral.F90:987-1010) to 1e-12, and the analytic
    derivative within the modified-wavenumber error of the modes present.
C2  256^3 k^-5/3 field (|k| <= 85, seed 2): CD10 ddx/ddy/ddz, d2dx2/d2dy2/
    d2dz2 and cf90 filterx/y/z against the operators applied exactly through
    the FFT, to 1e-12.
C5  384 x 384 x 256, periodic CD06 in x, y and bounded CD06 in z (one-sided
    closures), cf90 each step, 20 Shu-Osher RK3 steps of linear
    advection-diffusion (velocity (1, 0.5, 0.25), nu = 1e-3, CFL 0.5) of a
    Gaussian blob; gradient of the final L2 norm w.r.t. the initial field
    with per-step checkpointing, checked through the linearity of the
    20-step map and by central finite differences.
(C3 is bench.py's workload, its parity is covered by test_gpu_parity.py;
C4 is test_gpu_autograd.py.)
"""
import math

import numpy as np
import pytest
import torch

import padeops_amd as po
from padeops_amd import autograd as pag
from padeops_amd import fields, spectral

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def relerr(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_c1_cd06_fourier_modes():
    n = 64
    dx = 2 * math.pi / n
    d = po.derivatives()
    d.init(n, n, n, dx, dx, dx, True, True, True, "cd06", "cd06", "cd06")
    f, (k, amp, phase) = fields.fourier_modes(n, modes=32, kmin=1, kmax=21, seed=1, device=DEV)
    x = torch.arange(n, dtype=torch.float64, device=DEV) * dx
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    for axis, fn in enumerate((d.ddx, d.ddy, d.ddz)):
        out = po.fortran_empty((n, n, n), device=DEV)
        fn(f, out)
        disc = torch.zeros_like(X)
        exact = torch.zeros_like(X)
        bound = 0.0
        for m in range(32):
            theta = int(k[m, 0]) * X + int(k[m, 1]) * Y + int(k[m, 2]) * Z + phase[m]
            km = float(k[m, axis])
            kmod = float(spectral.eigenvalue("cd06", "d1", torch.tensor(km, dtype=torch.float64), dx).imag)
            disc -= amp[m] * kmod * torch.sin(theta)
            exact -= amp[m] * km * torch.sin(theta)
            bound += abs(amp[m]) * abs(km - kmod)
        assert relerr(out, disc) < 1e-12
        assert float((out - exact).abs().max()) <= bound * (1 + 1e-9)


def test_c2_cd10_cf90_256():
    n = 256
    dx = 2 * math.pi / n
    f = fields.kolmogorov((n, n, n), kmax=85, seed=2, device=DEV)
    fc = f.permute(2, 1, 0)                       # C view: dims (z, y, x)
    d = po.derivatives()
    d.init(n, n, n, dx, dx, dx, True, True, True, "cd10", "cd10", "cd10")
    flt = po.filters()
    flt.init(n, n, n, True, True, True, "cf90", "cf90", "cf90")
    out = po.fortran_empty((n, n, n), device=DEV)
    cases = [(d.ddx, "cd10", "d1", 2), (d.ddy, "cd10", "d1", 1), (d.ddz, "cd10", "d1", 0),
             (d.d2dx2, "cd10", "d2", 2), (d.d2dy2, "cd10", "d2", 1), (d.d2dz2, "cd10", "d2", 0),
             (flt.filterx, "cf90", "filter", 2), (flt.filtery, "cf90", "filter", 1),
             (flt.filterz, "cf90", "filter", 0)]
    for fn, kind, which, dim in cases:
        fn(f, out)
        ref = spectral.apply_fourier(kind, which, fc, dim, dx)
        err = relerr(out.permute(2, 1, 0), ref)
        assert err < 1e-12, (kind, which, dim, err)


class AdvDiff:
    """Linear advection-diffusion of a scalar with compact operators."""

    def __init__(self, shape, steps=20):
        nx, ny, nz = shape
        self.shape = shape
        self.dx = 2 * math.pi / nx
        self.dy = 2 * math.pi / ny
        self.lz = 2 * math.pi
        self.dz = self.lz / (nz - 1)
        self.c = (1.0, 0.5, 0.25)
        self.nu = 1e-3
        self.dt = 0.5 * min(self.dx / self.c[0], self.dy / self.c[1], self.dz / self.c[2])
        self.steps = steps
        self.der = po.derivatives()
        self.der.init(nx, ny, nz, self.dx, self.dy, self.dz, True, True, False, "cd06", "cd06", "cd06")
        self.d2 = [po.cd06(), po.cd06(), po.cd06()]
        assert self.d2[0].init(nx, self.dx, True) == 0
        assert self.d2[1].init(ny, self.dy, True) == 0
        assert self.d2[2].init(nz, self.dz, False, 0, 0) == 0
        self.flt = [po.cf90(), po.cf90(), po.cf90()]
        assert self.flt[0].init(nx, True) == 0
        assert self.flt[1].init(ny, True) == 0
        assert self.flt[2].init(nz, False) == 0

    def initial(self):
        nx, ny, nz = self.shape
        x = torch.arange(nx, dtype=torch.float64, device=DEV) * self.dx
        y = torch.arange(ny, dtype=torch.float64, device=DEV) * self.dy
        z = torch.arange(nz, dtype=torch.float64, device=DEV) * self.dz
        X, Y, Z = torch.meshgrid(x, y, z, indexing="ij")
        r2 = (X - math.pi) ** 2 + (Y - math.pi) ** 2 + (Z - 0.5 * self.lz) ** 2
        return po.as_fortran(torch.exp(-r2 / (2 * 0.4 ** 2)))

    def rhs(self, u):
        D = self.der
        adv = (self.c[0] * pag.ddx(D.operator("x"), u) + self.c[1] * pag.ddy(D.operator("y"), u)
               + self.c[2] * pag.ddz(D.operator("z"), u))
        lap = pag.d2dx2(self.d2[0], u) + pag.d2dy2(self.d2[1], u) + pag.d2dz2(self.d2[2], u)
        return -adv + self.nu * lap

    def step(self, u):
        dt = self.dt
        u1 = u + dt * self.rhs(u)
        u2 = 0.75 * u + 0.25 * u1 + 0.25 * dt * self.rhs(u1)
        u3 = u / 3.0 + (2.0 / 3.0) * u2 + (2.0 / 3.0) * dt * self.rhs(u2)
        return pag.filterz(self.flt[2], pag.filtery(self.flt[1], pag.filterx(self.flt[0], u3)))

    def run(self, u0, checkpoint=True):
        if checkpoint:
            return pag.checkpointed_steps(self.step, u0, self.steps)
        u = u0
        for _ in range(self.steps):
            u = self.step(u)
        return u


def test_c5_checkpointed_adjoint():
    m = AdvDiff((384, 384, 256), steps=20)
    u0 = m.initial().requires_grad_(True)
    uT = m.run(u0)
    L = torch.linalg.vector_norm(uT)
    (g,) = torch.autograd.grad(L, u0)
    assert torch.isfinite(g).all()
    gen = torch.Generator(device=DEV)
    gen.manual_seed(5)
    uTd = uT.detach()
    with torch.no_grad():
        # the 20-step map M is linear: <grad L, d> = <M u0, M d> / L
        for _ in range(3):
            dvec = torch.randn(u0.shape, generator=gen, dtype=torch.float64, device=DEV)
            Md = m.run(po.as_fortran(dvec), checkpoint=False)
            lhs = float((g * dvec).sum())
            rhs = float((uTd * Md).sum() / L.detach())
            scale = float((g * dvec).abs().sum())
            assert abs(lhs - rhs) <= 1e-9 * abs(rhs) + 1e-12 * scale, (lhs, rhs, scale)
        # central finite difference along a modulated-blob direction
        eps = 1e-6
        noise = torch.rand(u0.shape, generator=gen, dtype=torch.float64, device=DEV)
        dvec = po.as_fortran(u0.detach() * (1.0 + 0.5 * noise))
        lp = torch.linalg.vector_norm(m.run(u0.detach() + eps * dvec, checkpoint=False))
        lm = torch.linalg.vector_norm(m.run(u0.detach() - eps * dvec, checkpoint=False))
        fd = float((lp - lm) / (2 * eps))
        an = float((g * dvec).sum())
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_c5_checkpoint_equals_plain_small():
    m = AdvDiff((96, 96, 64), steps=4)
    grads = []
    for ck in (True, False):
        u0 = m.initial().requires_grad_(True)
        (g,) = torch.autograd.grad(torch.linalg.vector_norm(m.run(u0, checkpoint=ck)), u0)
        grads.append(g)
    assert torch.equal(grads[0], grads[1])


def test_facade_errors_and_bc_dispatch():
    n = 32
    d = po.derivatives()
    d.init(n, n, n, 0.1, 0.1, 0.1, False, True, False, "cd10", "cd06", "cd06")
    f = po.as_fortran(torch.randn((n, n, n), dtype=torch.float64, device=DEV))
    a = po.fortran_empty((n, n, n), device=DEV)
    b = po.fortran_empty((n, n, n), device=DEV)
    with pytest.raises(po.PadeError) as e:
        d.d2dy2(f, a)
    assert e.value.code == 21
    # cd10: bc used only when both are given (derivatives.F90:455-458)
    d.ddx(f, a, 1, -1)
    d.operator("x").dd1(f, b, n, n, 1, -1)
    assert torch.equal(a, b)
    d.ddx(f, a, 1)
    d.operator("x").dd1(f, b, n, n)
    assert torch.equal(a, b)
    # cd06 ignores the bc arguments
    d.ddz(f, a, 1, 1)
    d.operator("z").dd3(f, b, n, n)
    assert torch.equal(a, b)
    flt = po.filters()
    flt.init(n, n, n, False, False, False, "cf90", "cf90", "cf90")
    flt.filtery(f, a, 0, 1)
    flt.operator("y").filter2(f, b, n, n, 0, 1)
    assert torch.equal(a, b)
    with pytest.raises(po.PadeError) as e:
        po.derivatives().init(4, 4, 4, 0.1, 0.1, 0.1, True, True, True, "cd10", "cd10", "cd10")
    assert e.value.code == 11
    with pytest.raises(po.PadeError) as e:
        po.filters().init(4, 4, 4, True, True, True, "cf90", "cf90", "cf90")
    assert e.value.code == 51
    flt.destroy()
    flt.init(n, n, n, True, True, True, "cf90", "cf90", "cf90")   # re-init after destroy works
