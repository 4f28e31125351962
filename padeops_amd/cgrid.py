"""The compressible solver's Runge-Kutta substep on GPU-resident fields:
src/compressible/cgrid.F90 (advance_RK45 :663-717, getRHS :822-1056, LAD
:1058-1257, filter :1259-1336, get_tau / get_q :1338-1457) for one ideal-gas
species with the power-law viscosity model (IdealGasEOS.F90,
PowerLawViscosity.F90, MixtureEOS.F90 ns = 1) on a periodic grid.

Every derivative and filter runs through libpadeops_hip (the derivatives and
filters facades: cd10 / cd06 first and second derivatives, cf90 and the
periodic Gaussian filter); the flux divergences accumulate into the
right-hand side inside the derivative kernels (rhs = rhs - d(flux)/dx in one
pass).  Pointwise algebra is PyTorch on the same FP64 arrays.  On one GPU the
reference's y / x / z pencil transposes are identities: every field is one
Fortran-ordered (nx, ny, nz) array.

Conserved variables, as the reference orders them for one species:
W[0] = rho, W[1:4] = rho (u, v, w), W[4] = total energy rho (e + |u|^2 / 2).
"""
from __future__ import annotations

import math

import torch

from . import composite
from .facades import derivatives, filters
from .operators import as_fortran, fortran_empty

# low-storage RK45 (src/utilities/RKCoeffs.F90:6-17)
RK45_A = (0.0, -6234157559845.0 / 12983515589748.0, -6194124222391.0 / 4410992767914.0,
          -31623096876824.0 / 15682348800105.0, -12251185447671.0 / 11596622555746.0)
RK45_B = (494393426753.0 / 4806282396855.0, 4047970641027.0 / 5463924506627.0,
          9795748752853.0 / 13190207949281.0, 4009051133189.0 / 8539092990294.0,
          1348533437543.0 / 7166442652324.0)

_EPS = 1.0e-32


def _fo(t):
    """t itself when it already has Fortran strides, else a Fortran copy (the
    library reads (nx, ny, nz) arrays in Fortran order)."""
    return t if t.permute(2, 1, 0).is_contiguous() else as_fortran(t)


class cgrid:
    """One-species, periodic compressible grid.  ``init`` takes the sizes,
    spacings and the &CINPUT parameters (cgrid.F90:150-199: gam, Rgas, Cmu,
    Cbeta, Ckap, inviscid; the power-law viscosity mu_ref, T_ref, n, Pr,
    PowerLawViscosity.F90:8-15).  Fields are (nx, ny, nz) FP64 tensors in
    Fortran order on ``device``."""

    def __init__(self, nx, ny, nz, dx, dy, dz, derivative="cd10", filt="cf90", gam=1.4, Rgas=1.0,
                 Cmu=0.002, Cbeta=1.75, Ckap=0.01, inviscid=True, mu_ref=1.0, T_ref=1.0, visc_n=0.75,
                 Pr=0.75, device="cuda:0"):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = dx, dy, dz
        self.gam, self.Rgas = gam, Rgas
        self.Cmu, self.Cbeta, self.Ckap = Cmu, Cbeta, Ckap
        self.inviscid = inviscid
        self.mu_ref, self.T_ref, self.visc_n, self.Pr = mu_ref, T_ref, visc_n, Pr
        self.device = device
        self.der = derivatives()
        self.der.init(nx, ny, nz, dx, dy, dz, True, True, True, derivative, derivative, derivative)
        self.fil = filters()
        self.fil.init(nx, ny, nz, True, True, True, filt, filt, filt)
        self.gfil = filters()                                   # cgrid.F90:338-340
        self.gfil.init(nx, ny, nz, True, True, True, "gaussian", "gaussian", "gaussian")
        self.x_bc = self.y_bc = self.z_bc = (0, 0)
        e = self._empty
        self.rho, self.u, self.v, self.w = e(), e(), e(), e()
        self.p, self.T, self.e = e(), e(), e()
        self.mu, self.bulk, self.kap = e(), e(), e()
        self.W = fortran_empty((nx, ny, nz, 5), device=device)
        self.tsim = 0.0
        self.step = 0
        self.dt = 0.0

    def _empty(self):
        return fortran_empty((self.nx, self.ny, self.nz), device=self.device)

    def destroy(self):
        self.der.destroy()
        self.fil.destroy()
        self.gfil.destroy()

    # ------------------------------------------------------------ EOS
    def get_e_from_p(self):
        """IdealGasEOS.F90:54-61."""
        torch.div(self.p / (self.gam - 1.0), self.rho, out=self.e)

    def get_T(self):
        """IdealGasEOS.F90:63-70."""
        torch.mul(self.e, (self.gam - 1.0) / self.Rgas, out=self.T)

    def get_p(self):
        """IdealGasEOS.F90:45-52."""
        torch.mul(self.rho * self.e, self.gam - 1.0, out=self.p)

    def get_sos(self):
        """IdealGasEOS.F90:72-79."""
        return torch.sqrt(self.gam * self.p / self.rho)

    def post_bc(self):
        """cgrid.F90:813-820."""
        self.get_e_from_p()
        self.get_T()

    def get_primitive(self):
        """cgrid.F90:764-797 (one species)."""
        W = self.W
        self.rho.copy_(W[:, :, :, 0])
        onebyrho = 1.0 / self.rho
        torch.mul(W[:, :, :, 1], onebyrho, out=self.u)
        torch.mul(W[:, :, :, 2], onebyrho, out=self.v)
        torch.mul(W[:, :, :, 3], onebyrho, out=self.w)
        self.e.copy_(W[:, :, :, 4] * onebyrho - 0.5 * (self.u * self.u + self.v * self.v + self.w * self.w))
        self.get_p()
        self.get_T()

    def get_conserved(self):
        """cgrid.F90:799-811 (one species)."""
        W = self.W
        W[:, :, :, 0].copy_(self.rho)
        W[:, :, :, 1].copy_(self.rho * self.u)
        W[:, :, :, 2].copy_(self.rho * self.v)
        W[:, :, :, 3].copy_(self.rho * self.w)
        W[:, :, :, 4].copy_(self.rho * (self.e + 0.5 * (self.u * self.u + self.v * self.v + self.w * self.w)))

    def get_transport_properties(self):
        """MixtureEOS.F90:175-229 (ns = 1) with PowerLawViscosity.F90:38-75."""
        if self.inviscid:
            self.mu.zero_()
            self.bulk.zero_()
            self.kap.zero_()
            return
        Cp = self.gam * self.Rgas / (self.gam - 1.0)
        torch.mul(torch.pow(self.T / self.T_ref, self.visc_n), self.mu_ref, out=self.mu)
        self.bulk.zero_()
        torch.mul(self.mu, Cp / self.Pr, out=self.kap)

    # ------------------------------------------------------ operators
    def gradient(self, f, dfdx, dfdy, dfdz):
        """cgrid.F90:446-478."""
        composite.gradient(None, self.der, f, dfdx, dfdy, dfdz, self.x_bc, self.y_bc, self.z_bc)

    def _d4(self, axis, f, t1, t2):
        """Fourth derivative along one axis as the second derivative applied
        twice (cgrid.F90:1089-1105): result in t2."""
        d2 = {"x": self.der.d2dx2, "y": self.der.d2dy2, "z": self.der.d2dz2}[axis]
        d2(f, t1)
        d2(t1, t2)
        return t2

    def filter(self, arr, fil, numtimes=1):
        """cgrid.F90:1259-1336: y, then x, then z, each ``numtimes`` times;
        the result replaces ``arr``."""
        t1, t2 = self._empty(), self._empty()
        for f1 in (fil.filtery, fil.filterx, fil.filterz):
            f1(arr, t1)
            for _ in range(numtimes - 1):
                f1(t1, t2)
                t1, t2 = t2, t1
            arr.copy_(t1)

    # ------------------------------------------------------------- LAD
    def getLAD(self, du):
        """cgrid.F90:1058-1257 (no species diffusivity for one species)."""
        dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz = du
        t1, t2 = self._empty(), self._empty()
        dx6, dy6, dz6 = self.dx ** 6, self.dy ** 6, self.dz ** 6
        # artificial shear viscosity (:1083-1113)
        func = _fo(torch.sqrt(dudx ** 2 + 0.5 * (dvdx + dudy) ** 2 + 0.5 * (dwdx + dudz) ** 2
                              + dvdy ** 2 + 0.5 * (dwdy + dvdz) ** 2 + dwdz ** 2))
        mustar = self._d4("x", func, t1, t2) * dx6
        mustar += self._d4("z", func, t1, t2) * dz6
        mustar += self._d4("y", func, t1, t2) * dy6
        mustar = _fo(self.Cmu * self.rho * torch.abs(mustar))
        self.filter(mustar, self.gfil, 2)
        # artificial bulk viscosity (:1115-1161)
        func = _fo(dudx + dvdy + dwdz)
        g1, g2, g3 = self._empty(), self._empty(), self._empty()
        self.gradient(self.rho, g1, g2, g3)
        g1.mul_(g1)
        g2.mul_(g2)
        g3.mul_(g3)
        den = g1 + g2 + g3 + _EPS
        bulkstar = self._d4("x", func, t1, t2) * self.dx ** 4 * (self.dx * g1 / den) ** 2
        bulkstar += self._d4("z", func, t1, t2) * self.dz ** 4 * (self.dz * g3 / den) ** 2
        bulkstar += self._d4("y", func, t1, t2) * self.dy ** 4 * (self.dy * g2 / den) ** 2
        curl2 = (dwdy - dvdz) ** 2 + (dudz - dwdx) ** 2 + (dvdx - dudy) ** 2
        dil2 = func * func
        fsw = dil2 / (dil2 + curl2 + _EPS)
        fsw = torch.where(func >= 0.0, torch.zeros_like(fsw), fsw)
        bulkstar = _fo(self.Cbeta * self.rho * fsw * torch.abs(bulkstar))
        self.filter(bulkstar, self.gfil, 2)
        # artificial conductivity (:1163-1199)
        self.gradient(self.e, g1, g2, g3)
        g1.mul_(g1)
        g2.mul_(g2)
        g3.mul_(g3)
        den = g1 + g2 + g3 + _EPS
        kapstar = self._d4("x", self.e, t1, t2) * self.dx ** 4 * (self.dx * g1 / den)
        kapstar += self._d4("z", self.e, t1, t2) * self.dz ** 4 * (self.dz * g3 / den)
        kapstar += self._d4("y", self.e, t1, t2) * self.dy ** 4 * (self.dy * g2 / den)
        kapstar = _fo(self.Ckap * self.rho * self.get_sos() * torch.abs(kapstar) / self.T)
        self.filter(kapstar, self.gfil, 2)
        # (:1252-1254)
        self.mu += mustar
        self.bulk += bulkstar
        self.kap += kapstar

    # ------------------------------------------------------------- RHS
    def getRHS(self, rhs):
        """cgrid.F90:822-896 with getRHS_x/y/z (:898-1056), get_tau
        (:1338-1390) and get_q (:1392-1444).  ``rhs`` is (nx, ny, nz, 5)."""
        du = [self._empty() for _ in range(9)]
        dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz = du
        self.gradient(self.u, dudx, dudy, dudz)
        self.gradient(self.v, dvdx, dvdy, dvdz)
        self.gradient(self.w, dwdx, dwdy, dwdz)
        self.get_transport_properties()
        self.getLAD(du)
        # stress tensor (get_tau)
        mu, bulk = self.mu, self.bulk
        bambda = (4.0 / 3.0) * mu + bulk
        lam = bulk - (2.0 / 3.0) * mu
        tauxy = mu * (dudy + dvdx)
        tauxz = mu * (dudz + dwdx)
        tauyz = mu * (dvdz + dwdy)
        tauxx = bambda * dudx + lam * (dvdy + dwdz)
        tauyy = bambda * dvdy + lam * (dudx + dwdz)
        tauzz = bambda * dwdz + lam * (dudx + dvdy)
        del du, dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz
        # heat flux (get_q): q = -kap grad T
        qx, qy, qz = self._empty(), self._empty(), self._empty()
        self.gradient(self.T, qx, qy, qz)
        qx.mul_(-self.kap)
        qy.mul_(-self.kap)
        qz.mul_(-self.kap)
        rhs.zero_()
        W, u, v, w, p = self.W, self.u, self.v, self.w, self.p
        rhou, rhov, rhow, TE = W[:, :, :, 1], W[:, :, :, 2], W[:, :, :, 3], W[:, :, :, 4]
        flux = self._empty()

        def acc(axis, k, val):
            # rhs[k] -= d(flux)/d(axis), accumulated inside the derivative kernel
            flux.copy_(val)
            bc = {"x": self.x_bc, "y": self.y_bc, "z": self.z_bc}[axis]
            composite._d1(self.der, axis, flux, rhs[:, :, :, k], bc, alpha=-1.0, beta=1.0)

        # x fluxes (getRHS_x)
        acc("x", 0, rhou)
        acc("x", 1, rhou * u + p - tauxx)
        acc("x", 2, rhou * v - tauxy)
        acc("x", 3, rhou * w - tauxz)
        acc("x", 4, (TE + p - tauxx) * u - v * tauxy - w * tauxz + qx)
        # y fluxes (getRHS_y)
        acc("y", 0, rhov)
        acc("y", 1, rhov * u - tauxy)
        acc("y", 2, rhov * v + p - tauyy)
        acc("y", 3, rhov * w - tauyz)
        acc("y", 4, (TE + p - tauyy) * v - u * tauxy - w * tauyz + qy)
        # z fluxes (getRHS_z)
        acc("z", 0, rhow)
        acc("z", 1, rhow * u - tauxz)
        acc("z", 2, rhow * v - tauyz)
        acc("z", 3, rhow * w + p - tauzz)
        acc("z", 4, (TE + p - tauzz) * w - u * tauxz - v * tauyz + qz)

    # ------------------------------------------------------------ step
    def advance_RK45(self, dt=None):
        """cgrid.F90:663-717: five low-storage substeps, each followed by the
        filter of the conserved variables and the primitive update."""
        if dt is not None:
            self.dt = dt
        if self.step == 0:
            self.get_e_from_p()
            self.get_T()
        rhs = fortran_empty((self.nx, self.ny, self.nz, 5), device=self.device)
        Qtmp = torch.zeros_like(rhs)
        Qtmpt = 0.0
        for isub in range(5):
            self.get_conserved()
            if not bool(torch.isfinite(self.W).all()):
                raise FloatingPointError(f"NaN encountered in solution at substep {isub + 1} of step {self.step + 1}")
            self.getRHS(rhs)
            Qtmp.mul_(RK45_A[isub]).add_(rhs, alpha=self.dt)
            Qtmpt = self.dt + RK45_A[isub] * Qtmpt
            self.W.add_(Qtmp, alpha=RK45_B[isub])
            self.tsim += RK45_B[isub] * Qtmpt
            for k in range(5):
                arr = self.W[:, :, :, k]
                tmp = self._empty()
                tmp.copy_(arr)
                self.filter(tmp, self.fil, 1)
                arr.copy_(tmp)
            self.get_primitive()
            self.post_bc()
        self.step += 1

    def get_dt(self, CFL):
        """Convective and diffusive limits of cgrid.F90:719-762 (the first
        ten steps take a tenth)."""
        cs = self.get_sos()
        dtCFL = CFL / float(torch.max(torch.abs(self.u) / self.dx + torch.abs(self.v) / self.dy
                                      + torch.abs(self.w) / self.dz
                                      + cs * math.sqrt(1 / self.dx ** 2 + 1 / self.dy ** 2 + 1 / self.dz ** 2)))
        h = min(self.dx, self.dy, self.dz)
        eps = 1e-15
        dtmu = 0.2 * h ** 2 / (float(torch.max(self.mu / self.rho)) + eps)
        dtbulk = 0.2 * h ** 2 / (float(torch.max(self.bulk / self.rho)) + eps)
        dtkap = 0.2 / (float(torch.max(self.kap * self.T / (self.rho * h ** 4))) ** (1.0 / 3.0) + eps)
        dt = dtCFL
        if dt > dtmu:
            dt = dtmu
        elif dt > dtbulk:
            dt = dtbulk
        elif dt > dtkap:
            dt = dtkap
        if self.step <= 10:
            dt = dt / 10.0
        self.dt = dt
        return dt


def taylorgreen(g):
    """Initial fields of problems/compressible/taylorgreen_files/hooks.F90
    (meshgen and initfields): x = (i - 1) dx on [0, 2 pi)."""
    dev = g.device
    x = torch.arange(g.nx, dtype=torch.float64, device=dev) * g.dx
    y = torch.arange(g.ny, dtype=torch.float64, device=dev) * g.dy
    z = torch.arange(g.nz, dtype=torch.float64, device=dev) * g.dz
    X, Y, Z = torch.meshgrid(x, y, z, indexing="ij")
    g.rho.fill_(1.0)
    g.u.copy_(torch.sin(X) * torch.cos(Y) * torch.cos(Z))
    g.v.copy_(-torch.cos(X) * torch.sin(Y) * torch.cos(Z))
    g.w.zero_()
    g.p.copy_(100.0 + ((torch.cos(2 * Z) + 2.0) * (torch.cos(2 * X) + torch.cos(2 * Y)) - 2.0) / 16.0)
    g.post_bc()
    g.get_conserved()
