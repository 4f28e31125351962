"""CPU oracle of the compressible solver's Runge-Kutta step for one ideal-gas
species (src/compressible/cgrid.F90 with IdealGasEOS.F90,
PowerLawViscosity.F90 and MixtureEOS.F90 at ns = 1), periodic or bounded per
axis with the reference's bc pairs and their sign flips.

TEST INFRASTRUCTURE ONLY.  Nothing in the product may import this module;
only tests/ use it, as the checker.

Every derivative and filter is a dense FP64 matrix from ``pade_dense``
applied along one axis; the pointwise algebra is restated from the
reference's statements (line citations at each routine).  Conserved
variables W[..., 0..4] = rho, rho u, rho v, rho w, total energy.

Pinning (tests/test_oracle_cgrid.py): the RK45 coefficients by the order
conditions of the low-storage scheme; the inviscid and viscous fluxes
against the compressible Navier-Stokes equations evaluated with FFT
derivatives; conservation of mass, momentum and energy by the full step;
invariance of a uniform state; on bounded axes with the symmetric pair, the
step equals the periodic step of the mirror-extended state on its half (pins
every bc sign flip).  The artificial-property (LAD) formulas are
restated without an independent closed form ("parity unpinned" beyond
conservation and the uniform-state check).
"""
from __future__ import annotations

import numpy as np

from . import pade_dense as pd

RK45_A = (0.0, -6234157559845.0 / 12983515589748.0, -6194124222391.0 / 4410992767914.0,
          -31623096876824.0 / 15682348800105.0, -12251185447671.0 / 11596622555746.0)
RK45_B = (494393426753.0 / 4806282396855.0, 4047970641027.0 / 5463924506627.0,
          9795748752853.0 / 13190207949281.0, 4009051133189.0 / 8539092990294.0,
          1348533437543.0 / 7166442652324.0)       # src/utilities/RKCoeffs.F90:6-17


class Grid:
    def __init__(self, shape, spacing, derivative="cd10", filt="cf90", gam=1.4, Rgas=1.0, Cmu=0.002,
                 Cbeta=1.75, Ckap=0.01, inviscid=True, mu_ref=1.0, T_ref=1.0, visc_n=0.75, Pr=0.75,
                 periodic=(True, True, True), bc=((0, 0), (0, 0), (0, 0))):
        self.shape, self.h = tuple(shape), tuple(spacing)
        self.gam, self.Rgas = gam, Rgas
        self.Cmu, self.Cbeta, self.Ckap = Cmu, Cbeta, Ckap
        self.inviscid = inviscid
        self.mu_ref, self.T_ref, self.visc_n, self.Pr = mu_ref, T_ref, visc_n, Pr
        self.periodic = tuple(bool(p) for p in periodic)
        self.bcs = tuple((0, 0) if p else tuple(b) for p, b in zip(self.periodic, bc))
        self.kinds = {"d1": derivative, "d2": derivative, "fil": filt, "gfil": "gaussian"}
        self._cache = {}
        self.tsim = 0.0
        self.step = 0

    def _mat(self, which, a, sign):
        """Dense operator ``which`` along axis a with the axis's bc pair
        (negated when sign = -1: a quantity odd at the walls)."""
        b1, bn = self.bcs[a]
        key = (which, a, sign * b1, sign * bn)
        if key not in self._cache:
            kind, op = self.kinds[which], {"d1": "d1", "d2": "d2", "fil": "filter", "gfil": "filter"}[which]
            hh = self.h[a] if which in ("d1", "d2") else 1.0
            self._cache[key] = pd.dense_operator(kind, op, self.shape[a], hh, self.periodic[a], sign * b1, sign * bn)
        return self._cache[key]

    # dense operator along axis a of a 3-D array
    @staticmethod
    def _along(M, f, a):
        return np.moveaxis(np.tensordot(M, f, axes=([1], [a])), 0, a)

    def d1(self, f, a, sign=1):
        return self._along(self._mat("d1", a, sign), f, a)

    def d4(self, f, a):
        D2 = self._mat("d2", a, 1)
        return self._along(D2, self._along(D2, f, a), a)

    def grad(self, f, flip=None):
        """cgrid.F90:446-478; the bc of axis ``flip`` negated."""
        return [self.d1(f, a, -1 if a == flip else 1) for a in range(3)]

    def filt(self, f, which, numtimes=1, flip=None):
        """cgrid.F90:1259-1336: y, then x, then z."""
        for a in (1, 0, 2):
            M = self._mat(which, a, -1 if a == flip else 1)
            for _ in range(numtimes):
                f = self._along(M, f, a)
        return f

    # EOS (IdealGasEOS.F90:45-79)
    def primitive(self, W):
        """cgrid.F90:764-797."""
        rho = W[..., 0].copy()
        onebyrho = 1.0 / rho
        u, v, w = W[..., 1] * onebyrho, W[..., 2] * onebyrho, W[..., 3] * onebyrho
        e = W[..., 4] * onebyrho - 0.5 * (u * u + v * v + w * w)
        p = (self.gam - 1.0) * rho * e
        T = (self.gam - 1.0) * e / self.Rgas
        return dict(rho=rho, u=u, v=v, w=w, e=e, p=p, T=T)

    def conserved(self, s):
        """cgrid.F90:799-811."""
        rho, u, v, w, e = s["rho"], s["u"], s["v"], s["w"], s["e"]
        return np.stack([rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))], axis=-1)

    def from_p(self, s):
        """post_bc, cgrid.F90:813-820: e and T from (rho, p)."""
        s["e"] = s["p"] / (self.gam - 1.0) / s["rho"]
        s["T"] = s["e"] * (self.gam - 1.0) / self.Rgas
        return s

    def transport(self, s):
        """MixtureEOS.F90:175-229 (ns = 1), PowerLawViscosity.F90:38-75."""
        z = np.zeros(self.shape)
        if self.inviscid:
            return z, z.copy(), z.copy()
        mu = self.mu_ref * (s["T"] / self.T_ref) ** self.visc_n
        Cp = self.gam * self.Rgas / (self.gam - 1.0)
        return mu, z, Cp * mu / self.Pr

    def lad(self, s, du):
        """cgrid.F90:1058-1257 (one species)."""
        (dudx, dudy, dudz), (dvdx, dvdy, dvdz), (dwdx, dwdy, dwdz) = du
        hx, hy, hz = self.h
        rho = s["rho"]
        func = np.sqrt(dudx ** 2 + 0.5 * (dvdx + dudy) ** 2 + 0.5 * (dwdx + dudz) ** 2
                       + dvdy ** 2 + 0.5 * (dwdy + dvdz) ** 2 + dwdz ** 2)
        mustar = self.d4(func, 0) * hx ** 6
        mustar = mustar + self.d4(func, 2) * hz ** 6
        mustar = mustar + self.d4(func, 1) * hy ** 6
        mustar = self.filt(self.Cmu * rho * np.abs(mustar), "gfil", 2)
        func = dudx + dvdy + dwdz
        g1, g2, g3 = [g * g for g in self.grad(rho)]
        den = g1 + g2 + g3 + 1.0e-32
        bulk = self.d4(func, 0) * hx ** 4 * (hx * g1 / den) ** 2
        bulk = bulk + self.d4(func, 2) * hz ** 4 * (hz * g3 / den) ** 2
        bulk = bulk + self.d4(func, 1) * hy ** 4 * (hy * g2 / den) ** 2
        curl2 = (dwdy - dvdz) ** 2 + (dudz - dwdx) ** 2 + (dvdx - dudy) ** 2
        fsw = func * func / (func * func + curl2 + 1.0e-32)
        fsw = np.where(func >= 0.0, 0.0, fsw)
        bulk = self.filt(self.Cbeta * rho * fsw * np.abs(bulk), "gfil", 2)
        e = s["e"]
        g1, g2, g3 = [g * g for g in self.grad(e)]
        den = g1 + g2 + g3 + 1.0e-32
        kap = self.d4(e, 0) * hx ** 4 * (hx * g1 / den)
        kap = kap + self.d4(e, 2) * hz ** 4 * (hz * g3 / den)
        kap = kap + self.d4(e, 1) * hy ** 4 * (hy * g2 / den)
        sos = np.sqrt(self.gam * s["p"] / rho)
        kap = self.filt(self.Ckap * rho * sos * np.abs(kap) / s["T"], "gfil", 2)
        return mustar, bulk, kap

    def rhs(self, W, s, lad=True):
        """getRHS, cgrid.F90:822-1056 (get_tau :1338-1390, get_q :1392-1444)."""
        du = [self.grad(s[c], flip=a) for a, c in enumerate(("u", "v", "w"))]
        (dudx, dudy, dudz), (dvdx, dvdy, dvdz), (dwdx, dwdy, dwdz) = du
        mu, bulk, kap = self.transport(s)
        if lad:
            ms, bs, ks = self.lad(s, du)
            mu, bulk, kap = mu + ms, bulk + bs, kap + ks
        bambda = (4.0 / 3.0) * mu + bulk
        lam = bulk - (2.0 / 3.0) * mu
        txy = mu * (dudy + dvdx)
        txz = mu * (dudz + dwdx)
        tyz = mu * (dvdz + dwdy)
        txx = bambda * dudx + lam * (dvdy + dwdz)
        tyy = bambda * dvdy + lam * (dudx + dwdz)
        tzz = bambda * dwdz + lam * (dudx + dvdy)
        qx, qy, qz = [-kap * g for g in self.grad(s["T"])]
        u, v, w, p = s["u"], s["v"], s["w"], s["p"]
        ru, rv, rw, TE = W[..., 1], W[..., 2], W[..., 3], W[..., 4]
        fx = [ru, ru * u + p - txx, ru * v - txy, ru * w - txz, (TE + p - txx) * u - v * txy - w * txz + qx]
        fy = [rv, rv * u - txy, rv * v + p - tyy, rv * w - tyz, (TE + p - tyy) * v - u * txy - w * tyz + qy]
        fz = [rw, rw * u - txz, rw * v - tyz, rw * w + p - tzz, (TE + p - tzz) * w - u * txz - v * tyz + qz]
        R = np.zeros(self.shape + (5,))
        for k in range(5):
            R[..., k] = R[..., k] - self.d1(fx[k], 0, 1 if k == 1 else -1)
        for k in range(5):
            R[..., k] = R[..., k] - self.d1(fy[k], 1, 1 if k == 2 else -1)
        for k in range(5):
            R[..., k] = R[..., k] - self.d1(fz[k], 2, 1 if k == 3 else -1)
        return R

    def advance(self, s, dt, lad=True):
        """advance_RK45, cgrid.F90:663-717.  ``s``: primitive state."""
        if self.step == 0:
            s = self.from_p(dict(s))
        Qtmp = np.zeros(self.shape + (5,))
        Qtmpt = 0.0
        for isub in range(5):
            W = self.conserved(s)
            R = self.rhs(W, s, lad)
            Qtmp = dt * R + RK45_A[isub] * Qtmp
            Qtmpt = dt + RK45_A[isub] * Qtmpt
            W = W + RK45_B[isub] * Qtmp
            self.tsim += RK45_B[isub] * Qtmpt
            for k in range(5):
                W[..., k] = self.filt(W[..., k], "fil", 1, (None, 0, 1, 2, None)[k])
            s = self.from_p(self.primitive(W))
        self.step += 1
        return s


def taylorgreen(shape):
    """problems/compressible/taylorgreen_files/hooks.F90 meshgen/initfields."""
    h = tuple(2 * np.pi / n for n in shape)
    x, y, z = [np.arange(n) * hh for n, hh in zip(shape, h)]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    s = dict(rho=np.ones(shape), u=np.sin(X) * np.cos(Y) * np.cos(Z), v=-np.cos(X) * np.sin(Y) * np.cos(Z),
             w=np.zeros(shape), p=100.0 + ((np.cos(2 * Z) + 2.0) * (np.cos(2 * X) + np.cos(2 * Y)) - 2.0) / 16.0)
    return s, h
