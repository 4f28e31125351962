"""Pinning of the compressible-step oracle (oracle/cgrid_step.py), CPU.

  * RK45 coefficients (src/utilities/RKCoeffs.F90): the low-storage scheme's
    amplification factor for y' = z y matches exp(z) through z^4 (fourth
    order), and its time weights add up to the step;
  * fluxes (getRHS_x/y/z, get_tau, get_q): with the artificial properties off,
    the right-hand side equals the compressible Navier-Stokes equations
    evaluated with FFT derivatives on smooth periodic fields (inviscid and
    with the power-law viscosity), so every flux term, sign and component is
    fixed by the equations themselves;
  * the periodic Gaussian filter: its transfer function
    agf + 2 sum_m c_m cos(m k), and unit gain at k = 0;
  * the full step with LAD and the filter conserves mass, momentum and
    energy and keeps a uniform state;
  * bounded axes: the step with the symmetric bc pair equals the periodic
    step of the mirror-extended state on its half.
"""
import numpy as np
import pytest

from oracle import cgrid_step as cg
from oracle import pade_dense as pd


def test_rk45_order_conditions():
    # one step of y' = z y: the low-storage recursion
    def amp(z):
        y, q = 1.0, 0.0
        for a, b in zip(cg.RK45_A, cg.RK45_B):
            q = z * y + a * q
            y = y + b * q
        return y
    # Taylor coefficients of the amplification factor from a polynomial fit
    zs = np.linspace(-0.05, 0.05, 11)
    c = np.polynomial.polynomial.polyfit(zs, [amp(z) for z in zs], 5)
    for k, want in enumerate([1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0]):
        assert abs(c[k] - want) < 1e-9, (k, c[k], want)
    # the time advance over the substeps is the step
    t, q = 0.0, 0.0
    for a, b in zip(cg.RK45_A, cg.RK45_B):
        q = 1.0 + a * q
        t += b * q
    assert abs(t - 1.0) < 1e-14


def test_gaussian_transfer_function():
    g = pd.gaussian_coeffs()
    assert abs(g[0] + 2 * sum(g[1:]) - 1.0) < 1e-15
    n = 32
    G = pd.dense_operator("gaussian", "filter", n, 1.0, True)
    x = 2 * np.pi * np.arange(n) / n
    for m in range(n // 2 + 1):
        k = 2 * np.pi * m / n
        T = g[0] + 2 * sum(g[j] * np.cos(j * k) for j in range(1, 5))
        f = np.cos(m * x + 0.3)
        assert np.abs(G @ f - T * f).max() < 1e-14


def _smooth_state(shape, seed):
    rng = np.random.default_rng(seed)
    h = tuple(2 * np.pi / n for n in shape)
    X, Y, Z = np.meshgrid(*[np.arange(n) * hh for n, hh in zip(shape, h)], indexing="ij")

    def modes(a0, amp):
        f = np.full(shape, a0)
        for _ in range(3):
            k = rng.integers(-2, 3, size=3)
            f = f + amp * rng.standard_normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6.28))
        return f
    s = dict(rho=modes(1.0, 0.05), u=modes(0.3, 0.2), v=modes(-0.2, 0.2), w=modes(0.1, 0.2), p=modes(2.0, 0.1))
    return s, h


def _fft_d(f, a, L=2 * np.pi):
    n = f.shape[a]
    k = np.fft.fftfreq(n, d=L / n) * 2 * np.pi
    shp = [1, 1, 1]
    shp[a] = n
    F = np.fft.fft(f, axis=a) * (1j * k.reshape(shp))
    if n % 2 == 0:
        idx = [slice(None)] * 3
        idx[a] = n // 2
        F[tuple(idx)] = 0.0
    return np.real(np.fft.ifft(F, axis=a))


def _ns_rhs_fft(s, gam, Rgas, mu_fn=None, Pr=0.75):
    """Compressible Navier-Stokes right-hand side with FFT derivatives:
    d_t rho = -div(rho u); d_t (rho u_i) = -div(rho u_i u + p e_i - tau_i);
    d_t E = -div((E + p) u - tau.u + q), tau = mu (grad u + grad u^T -
    2/3 div u I), q = -kappa grad T, kappa = Cp mu / Pr."""
    rho, p = s["rho"], s["p"]
    U = [s["u"], s["v"], s["w"]]
    e = p / ((gam - 1.0) * rho)
    T = (gam - 1.0) * e / Rgas
    E = rho * (e + 0.5 * sum(c * c for c in U))
    G = [[_fft_d(U[i], j) for j in range(3)] for i in range(3)]     # G[i][j] = d u_i / d x_j
    if mu_fn is None:
        mu = np.zeros_like(rho)
    else:
        mu = mu_fn(T)
    kap = gam * Rgas / (gam - 1.0) * mu / Pr
    div = G[0][0] + G[1][1] + G[2][2]
    tau = [[mu * (G[i][j] + G[j][i]) - (2.0 / 3.0) * mu * div * (i == j) for j in range(3)] for i in range(3)]
    q = [-kap * _fft_d(T, j) for j in range(3)]
    R = np.zeros(rho.shape + (5,))
    for j in range(3):
        R[..., 0] -= _fft_d(rho * U[j], j)
        for i in range(3):
            R[..., 1 + i] -= _fft_d(rho * U[i] * U[j] + p * (i == j) - tau[i][j], j)
        R[..., 4] -= _fft_d((E + p) * U[j] - sum(tau[i][j] * U[i] for i in range(3)) + q[j], j)
    return R


@pytest.mark.parametrize("inviscid", [True, False])
def test_fluxes_are_navier_stokes(inviscid):
    shape = (40, 36, 32)
    s, h = _smooth_state(shape, 5)
    g = cg.Grid(shape, h, gam=1.4, Rgas=1.0, inviscid=inviscid, mu_ref=0.05, T_ref=1.0, visc_n=0.75, Pr=0.75)
    s = g.from_p(dict(s))
    R = g.rhs(g.conserved(s), s, lad=False)
    want = _ns_rhs_fft(s, 1.4, 1.0, None if inviscid else (lambda T: 0.05 * T ** 0.75), 0.75)
    for k in range(5):
        err = np.abs(R[..., k] - want[..., k]).max() / np.abs(want[..., k]).max()
        assert err < 5e-6, (k, err)


def test_step_conserves_and_keeps_uniform_state():
    shape = (16, 16, 16)
    s, h = cg.taylorgreen(shape)
    g = cg.Grid(shape, h, gam=5.0 / 3.0)
    W0 = g.conserved(g.from_p(dict(s)))
    s1 = g.advance(s, 2e-3)
    W1 = g.conserved(s1)
    for k in range(5):
        tot0, tot1 = W0[..., k].sum(), W1[..., k].sum()
        scale = max(np.abs(W0[..., k]).sum(), np.abs(W0[..., 1:4]).sum())   # w starts at zero
        assert abs(tot1 - tot0) <= 1e-13 * scale, (k, tot0, tot1)
    assert np.abs(W1 - W0).max() > 1e-6          # the flow did evolve
    # a uniform state stays uniform (every derivative and filter of a
    # constant is exact)
    shape = (12, 10, 14)
    h = tuple(2 * np.pi / n for n in shape)
    u = dict(rho=np.full(shape, 1.3), u=np.full(shape, 0.4), v=np.full(shape, -0.2), w=np.full(shape, 0.1),
             p=np.full(shape, 2.5))
    g = cg.Grid(shape, h, inviscid=False)
    s1 = g.advance(u, 1e-2)
    for c in ("rho", "u", "v", "w", "p"):
        assert np.abs(s1[c] - u[c]).max() < 1e-12 * np.abs(u[c]).max(), c


def test_bounded_mirror_step_is_half_of_periodic_step():
    """Bounded axes with the symmetric pair (1, 1) on a state that is even
    about x = 0, pi and y = 0, pi (u odd in x, v odd in y): one full step with
    LAD and viscosity equals the periodic step on the doubled grid restricted
    to the half (nodes 0..n/2).  A wrong flux, gradient or filter sign flip
    (cgrid.F90:446-478, :663-717, getRHS_x/y/z) breaks the symmetry."""
    rng = np.random.default_rng(9)
    full, half = (32, 24, 8), (17, 13, 8)
    h = tuple(2 * np.pi / n for n in full)
    X, Y, Z = np.meshgrid(*[np.arange(n) * hh for n, hh in zip(full, h)], indexing="ij")

    def modes(a0, amp, fx, fy):
        f = np.full(full, a0)
        for _ in range(3):
            kx, ky, kz = rng.integers(1, 3, size=3)
            f = f + amp * rng.standard_normal() * fx(kx * X) * fy(ky * Y) * np.cos(kz * Z + rng.uniform(0, 6.28))
        return f
    s = dict(rho=modes(1.0, 0.05, np.cos, np.cos), u=modes(0.0, 0.2, np.sin, np.cos),
             v=modes(0.0, 0.2, np.cos, np.sin), w=modes(0.1, 0.2, np.cos, np.cos),
             p=modes(2.0, 0.1, np.cos, np.cos))
    kw = dict(gam=1.4, Rgas=1.0, inviscid=False, mu_ref=0.02, T_ref=1.0, visc_n=0.75, Pr=0.7)
    sp = cg.Grid(full, h, **kw).advance(s, 2e-3)
    sh = {c: s[c][:17, :13].copy() for c in s}
    sb = cg.Grid(half, h, periodic=(False, False, True), bc=((1, 1), (1, 1), (0, 0)), **kw).advance(sh, 2e-3)
    for c in ("rho", "u", "v", "w", "p", "T"):
        err = np.abs(sb[c] - sp[c][:17, :13]).max() / np.abs(sp[c]).max()
        assert err < 1e-12, (c, err)
