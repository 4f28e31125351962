"""Seeded synthetic fields of the benchmark configurations (on any torch device).

* ``fourier_modes``: sum of M cosine modes with integer wavenumbers per axis
  uniform in [kmin, kmax], amplitudes N(0,1), phases U(0, 2 pi) (64^3 case).
* ``kolmogorov``: isotropic random field whose shell energy spectrum follows
  E(k) ~ k^(-5/3) for 1 <= |k| <= kmax: white Gaussian noise filtered by
  |k|^(-11/6) (so |u_hat|^2 4 pi k^2 ~ k^(-5/3)), zero mean, unit RMS.

Arrays are returned as Fortran-ordered tensors of shape (n1, n2, n3) on the
periodic box [0, 2 pi)^3.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _fortran(t):
    n1, n2, n3 = t.shape
    out = torch.empty((n3, n2, n1), dtype=t.dtype, device=t.device).permute(2, 1, 0)
    out.copy_(t)
    return out


def fourier_modes(n, modes=32, kmin=1, kmax=21, seed=0, device="cpu", dtype=torch.float64):
    """Returns (f, params) with params = (k[modes,3], amp[modes], phase[modes])."""
    rng = np.random.default_rng(seed)
    k = rng.integers(kmin, kmax + 1, size=(modes, 3))
    amp = rng.standard_normal(modes)
    phase = rng.uniform(0.0, 2 * math.pi, size=modes)
    x = torch.arange(n, dtype=torch.float64, device=device) * (2 * math.pi / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    f = torch.zeros((n, n, n), dtype=torch.float64, device=device)
    for m in range(modes):
        f += amp[m] * torch.cos(int(k[m, 0]) * X + int(k[m, 1]) * Y + int(k[m, 2]) * Z + phase[m])
    return _fortran(f.to(dtype)), (k, amp, phase)


def kolmogorov(shape, kmax, seed=0, device="cpu", dtype=torch.float64):
    n1, n2, n3 = shape
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    # white noise with the memory order of a Fortran (n1, n2, n3) array
    w = torch.randn((n3, n2, n1), generator=g, dtype=torch.float64, device=device)
    W = torch.fft.rfftn(w)
    k3 = torch.fft.fftfreq(n3, d=1.0 / n3, dtype=torch.float64, device=device).reshape(n3, 1, 1)
    k2 = torch.fft.fftfreq(n2, d=1.0 / n2, dtype=torch.float64, device=device).reshape(1, n2, 1)
    k1 = torch.fft.rfftfreq(n1, d=1.0 / n1, dtype=torch.float64, device=device).reshape(1, 1, -1)
    kk = torch.sqrt(k1 * k1 + k2 * k2 + k3 * k3)
    amp = torch.where((kk >= 1.0) & (kk <= kmax), kk.clamp_min(1.0) ** (-11.0 / 6.0),
                      torch.zeros((), dtype=torch.float64, device=device))
    del kk
    W *= amp
    del amp
    u = torch.fft.irfftn(W, s=(n3, n2, n1))
    del W
    u -= u.mean()
    u /= torch.sqrt((u * u).mean())
    return u.to(dtype).permute(2, 1, 0)
