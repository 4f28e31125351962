"""Modified wavenumbers / transfer functions of the periodic operators.

Mirrors the reference's closed forms in src/utilities/spectral.F90:938-1010
(GetCF90TransferFunction, GetCD10ModWaveNum, GetCD10D2ModWaveNum,
GetCD06ModWaveNum) and extends them to the cd06 second derivative.  A Fourier
mode exp(i k x) is an eigenvector of every periodic operator; these give the
eigenvalue, so applying them in Fourier space reproduces the banded operator
to round-off (used by bench.py to report errors at full size).
"""
from __future__ import annotations

import math

import torch

_CD10_D1 = dict(alpha=0.5, beta=1.0 / 20.0, a=(17.0 / 12.0) / 2.0, b=(101.0 / 150.0) / 4.0,
                c=(1.0 / 100.0) / 6.0)
_CD10_D2 = dict(alpha=334.0 / 899.0, beta=43.0 / 1798.0, a=1065.0 / 1798.0, b=(1038.0 / 899.0) / 4.0,
                c=(79.0 / 1798.0) / 9.0)
_CD06_D1 = dict(alpha=1.0 / 3.0, a=(14.0 / 9.0) / 2.0, b=(1.0 / 9.0) / 4.0)
_CD06_D2 = dict(alpha=2.0 / 11.0, a=12.0 / 11.0, b=(3.0 / 11.0) / 4.0)
_CF90 = dict(alpha=6.6624e-1, beta=1.6688e-1, a=9.9965e-1, b=6.6652e-1, c=1.6674e-1, d=4.0e-5, e=-5.0e-6)


def eigenvalue(kind, which, k, dx):
    """Complex eigenvalue of the periodic operator on exp(i k x) (k a tensor)."""
    w = k * dx
    if kind == "cd10" and which == "d1":
        c = _CD10_D1
        num = 2 * c["a"] * torch.sin(w) + 2 * c["b"] * torch.sin(2 * w) + 2 * c["c"] * torch.sin(3 * w)
        den = 1 + 2 * c["alpha"] * torch.cos(w) + 2 * c["beta"] * torch.cos(2 * w)
        return 1j * num / den / dx
    if kind == "cd06" and which == "d1":
        c = _CD06_D1
        num = 2 * c["a"] * torch.sin(w) + 2 * c["b"] * torch.sin(2 * w)
        den = 1 + 2 * c["alpha"] * torch.cos(w)
        return 1j * num / den / dx
    if kind == "cd10" and which == "d2":
        c = _CD10_D2
        num = 2 * c["a"] * (1 - torch.cos(w)) + 2 * c["b"] * (1 - torch.cos(2 * w)) + 2 * c["c"] * (1 - torch.cos(3 * w))
        den = 1 + 2 * c["alpha"] * torch.cos(w) + 2 * c["beta"] * torch.cos(2 * w)
        return -(num / den) / dx ** 2 + 0j
    if kind == "cd06" and which == "d2":
        c = _CD06_D2
        num = 2 * c["a"] * (1 - torch.cos(w)) + 2 * c["b"] * (1 - torch.cos(2 * w))
        den = 1 + 2 * c["alpha"] * torch.cos(w)
        return -(num / den) / dx ** 2 + 0j
    if kind == "cf90" and which == "filter":
        c = _CF90
        num = (c["a"] + 2 * c["b"] * torch.cos(w) + 2 * c["c"] * torch.cos(2 * w) + 2 * c["d"] * torch.cos(3 * w)
               + 2 * c["e"] * torch.cos(4 * w))
        den = 1 + 2 * c["alpha"] * torch.cos(w) + 2 * c["beta"] * torch.cos(2 * w)
        return num / den + 0j
    raise ValueError(f"unknown operator {kind}/{which}")


def apply_fourier(kind, which, f, dim, dx):
    """Periodic operator applied along tensor dimension `dim` through the FFT."""
    n = f.shape[dim]
    k = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64, device=f.device) * (2 * math.pi / (n * dx))
    lam = eigenvalue(kind, which, k, dx)
    if which == "d1" and n % 2 == 0:
        lam[-1] = 0.0  # the Nyquist mode of an odd symbol vanishes
    shape = [1] * f.dim()
    shape[dim] = lam.numel()
    F = torch.fft.rfft(f, dim=dim)
    return torch.fft.irfft(F * lam.reshape(shape), n=n, dim=dim)
