"""Convolutional PML absorbing boundary (no upstream counterpart).

The upstream solver absorbs with a multiplicative sponge folded into the
damping array (sim/core/fdtd.py:211-233, 1693-1700).  This module adds a
convolutional perfectly matched layer, selected per shot, as an alternative.
Each spatial derivative dd of the staggered scheme, taken along an axis
inside that axis's absorbing band, is replaced by

    psi' = b * psi + a * dd          (recursive convolution, one per derivative)
    D    = dd / kappa + psi'

with, at normalised depth xi in (0, 1] into the band (xi = 0 elsewhere),

    d(xi)     = d0 * xi**N,   d0 = -(N + 1) * vmax * ln(R) / (2 * L * h)
    kappa(xi) = 1 + (kappa_max - 1) * xi**N
    alpha(xi) = alpha_max * (1 - xi),   alpha_max = pi * f0
    b = exp(-(d / kappa + alpha) * dt)
    a = d * (b - 1) / (kappa * (d + kappa * alpha))      (0 where d = 0)

(the CFS-PML recursion of Roden & Gedney 2000, in the form of Komatitsch &
Martin 2007, Geophysics 72(5) SM155).  A band cell is computed as the plain
cell plus corrections: each derivative along a band axis contributes
corr = dd * (1/kappa - 1) + psi' to its strain or acceleration, so that
dd + corr = D; the memory variables advance in place, once per step.

Depth is measured from the last interior point: on the low side xi = (L - p) / L for p < L, on the high
side xi = (p - (n - 1 - L)) / L for p > n - 1 - L, where p = i for integer
points and i + 1/2 for half points (the two staggers of the grid), capped
at 1.  The kernels keep memory variables on a band of L + 1 cells at each
end of every axis (include/pefdtd.h, struct pe_shot).
"""
from __future__ import annotations

import math

import numpy as np

#: default reflection coefficient and profile power
R_DEFAULT = 1e-4
POWER = 2


def profiles(n: int, width: int, h: float, dt: float, vmax: float, f0: float,
             R: float = R_DEFAULT, kappa_max: float = 1.0, power: int = POWER) -> np.ndarray:
    """float64 (2, 3, n): [stagger 0 integer / 1 half points][b, a, 1/kappa]
    for one axis of n points and a band of ``width`` cells."""
    if width < 1:
        raise ValueError("CPML width must be >= 1")
    if n < 2 * width + 2:
        raise ValueError("CPML width %d needs an axis of >= %d cells (got %d)"
                         % (width, 2 * width + 2, n))
    L = float(width)
    d0 = -(power + 1) * float(vmax) * math.log(R) / (2.0 * L * float(h))
    alpha_max = math.pi * float(f0)
    out = np.empty((2, 3, n))
    for st in (0, 1):
        p = np.arange(n, dtype=float) + 0.5 * st
        xi = np.zeros(n)
        lo = p < L
        hi = p > (n - 1) - L
        xi[lo] = (L - p[lo]) / L
        xi[hi] = (p[hi] - ((n - 1) - L)) / L
        xi = np.minimum(xi, 1.0)
        d = d0 * xi ** power
        kappa = 1.0 + (kappa_max - 1.0) * xi ** power
        alpha = np.where(xi > 0, alpha_max * (1.0 - xi), 0.0)
        b = np.exp(-(d / kappa + alpha) * dt)
        den = kappa * (d + kappa * alpha)
        a = np.where(d > 0, d * (b - 1.0) / np.where(den > 0, den, 1.0), 0.0)
        out[st, 0] = b
        out[st, 1] = a
        out[st, 2] = 1.0 / kappa
    return out


def band_volumes(shape, width: int):
    """Cells of one memory-variable component per axis (z, y, x)."""
    nz, ny, nx = shape
    W = 2 * width + 2
    return (W * ny * nx, nz * W * nx, nz * ny * W)


def dominant_frequency(wavelet, dt: float) -> float:
    """Frequency (Hz) of the amplitude-spectrum peak of a sampled wavelet."""
    w = np.asarray(wavelet, float)
    nfft = 1 << int(math.ceil(math.log2(max(16, 4 * len(w)))))
    spec = np.abs(np.fft.rfft(w, nfft))
    spec[0] = 0.0
    return float(np.argmax(spec)) / (nfft * dt)


class Cpml:
    """Host description of a CPML boundary for one shot: width (cells), the
    per-axis profiles (float64, see ``profiles``) and the band volumes."""

    def __init__(self, shape, width, h, dt, vmax, f0, R=R_DEFAULT, kappa_max=1.0):
        self.shape = tuple(int(s) for s in shape)
        self.width = int(width)
        self.prof = [profiles(n, self.width, h, dt, vmax, f0, R, kappa_max) for n in self.shape]
        self.vol = band_volumes(self.shape, self.width)
        self.params = dict(width=self.width, h=float(h), dt=float(dt), vmax=float(vmax),
                           f0=float(f0), R=float(R), kappa_max=float(kappa_max))


def vmax_table(Ct, rho_t) -> float:
    """Largest P-wave speed sqrt(max(C11, C22, C33) / rho) over the table rows."""
    Ct = np.asarray(Ct, float).reshape(-1, 36)
    rho = np.asarray(rho_t, float)
    cmax = Ct[:, [0, 7, 14]].max(axis=1)
    ok = rho > 0
    return float(np.sqrt(cmax[ok] / rho[ok]).max())
