"""CPU ORACLE / CPU BASELINE -- test and benchmark infrastructure only.

A NumPy restatement of the upstream array-path solver loop
``fdtd.forward`` (sim/core/fdtd.py:236-376, derivative and move helpers
:160-207), float64 on the CPU.  Two uses:

* the benchmark's ``cpu_baseline`` leg times it ("the reference solver's
  update loop run with NumPy", single-threaded);
* tests check the C oracle's float64 step against it: without damping the
  two formulations are the same scheme evaluated in a different
  floating-point order (slice-wise derivatives scaled per term, moves as
  successive two-point averages), so they agree to rounding.

Inputs follow the label path: a material index volume and (Ct, rho_t)
tables, expanded here to the dense 21-map form the array path uses.
Damping, when given, is applied to v and s after the velocity update, as in
the array path.
"""
from __future__ import annotations

import numpy as np

_POS = {"c": (0, 0, 0), "yz": (1, 1, 0), "xz": (1, 0, 1), "xy": (0, 1, 1)}
_VPOS = {1: "c", 2: "c", 3: "c", 4: "yz", 5: "xz", 6: "xy"}


def deriv(f, ax, coeffs, inv_h, out, forward):
    """Staggered derivative along ``ax`` (sim/core/fdtd.py:160-184):
    forward: out[i] += c_j/h (f[i+j+1] - f[i-j]); backward: out[i] += c_j/h
    (f[i+j] - f[i-j-1]); each term only where both taps exist."""
    out.fill(0)
    n = f.shape[ax]
    for j, c in enumerate(coeffs):
        hi, lo = (j + 1, j) if forward else (j, j + 1)
        m = n - hi - lo
        if m <= 0:
            break
        so = [slice(None)] * f.ndim
        sh = [slice(None)] * f.ndim
        sl = [slice(None)] * f.ndim
        so[ax] = slice(lo, lo + m)
        sh[ax] = slice(hi + lo, hi + lo + m)
        sl[ax] = slice(0, m)
        out[tuple(so)] += (c * inv_h) * (f[tuple(sh)] - f[tuple(sl)])
    return out


def _avg(f, ax, d):
    out = np.zeros_like(f)
    n = f.shape[ax]
    a = [slice(None)] * f.ndim
    b = [slice(None)] * f.ndim
    a[ax] = slice(0, n - 1)
    b[ax] = slice(1, n)
    if d > 0:
        out[tuple(a)] = 0.5 * (f[tuple(a)] + f[tuple(b)])
    else:
        out[tuple(b)] = 0.5 * (f[tuple(a)] + f[tuple(b)])
    return out


def move(f, src, dst):
    """Average a field from staggered position ``src`` to ``dst``, one axis
    at a time (sim/core/fdtd.py:187-207)."""
    for ax in range(3):
        d = _POS[dst][ax] - _POS[src][ax]
        if d:
            f = _avg(f, ax, d)
    return f


def face_density(rho, ax):
    a = np.asarray(rho, float)
    m = a.copy()
    n = a.shape[ax]
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[ax] = slice(0, n - 1)
    hi[ax] = slice(1, n)
    m[tuple(lo)] = 0.5 * (a[tuple(lo)] + a[tuple(hi)])
    return m


class ArrayPathShot:
    """State and update loop of the array-path solver for one shot."""

    def __init__(self, mats, Ct, rho_t, coeffs, h, dt, src, rec, damp=None):
        mats = np.asarray(mats)
        Ct = np.asarray(Ct, float).reshape(-1, 36)
        rho_t = np.asarray(rho_t, float)
        self.shape = mats.shape
        self.C = {}
        for I in range(1, 7):
            for J in range(I, 7):
                self.C[(I, J)] = Ct[:, (I - 1) * 6 + (J - 1)][mats]
        self.active = {(I, J): bool(np.any(self.C[(min(I, J), max(I, J))]))
                       for I in range(1, 7) for J in range(1, 7)}
        rho = rho_t[mats]
        self.rv = [face_density(rho, 2), face_density(rho, 1), face_density(rho, 0)]
        self.coeffs = np.asarray(coeffs, float)
        self.inv_h = 1.0 / h
        self.dt = dt
        self.v = [np.zeros(self.shape) for _ in range(3)]
        self.s = {I: np.zeros(self.shape) for I in range(1, 7)}
        self.buf = [np.zeros(self.shape) for _ in range(3)]
        self.acc = np.zeros(self.shape)
        self.src = src
        self.rec = rec
        self.damp = None if damp is None else np.asarray(damp, float)

    def step(self, wn):
        """One time step (sim/core/fdtd.py:310-375); returns the receiver row."""
        X, Y, Z = -1, -2, -3
        co, ih, b = self.coeffs, self.inv_h, self.buf
        vx, vy, vz = self.v
        s = self.s
        g = {1: deriv(vx, X, co, ih, b[0], False).copy(),
             2: deriv(vy, Y, co, ih, b[0], False).copy(),
             3: deriv(vz, Z, co, ih, b[0], False).copy(),
             4: deriv(vz, Y, co, ih, b[0], True).copy() + deriv(vy, Z, co, ih, b[1], True),
             5: deriv(vz, X, co, ih, b[0], True).copy() + deriv(vx, Z, co, ih, b[1], True),
             6: deriv(vy, X, co, ih, b[0], True).copy() + deriv(vx, Y, co, ih, b[1], True)}
        moved = {}
        for I in range(1, 7):
            pos = _VPOS[I]
            self.acc.fill(0)
            anyterm = False
            for J in range(1, 7):
                if not self.active[(I, J)]:
                    continue
                key = (J, pos)
                if key not in moved:
                    moved[key] = g[J] if _VPOS[J] == pos else move(g[J], _VPOS[J], pos)
                self.acc += self.C[(min(I, J), max(I, J))] * moved[key]
                anyterm = True
            if anyterm:
                s[I] += self.dt * self.acc
        idx, w = self.src
        for I in (1, 2, 3):
            s[I].flat[idx] += wn * np.asarray(w)
        vx += (self.dt / self.rv[0]) * (deriv(s[1], X, co, ih, b[0], True)
                                        + deriv(s[6], Y, co, ih, b[1], False)
                                        + deriv(s[5], Z, co, ih, b[2], False))
        vy += (self.dt / self.rv[1]) * (deriv(s[6], X, co, ih, b[0], False)
                                        + deriv(s[2], Y, co, ih, b[1], True)
                                        + deriv(s[4], Z, co, ih, b[2], False))
        vz += (self.dt / self.rv[2]) * (deriv(s[5], X, co, ih, b[0], False)
                                        + deriv(s[4], Y, co, ih, b[1], False)
                                        + deriv(s[3], Z, co, ih, b[2], True))
        if self.damp is not None:
            for a in self.v:
                a *= self.damp
            for I in range(1, 7):
                s[I] *= self.damp
        p = -(s[1] + s[2] + s[3]) / 3.0
        ptr, ridx, rw = self.rec
        pf = p.reshape(-1)
        return np.array([pf[ridx[ptr[r]:ptr[r + 1]]] @ np.asarray(rw[ptr[r]:ptr[r + 1]], float)
                         for r in range(len(ptr) - 1)])

    def run(self, wavelet, nt):
        return np.stack([self.step(float(wavelet[n])) for n in range(int(nt))])
