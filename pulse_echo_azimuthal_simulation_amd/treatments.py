"""Equivalent-medium stiffness tables: material treatments that reuse the
label-indexed kernels unchanged by adding table rows (host, NumPy).

* Contrast modes of the upstream production runner
  (sim/runs/master_run.py:91-105): ``"zero"`` gives every grain the first
  grain's tensor (no acoustic contrast), ``"contrast"`` pulls every grain a
  fraction ``frac`` of the way from the orientation mean to its own tensor
  (scattering scales as frac**2), in float32 as upstream.
* Interface treatments of the tilted-interface testbed
  (sim/validation/tilt_testbed.py:87-180): a plane between two grains is
  represented by extra labels, one per volume-fraction bin, whose rows are
  ``"naive"`` (the majority grain), ``"aa"`` (stiffness linear in the
  fraction) or ``"sm"`` (the Schoenberg-Muir / Backus equivalent medium of
  the two layers about the exact plane normal).

The Schoenberg-Muir average (Schoenberg & Muir 1989, Geophysics 54(5) 581)
of layers normal to x3 splits Voigt indices into the normal group
N = {3, 4, 5} and the tangential group T = {1, 2, 6}; with <.> the
volume-fraction mean:

    C_NN* = <C_NN^-1>^-1
    C_TN* = <C_TN C_NN^-1> C_NN*
    C_TT* = <C_TT - C_TN C_NN^-1 C_NT> + <C_TN C_NN^-1> C_NN* <C_NN^-1 C_NT>
"""
from __future__ import annotations

import numpy as np

from . import materials

_NRM = [2, 3, 4]          # Voigt 3, 4, 5
_TAN = [0, 1, 5]          # Voigt 1, 2, 6


# ------------------------------------------------------------ contrast modes
def contrast_tables(Ct, rho_t, mode: str, frac: float | None = None):
    """(Ct, rho_t) with a contrast mode applied to the grain rows 1.. (row 0,
    the couplant, is kept).  mode: "none", "zero" or "contrast" (needs frac)."""
    Ct = np.asarray(Ct, np.float32).reshape(len(rho_t), 36)
    if mode == "none":
        return Ct, rho_t
    out = Ct.copy()
    if mode == "zero":
        out[1:] = out[1]
    elif mode == "contrast":
        if frac is None:
            raise ValueError("mode 'contrast' needs frac")
        mean = out[1:].mean(axis=0, keepdims=True)
        out[1:] = mean + frac * (out[1:] - mean)
    else:
        raise ValueError(f"unknown contrast mode {mode!r}")
    return out, rho_t


# ------------------------------------------------------ interface treatments
def schoenberg_muir(Cs, fs) -> np.ndarray:
    """Equivalent 6x6 stiffness of layers normal to x3 with volume fractions
    fs (summing to 1), float64."""
    inv_nn = np.zeros((3, 3))
    tn_inv = np.zeros((3, 3))
    inv_nt = np.zeros((3, 3))
    tt_red = np.zeros((3, 3))
    for C, f in zip(Cs, fs):
        C = np.asarray(C, float)
        Cnn = C[np.ix_(_NRM, _NRM)]
        Ctn = C[np.ix_(_TAN, _NRM)]
        Ctt = C[np.ix_(_TAN, _TAN)]
        Cnn_i = np.linalg.inv(Cnn)
        inv_nn += f * Cnn_i
        tn_inv += f * (Ctn @ Cnn_i)
        inv_nt += f * (Cnn_i @ Ctn.T)
        tt_red += f * (Ctt - Ctn @ Cnn_i @ Ctn.T)
    nn = np.linalg.inv(inv_nn)
    tn = tn_inv @ nn
    tt = tt_red + tn_inv @ nn @ inv_nt
    E = np.zeros((6, 6))
    E[np.ix_(_NRM, _NRM)] = nn
    E[np.ix_(_TAN, _NRM)] = tn
    E[np.ix_(_NRM, _TAN)] = tn.T
    E[np.ix_(_TAN, _TAN)] = tt
    return 0.5 * (E + E.T)


def _rotation_to_x3(normal) -> np.ndarray:
    """Rotation R with R @ normal = e3 (Rodrigues).  A normal already of
    unit length to 1e-12 is used as given (the rounding of its components
    carries into R, as in tilt_testbed.py:157-166)."""
    n = np.asarray(normal, float)
    nn = np.linalg.norm(n)
    if abs(nn - 1.0) > 1e-12:
        n = n / nn
    e3 = np.array([0.0, 0.0, 1.0])
    v = np.cross(n, e3)
    s, c = np.linalg.norm(v), float(n @ e3)
    if s < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    K = np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])
    return np.eye(3) + K + K @ K * ((1.0 - c) / s ** 2)


def grain_stiffness(axis) -> np.ndarray:
    """6x6 float64 stiffness of an ice grain with the given c-axis."""
    return materials.ti_stiffness_3d(np.asarray(axis, float))


def interface_tables(CA, CB, normal, treatment: str, nbin: int = 21,
                     rho: float = materials.ICE["rho"]):
    """(nbin + 1, 36) float32 table and (nbin + 1,) densities for a plane
    between grain A (near side) and grain B: row 1 + i holds volume
    fraction f_i = i / (nbin - 1) of A; row 0 (label -1) repeats row 1.
    ``normal`` is the plane normal in (x, y, z) order."""
    if treatment not in ("naive", "aa", "sm"):
        raise ValueError(f"unknown treatment {treatment!r}")
    CA = np.asarray(CA, float)
    CB = np.asarray(CB, float)
    R = _rotation_to_x3(normal)
    if treatment == "sm":
        CAr = materials.rotate_voigt_bond(CA, R)
        CBr = materials.rotate_voigt_bond(CB, R)
    rows = np.zeros((nbin, 36), np.float32)
    for i, fa in enumerate(np.linspace(0.0, 1.0, nbin)):
        if treatment == "naive":
            E = CA if fa > 0.5 else CB
        elif treatment == "aa":
            E = fa * CA + (1.0 - fa) * CB
        else:
            E = materials.rotate_voigt_bond(schoenberg_muir([CAr, CBr], [fa, 1.0 - fa]), R.T)
        rows[i] = np.asarray(E, np.float32).reshape(-1)
    table = np.vstack([rows[:1], rows])
    return table, np.full(nbin + 1, rho, np.float32)


def fraction_field(shape, h: float, normal, x0, sup: int = 4) -> np.ndarray:
    """float32 (nz, ny, nx): volume fraction of the near half-space
    n . (x - x0) < 0 in each cell (centre (i + 1/2) h), by sup**3 point
    sampling; ``normal`` and ``x0`` in (x, y, z) order."""
    nz, ny, nx = shape
    n = np.asarray(normal, float)
    x0 = np.asarray(x0, float)
    offs = (np.arange(sup) + 0.5) / sup - 0.5
    hits = np.zeros(shape, np.float32)
    for dz in offs:
        z = (np.arange(nz) + 0.5 + dz) * h
        for dy in offs:
            y = (np.arange(ny) + 0.5 + dy) * h
            for dx in offs:
                x = (np.arange(nx) + 0.5 + dx) * h
                s = (n[2] * (z - x0[2])[:, None, None] + n[1] * (y - x0[1])[None, :, None]
                     + n[0] * (x - x0[0])[None, None, :])
                hits += (s < 0).astype(np.float32)
    return hits / sup ** 3


def fraction_labels(frac, nbin: int = 21) -> np.ndarray:
    """int32 labels (bin of the nearest fraction i / (nbin - 1))."""
    return np.rint(np.asarray(frac) * (nbin - 1)).astype(np.int32)
