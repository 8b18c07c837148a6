"""Per-grain stiffness and density tables for the label-indexed solver.

The device kernels never see a stiffness volume: a cell carries an integer
material index m = label + 1, and the kernels read row m of a small
(n_mat, 36) stiffness table and an (n_mat,) density table.  This module
builds those tables on the host in float64 and casts them to the solver's
precision.

Reference behaviour followed (upstream file:line):

* ``ti_stiffness_6``      -- vendor/ringfwi/anisotropy.py:373-382
* ``voigt6_to_tensor`` / ``tensor_to_voigt6`` -- anisotropy.py:311-329
* ``rotate_voigt_3d``     -- anisotropy.py:332-336 (einsum over the 4-tensor)
* ``rotation_z_to``       -- anisotropy.py:339-351 (Rodrigues, z -> axis)
* ``polycrystal row``     -- anisotropy.py:418-456 (couplant: K = rho c^2 in
                             the upper 3x3, zero shear)
* ``material_tables``     -- sim/core/fdtd.py:1580-1603 (row 0 = couplant)
* ``bond_matrix``/``rotate_voigt_bond`` -- sim/validation/tilt_testbed.py:105-124
"""
from __future__ import annotations

import numpy as np

# Single-crystal ice Ih (Pa, kg/m^3), c-axis along x3.
ICE = dict(C11=13.93e9, C33=15.01e9, C44=3.01e9, C12=7.08e9, C13=5.77e9, rho=917.0)

_VOIGT_PAIRS = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def ti_stiffness_6(C11, C33, C44, C12, C13, **_unused) -> np.ndarray:
    """6x6 Voigt stiffness of a transversely isotropic solid, axis along z,
    with C66 = (C11 - C12) / 2."""
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = C11
    C[2, 2] = C33
    C[0, 1] = C[1, 0] = C12
    C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = C13
    C[3, 3] = C[4, 4] = C44
    C[5, 5] = (C11 - C12) / 2.0
    return C


def voigt6_to_tensor(C) -> np.ndarray:
    """6x6 Voigt stiffness (engineering shear strains) -> c_ijkl."""
    C = np.asarray(C, float)
    T = np.zeros((3, 3, 3, 3))
    for a, (i, j) in enumerate(_VOIGT_PAIRS):
        for b, (k, l) in enumerate(_VOIGT_PAIRS):
            for (p, q) in {(i, j), (j, i)}:
                for (r, s) in {(k, l), (l, k)}:
                    T[p, q, r, s] = C[a, b]
    return T


def tensor_to_voigt6(T) -> np.ndarray:
    """c_ijkl -> 6x6 Voigt stiffness."""
    out = np.zeros((6, 6))
    for a, (i, j) in enumerate(_VOIGT_PAIRS):
        for b, (k, l) in enumerate(_VOIGT_PAIRS):
            out[a, b] = T[i, j, k, l]
    return out


def rotation_z_to(axis) -> np.ndarray:
    """Rotation matrix taking +z onto the unit vector along ``axis``
    (Rodrigues: I + [v]x + [v]x^2 / (1 + c), v = z x a, c = z . a; the
    antiparallel case is diag(1, -1, -1))."""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    v = np.array([-a[1], a[0], 0.0])       # z x a
    c = float(a[2])
    if np.linalg.norm(v) < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    K = np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])
    return np.eye(3) + K + K @ K / (1.0 + c)


def rotate_voigt_3d(C, R) -> np.ndarray:
    """Rotate a 6x6 Voigt stiffness by R through the full 4-tensor:
    c'_pqrs = R_pi R_qj R_rk R_sl c_ijkl."""
    T = voigt6_to_tensor(C)
    return tensor_to_voigt6(np.einsum("pi,qj,rk,sl,ijkl->pqrs", R, R, R, R, T))


def ti_stiffness_3d(axis, material=ICE) -> np.ndarray:
    """Full 6x6 stiffness of a TI ``material`` with its symmetry axis along ``axis``."""
    return rotate_voigt_3d(ti_stiffness_6(**material), rotation_z_to(axis))


def fluid_stiffness_6(c_couplant: float, rho_couplant: float) -> np.ndarray:
    """Inviscid couplant: bulk modulus K = rho c^2 in the upper 3x3, no shear."""
    C = np.zeros((6, 6))
    C[:3, :3] = rho_couplant * c_couplant * c_couplant
    return C


def material_tables(axes, c_couplant: float = 1500.0, rho_couplant: float = 300.0,
                    material=ICE, dtype=np.float32):
    """(Ct (n_mat, 36), rho_t (n_mat,)) for grains with c-axes ``axes`` (G, 3).

    Row 0 is the couplant (label -1); row g+1 is grain g's rotated stiffness
    (sim/core/fdtd.py:1580-1603).  Entry [m, 6I+J] holds C_IJ, symmetric.
    """
    axes = np.asarray(axes, float).reshape(-1, 3)
    n_mat = len(axes) + 1
    Ct = np.empty((n_mat, 36))
    rho = np.empty(n_mat)
    Ct[0] = fluid_stiffness_6(c_couplant, rho_couplant).reshape(36)
    rho[0] = rho_couplant
    base = ti_stiffness_6(**material)
    for g, ax in enumerate(axes):
        Ct[g + 1] = rotate_voigt_3d(base, rotation_z_to(ax)).reshape(36)
        rho[g + 1] = material["rho"]
    return Ct.astype(dtype), rho.astype(dtype)


def bond_matrix(R) -> np.ndarray:
    """6x6 Bond matrix M with C' = M C M^T for a rotation R of the frame."""
    R = np.asarray(R, float)
    M = np.zeros((6, 6))
    for a, (i, j) in enumerate(_VOIGT_PAIRS):
        for b, (k, l) in enumerate(_VOIGT_PAIRS):
            if b < 3:
                M[a, b] = R[i, k] * R[j, k]
            else:
                M[a, b] = R[i, k] * R[j, l] + R[i, l] * R[j, k]
    return M


def rotate_voigt_bond(C, R) -> np.ndarray:
    """Rotate a 6x6 Voigt stiffness with the Bond matrix: M C M^T."""
    M = bond_matrix(R)
    return M @ np.asarray(C, float) @ M.T


def active_mask(Ct) -> int:
    """Bit 6I+J set when any material has a nonzero C_IJ (sim/core/fdtd.py:1678-1679).

    Stress terms whose bit is clear are skipped, as the reference does."""
    Ct = np.asarray(Ct).reshape(-1, 36)
    flags = (np.abs(Ct) > 0).any(axis=0)
    mask = 0
    for k in range(36):
        if flags[k]:
            mask |= 1 << k
    return mask


#: the 21 upper-triangle Voigt keys of a dense stiffness volume set
KEYS21 = tuple("C%d%d" % (i, j) for i in range(1, 7) for j in range(i, 7))


def polycrystal_stiffness_3d(labels, axes, c_couplant: float = 1500.0,
                             rho_couplant: float = 300.0, material=ICE):
    """Dense per-cell stiffness and density of a label volume (label -1 =
    couplant, g = grain g): (dict of the 21 upper-triangle Voigt volumes
    "Cij" (float64), rho (float64)), each grain's values those of its
    material-table row (vendor/ringfwi/anisotropy.py:418-456)."""
    labels = np.asarray(labels)
    Ct, rho_t = material_tables(axes, c_couplant, rho_couplant, material, dtype=np.float64)
    m = labels.astype(np.int64) + 1
    C = {}
    for key in KEYS21:
        i, j = int(key[1]) - 1, int(key[2]) - 1
        C[key] = Ct[:, 6 * i + j][m]
    return C, rho_t[m]
