"""Per-grain c-axis angles -> stiffness table, with exact angle derivatives.

A grain's c-axis is given by colatitude theta (from +z) and azimuth phi
(from +x towards +y): a = (sin t cos p, sin t sin p, cos t) in (x, y, z).
Its stiffness is the transversely isotropic crystal stiffness C0 (axis along
z) rotated by R = Rz(phi) Ry(theta) with the Bond matrix, C = M C0 M^T, and
the angle derivatives follow from the bilinearity of the Bond form,
dM = bond2(dR, R) + bond2(R, dR) (vendor/ringfwi/axisfield.py:37-101,
``rotations``, ``_bond2``, ``axis_stiffness_and_derivs``).  R differs from
the Rodrigues rotation of ``materials.rotation_z_to`` by a turn about the
crystal axis, which leaves a TI stiffness unchanged, so
``stiffness_table(theta, phi)`` equals ``materials.material_tables(axes)``
for the same axes (to rounding).

``angle_projections`` packs dC/dtheta and dC/dphi per material as the
projection rows of the adjoint's gradient reduction, so the GPU returns the
per-grain angle gradient directly (axisfield.py:151-160 chain rule);
``stiffness_table_torch`` is the same map in torch for autograd.
"""
from __future__ import annotations

import numpy as np

from . import materials

_PAIRS = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def axes_from_angles(theta, phi):
    t = np.asarray(theta, float)
    p = np.asarray(phi, float)
    return np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], -1)


def angles_from_axes(axes):
    a = np.asarray(axes, float)
    a = a / np.linalg.norm(a, axis=-1, keepdims=True)
    return np.arccos(np.clip(a[..., 2], -1.0, 1.0)), np.arctan2(a[..., 1], a[..., 0])


def rotations(theta, phi):
    """R = Rz(phi) Ry(theta) per grain and its two angle derivatives, (G, 3, 3) each."""
    t = np.asarray(theta, float).ravel()
    p = np.asarray(phi, float).ravel()
    ct, st, cp, sp = np.cos(t), np.sin(t), np.cos(p), np.sin(p)
    G = t.size
    Ry = np.zeros((G, 3, 3))
    dRy = np.zeros((G, 3, 3))
    Ry[:, 0, 0], Ry[:, 0, 2], Ry[:, 1, 1], Ry[:, 2, 0], Ry[:, 2, 2] = ct, st, 1.0, -st, ct
    dRy[:, 0, 0], dRy[:, 0, 2], dRy[:, 2, 0], dRy[:, 2, 2] = -st, ct, -ct, -st
    Rz = np.zeros((G, 3, 3))
    dRz = np.zeros((G, 3, 3))
    Rz[:, 0, 0], Rz[:, 0, 1], Rz[:, 1, 0], Rz[:, 1, 1], Rz[:, 2, 2] = cp, -sp, sp, cp, 1.0
    dRz[:, 0, 0], dRz[:, 0, 1], dRz[:, 1, 0], dRz[:, 1, 1] = -sp, -cp, cp, -sp
    return Rz @ Ry, Rz @ dRy, dRz @ Ry


def bond2(A, B):
    """Bilinear Bond form (G,3,3) x (G,3,3) -> (G,6,6); bond2(R, R) is the
    Bond matrix of R (C' = M C M^T for engineering-shear Voigt stiffness)."""
    G = A.shape[0]
    M = np.zeros((G, 6, 6))
    for a, (i, j) in enumerate(_PAIRS):
        for b, (k, l) in enumerate(_PAIRS):
            M[:, a, b] = A[:, i, k] * B[:, j, l]
            if k != l:
                M[:, a, b] += A[:, i, l] * B[:, j, k]
    return M


def ti_stiffness_angles(theta, phi, material=materials.ICE):
    """(C, dC/dtheta, dC/dphi), each (G, 6, 6), of grains with c-axes (theta, phi)."""
    C0 = materials.ti_stiffness_6(**material)
    R, dRt, dRp = rotations(theta, phi)
    M = bond2(R, R)
    dMt = bond2(dRt, R) + bond2(R, dRt)
    dMp = bond2(dRp, R) + bond2(R, dRp)
    MC = M @ C0
    Mt = np.swapaxes(M, 1, 2)
    C = MC @ Mt
    dCt = (dMt @ C0) @ Mt + MC @ np.swapaxes(dMt, 1, 2)
    dCp = (dMp @ C0) @ Mt + MC @ np.swapaxes(dMp, 1, 2)
    return C, dCt, dCp


def stiffness_table(theta, phi, c_couplant=1500.0, rho_couplant=300.0,
                    material=materials.ICE, dtype=np.float64):
    """(Ct (G+1, 36), rho_t (G+1,)) with row 0 the couplant, as
    ``materials.material_tables`` (sim/core/fdtd.py:1580-1603)."""
    C, _, _ = ti_stiffness_angles(theta, phi, material)
    G = C.shape[0]
    Ct = np.empty((G + 1, 36))
    Ct[0] = materials.fluid_stiffness_6(c_couplant, rho_couplant).reshape(36)
    Ct[1:] = C.reshape(G, 36)
    rho = np.full(G + 1, float(material["rho"]))
    rho[0] = rho_couplant
    return Ct.astype(dtype), rho.astype(dtype)


def angle_projections(theta, phi, material=materials.ICE):
    """(G+1, 2, 36): rows dC/dtheta and dC/dphi of every material (zero for
    the couplant row 0), the projection table of the adjoint reduction."""
    _, dCt, dCp = ti_stiffness_angles(theta, phi, material)
    G = dCt.shape[0]
    P = np.zeros((G + 1, 2, 36))
    P[1:, 0] = dCt.reshape(G, 36)
    P[1:, 1] = dCp.reshape(G, 36)
    return P


def stiffness_table_torch(theta, phi, c_couplant=1500.0, rho_couplant=300.0,
                          material=materials.ICE):
    """Differentiable (G+1, 36) float64 table from torch angle tensors (G,)."""
    import torch
    t = theta.to(torch.float64)
    p = phi.to(torch.float64)
    ct, st, cp, sp = torch.cos(t), torch.sin(t), torch.cos(p), torch.sin(p)
    z, o = torch.zeros_like(t), torch.ones_like(t)
    Ry = torch.stack([torch.stack([ct, z, st], -1), torch.stack([z, o, z], -1),
                      torch.stack([-st, z, ct], -1)], -2)
    Rz = torch.stack([torch.stack([cp, -sp, z], -1), torch.stack([sp, cp, z], -1),
                      torch.stack([z, z, o], -1)], -2)
    R = Rz @ Ry
    rows = []
    for (i, j) in _PAIRS:
        cols = []
        for (k, l) in _PAIRS:
            e = R[:, i, k] * R[:, j, l]
            if k != l:
                e = e + R[:, i, l] * R[:, j, k]
            cols.append(e)
        rows.append(torch.stack(cols, -1))
    M = torch.stack(rows, -2)
    C0 = torch.as_tensor(materials.ti_stiffness_6(**material), dtype=torch.float64,
                         device=t.device)
    C = M @ C0 @ M.transpose(1, 2)
    fluid = torch.as_tensor(materials.fluid_stiffness_6(c_couplant, rho_couplant).reshape(1, 36),
                            dtype=torch.float64, device=t.device)
    return torch.cat([fluid, C.reshape(-1, 36)], 0)
