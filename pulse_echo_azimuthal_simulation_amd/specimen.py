"""Seeded synthetic ice-disk specimens and the benchmark shot configurations.

The upstream specimen builder (sim/core/specimen.py:165-351) is a
Laguerre/Voronoi tessellation of a disk with Watson-distributed c-axes.  This
module builds the same kind of input for the solver -- a Voronoi label
volume (int32, -1 = couplant) and per-grain c-axes -- from seeded uniform
centres, with the Watson rejection sampler of sim/core/specimen.py:45-69.
It is an input generator, not a port of the upstream builder (no size-CV
weight calibration, no spatial smoothing).

Two shot configurations are defined:

* ``config_c0``: a 96 x 96 x 48-cell disk raster (dx = 0.1 mm), 200 grains,
  Watson kappa = 5 about azimuth 30 deg, 1 MHz 3-cycle Hann burst from a
  6 mm piston on the top face, padded by absorber + 3 couplant cells on
  every side as sim/core/rotation_test.py:129-132 pads, 2000 steps.
* ``config_c1``: a 400 x 400 x 200 solver grid (32M cells) at
  h = 3850 / 2.25 MHz / 6, holding a 100 mm x 35 mm disk of 3000 grains with
  a Watson girdle (kappa = -3) about a seeded normal, 2.25 MHz Ricker from a
  6 mm piston on the top face, 12000 steps.

Both use the upstream absorbing boundary (20-cell sponge, strength 0.09,
times the exp(-0.02 EDT) couplant mask) and dt = safe_dt_labels(..., 0.5)
with the multistart kh_max 2.0 coefficients.
"""
from __future__ import annotations

import numpy as np

from . import host, materials

C_REF = 3850.0


def sample_watson(n, axis, kappa, rng):
    """n axial unit vectors from a Watson distribution about ``axis``
    (sim/core/specimen.py:45-69): cos(theta) by rejection against
    exp(kappa t^2) (a girdle for kappa < 0), uniform azimuth, rotated from
    +z onto ``axis``."""
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    if abs(kappa) < 1e-9:
        v = rng.normal(size=(n, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    t = np.empty(n)
    got = 0
    while got < n:
        u = rng.uniform(-1, 1, n)
        keep = rng.uniform(0, 1, n) < np.exp(kappa * (u ** 2 - (1 if kappa > 0 else 0)))
        k = u[keep][: n - got]
        t[got:got + k.size] = k
        got += k.size
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(np.clip(1 - t ** 2, 0, None))
    local = np.stack([s * np.cos(phi), s * np.sin(phi), t], axis=1)
    return local @ materials.rotation_z_to(axis).T


def disk_seeds(n, radius, thickness, rng):
    """n uniform points in a disk of ``radius`` and ``thickness`` centred at 0
    (x, y, z order), by rejection from the bounding box."""
    pts = np.empty((0, 3))
    while len(pts) < n:
        p = rng.uniform(-1, 1, (n * 2, 3))
        pts = np.vstack([pts, p[(p[:, 0] ** 2 + p[:, 1] ** 2) <= 1.0]])
    s = pts[:n].copy()
    s[:, :2] *= radius
    s[:, 2] *= thickness / 2
    return s


def voronoi_disk(shape, h, diameter, thickness, n_grains, rng, workers=-1):
    """Label volume (nz, ny, nx) int32 of a centred disk: cell -> index of the
    nearest seed (Voronoi), -1 outside the disk.  Returns (lab, seeds)."""
    from scipy.spatial import cKDTree
    nz, ny, nx = shape
    R = diameter / 2
    seeds = disk_seeds(n_grains, R, thickness, rng)
    x = (np.arange(nx) - (nx - 1) / 2) * h
    y = (np.arange(ny) - (ny - 1) / 2) * h
    z = (np.arange(nz) - (nz - 1) / 2) * h
    inside_xy = (x[None, :] ** 2 + y[:, None] ** 2) <= R * R          # (ny, nx)
    inside_z = np.abs(z) <= thickness / 2
    lab = np.full(shape, -1, np.int32)
    tree = cKDTree(seeds)
    yy, xx = np.nonzero(inside_xy)
    for iz in np.nonzero(inside_z)[0]:
        P = np.stack([x[xx], y[yy], np.full(len(xx), z[iz])], axis=1)
        _, idx = tree.query(P, k=1, workers=workers)
        plane = lab[iz]
        plane[yy, xx] = idx.astype(np.int32)
    return lab, seeds


def piston_footprint(lab, radius_cells):
    """Cells of a circular piston on the top (highest-z) ice face, centred on
    the disk axis: the topmost ice cell of every column within
    ``radius_cells`` of the centre column.  Weights 1/K (pulse-echo uses the
    same footprint to transmit and receive, as sim/pipeline/sweep_runner.py:342-351)."""
    nz, ny, nx = lab.shape
    ice = lab >= 0
    zmax = np.nonzero(ice.any(axis=(1, 2)))[0].max()
    cy, cx = (ny - 1) / 2, (nx - 1) / 2
    pts = []
    for iy in range(ny):
        for ix in range(nx):
            if (iy - cy) ** 2 + (ix - cx) ** 2 <= radius_cells ** 2 and ice[zmax, iy, ix]:
                pts.append((int(zmax), iy, ix))
    w = 1.0 / len(pts)
    return pts, w


def shot_config(lab, axes, h, nt, wavelet_fn, piston_d=6e-3, sponge=20, fluid_damp=0.02,
                coeffs=None):
    """Solver arguments of a pulse-echo shot on a given specimen."""
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True) if coeffs is None else coeffs
    Ct, rho_t = materials.material_tables(axes)
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, co, safety=0.5)
    er = max(int(piston_d / 2 / h), 1)
    pts, w = piston_footprint(lab, er)
    dm = host.fluid_damp_mask(lab, fluid_damp)
    return dict(lab=lab, axes=axes, h=h, dt=dt, nt=int(nt), coeffs=co,
                src_pts=[(p, w) for p in pts], rec_groups=[(pts, np.full(len(pts), w))],
                wavelet=wavelet_fn(dt, int(nt)), wavelet_fn=wavelet_fn, sponge_width=sponge,
                damp_mask=dm, mat_tables=(Ct, rho_t))


def config_c0(seed=7, nt=2000):
    """96x96x48 disk raster (0.1 mm), 200 grains, Watson kappa 5 about 30 deg,
    1 MHz 3-cycle Hann, 6 mm top piston, sponge 20 + 3 couplant cells pad."""
    h = 1e-4
    core = (48, 96, 96)
    rng = np.random.default_rng(seed)
    lab_core, _ = voronoi_disk(core, h, 96 * h, 48 * h, 200, rng)
    m = 20 + 3
    lab = np.full((core[0] + 2 * m, core[1] + 2 * m, core[2] + 2 * m), -1, np.int32)
    lab[m:m + core[0], m:m + core[1], m:m + core[2]] = lab_core
    az = np.radians(30.0)
    axes = sample_watson(200, [np.cos(az), np.sin(az), 0.0], 5.0, rng)
    return shot_config(lab, axes, h, nt,
                       lambda dt, n: host.hann_tone_burst(n, dt, 1e6, n_cycles=3))


def config_c1(seed=11, nt=12000, shape=(200, 400, 400), n_grains=3000):
    """400x400x200 grid at ppw 6 for 2.25 MHz, 100 mm x 35 mm disk of 3000
    grains, Watson girdle kappa -3 about a seeded normal, 2.25 MHz Ricker."""
    f0 = 2.25e6
    h = C_REF / f0 / 6
    rng = np.random.default_rng(seed)
    lab, _ = voronoi_disk(shape, h, 100e-3, 35e-3, n_grains, rng)
    normal = rng.normal(size=3)
    axes = sample_watson(n_grains, normal / np.linalg.norm(normal), -3.0, rng)
    return shot_config(lab, axes, h, nt, lambda dt, n: host.ricker(f0, dt, n))


def perturb_axes(axes, sigma_deg, rng):
    """Tilt every axis by a Gaussian tangent-plane step of sigma_deg per component."""
    a = np.asarray(axes, float)
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    ref = np.where(np.abs(a[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    e1 = np.cross(a, ref)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(a, e1)
    s = np.radians(sigma_deg)
    b = a + s * (rng.normal(size=(len(a), 1)) * e1 + rng.normal(size=(len(a), 1)) * e2)
    return b / np.linalg.norm(b, axis=1, keepdims=True)


def config_c3(seed=7, nt=4000, shape=(80, 160, 160), n_grains=500, sigma_deg=5.0):
    """Adjoint-gradient case: config_c0's geometry on an 80x160x160 grid
    (0.1 mm; disk raster inside 23-cell couplant/sponge pads), 500 grains,
    Watson kappa 5 about 30 deg, 1 MHz 3-cycle Hann, 4000 steps.  Adds
    ``axes_true``: the axes tilted by N(0, sigma_deg), the fabric that makes
    the target traces."""
    h = 1e-4
    m = 20 + 3
    core = (shape[0] - 2 * m, shape[1] - 2 * m, shape[2] - 2 * m)
    rng = np.random.default_rng(seed)
    lab_core, _ = voronoi_disk(core, h, core[2] * h, core[0] * h, n_grains, rng)
    lab = np.full(shape, -1, np.int32)
    lab[m:m + core[0], m:m + core[1], m:m + core[2]] = lab_core
    az = np.radians(30.0)
    axes = sample_watson(n_grains, [np.cos(az), np.sin(az), 0.0], 5.0, rng)
    cfg = shot_config(lab, axes, h, nt,
                      lambda dt, n: host.hann_tone_burst(n, dt, 1e6, n_cycles=3))
    cfg["axes_true"] = perturb_axes(axes, sigma_deg, rng)
    return cfg
