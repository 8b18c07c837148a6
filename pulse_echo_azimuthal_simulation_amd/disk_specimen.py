"""Laguerre-Voronoi ice disk with Watson c-axes: the upstream specimen builder.

Restates ``DiskSpecimen.build`` (sim/core/specimen.py:166-278) step for
step, with the same random-number consumption, so that a given seed gives
the upstream's grains: Monte-Carlo points and seed points drawn by
rejection in the disk, lognormal grain radii, Laguerre (power-diagram)
weights whose scale is bisected until the realised grain-size coefficient of
variation matches ``size_cv`` (over-seeding when the weights leave cells
empty), the label raster, empty cells dropped and labels renumbered, Watson
axes (sim/core/specimen.py:45-69), and optional neighbour smoothing of the
axes with the bulk eigenvalue restored (sim/core/specimen.py:72-112,
245-265).

Two labellers (sim/core/specimen.py:212-227, 281-351):

* ``labeller="all-seeds"``: the exact power-diagram argmin over every seed
  in float64, plane by plane, with the upstream GPU labeller's evaluation
  order ((dx^2 + dy^2) - w) + dz^2;
* ``labeller="shortlist"``: the upstream CPU path, the 12 Euclidean-nearest
  seeds (KD-tree) then the power argmin among them, (dx^2 + dy^2 + dz^2) - w.

They differ in a handful of near-tie cells per ten million.  Labels are
int16 in (x, y, z) order as upstream; ``solver_grid`` turns a build into a
padded (z, y, x) int32 solver grid (sim/fw_checks/coda_convergence.py:40-53).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .specimen import sample_watson

__all__ = ["DiskSpecimen", "orientation_tensor", "renormalise_fabric", "watson_eigenvalue",
           "kappa_for_eigenvalue", "solver_grid"]


def orientation_tensor(axes):
    """Eigenvalues (descending) and eigenvectors of the second-order
    orientation tensor <a a^T> (sim/core/specimen.py:138-147)."""
    a = np.asarray(axes, float)
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    T = (a[:, :, None] * a[:, None, :]).mean(axis=0)
    w, V = np.linalg.eigh(T)
    order = np.argsort(w)[::-1]
    return np.sort(w)[::-1], V[:, order]


def watson_eigenvalue(kappa, n=4001):
    """Largest orientation-tensor eigenvalue of a Watson distribution,
    int x^2 e^(k x^2) / int e^(k x^2) over [0, 1] by the trapezoid rule on n
    points (sim/core/specimen.py:114-123)."""
    x = np.linspace(0.0, 1.0, n)
    wgt = np.exp(kappa * (x ** 2 - 1.0))
    return float(np.trapezoid(x ** 2 * wgt, x) / np.trapezoid(wgt, x))


def kappa_for_eigenvalue(a1, lo=0.0, hi=400.0, tol=1e-6):
    """Bisection inverse of watson_eigenvalue (sim/core/specimen.py:126-135)."""
    a1 = float(np.clip(a1, 1.0 / 3.0 + 1e-9, 1.0 - 1e-6))
    if a1 <= 1.0 / 3.0 + 1e-9:
        return 0.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if watson_eigenvalue(mid) < a1:
            lo = mid
        else:
            hi = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)


def renormalise_fabric(axes, target_a1, n_iter=60):
    """Rescale every axis's angle from the dominant eigenvector until the
    bulk eigenvalue is back at target_a1, azimuths kept
    (sim/core/specimen.py:72-112)."""
    w, V = orientation_tensor(axes)
    if abs(w[0] - target_a1) < 1e-4:
        return axes
    e1 = V[:, 0]
    sgn = np.sign(axes @ e1)
    sgn[sgn == 0] = 1.0
    a = axes * sgn[:, None]
    cos = np.clip(a @ e1, -1.0, 1.0)
    psi = np.arccos(cos)
    perp = a - cos[:, None] * e1[None, :]
    pn = np.linalg.norm(perp, axis=1, keepdims=True)
    perp = np.divide(perp, np.maximum(pn, 1e-12), out=np.zeros_like(perp), where=pn > 1e-12)

    def rebuild(k):
        p2 = np.minimum(psi * k, np.pi / 2)
        return np.cos(p2)[:, None] * e1[None, :] + np.sin(p2)[:, None] * perp

    lo, hi = 0.0, 6.0
    for _ in range(n_iter):
        k = 0.5 * (lo + hi)
        if orientation_tensor(rebuild(k))[0][0] > target_a1:
            lo = k
        else:
            hi = k
    out = rebuild(0.5 * (lo + hi))
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def _seed_points(R, thickness, n, rng):
    """sim/core/specimen.py:280-286: batches of 3n box points kept inside
    the unit disk, the first n scaled to the disk."""
    pts = []
    while len(pts) < n:
        p = rng.uniform(-1, 1, (n * 3, 3))
        pts.extend(p[(p[:, 0] ** 2 + p[:, 1] ** 2) <= 1.0].tolist())
    s = np.array(pts[:n])
    s[:, :2] *= R
    s[:, 2] *= thickness / 2
    return s


def _mc_points(R, thickness, n, rng):
    """sim/core/specimen.py:288-294: one batch of 1.4 n box points."""
    p = rng.uniform(-1, 1, (int(n * 1.4), 3))
    p = p[(p[:, 0] ** 2 + p[:, 1] ** 2) <= 1.0][:n]
    p[:, :2] *= R
    p[:, 2] *= thickness / 2
    return p


def _assign(P, seeds, weights, k=12, chunk=1 << 20):
    """Power-diagram assignment among the k Euclidean-nearest seeds
    (sim/core/specimen.py:336-351), int16."""
    from scipy.spatial import cKDTree
    k = min(k, len(seeds))
    tree = cKDTree(seeds)
    out = np.empty(len(P), np.int16)
    for a in range(0, len(P), chunk):
        Q = P[a:a + chunk]
        _, idx = tree.query(Q, k=k, workers=-1)
        idx = np.atleast_2d(idx.T).T
        d2 = ((Q[:, None, :] - seeds[idx]) ** 2).sum(-1) - weights[idx]
        out[a:a + chunk] = idx[np.arange(len(Q)), d2.argmin(1)].astype(np.int16)
    return out


def _label_all_seeds(x, z, R, seeds, weights, chunk=1 << 16):
    """Exact power argmin over all seeds, float64, (x, y, z) raster
    (sim/core/specimen.py:298-334: ((dx^2 + dy^2) - w) + dz^2)."""
    nx, nz = len(x), len(z)
    X, Y = np.meshgrid(x, x, indexing="ij")
    inside = (X ** 2 + Y ** 2) <= R ** 2
    Pxy = np.stack([X[inside], Y[inside]], axis=1)
    raw = np.full((nx, nx, nz), -1, np.int32)
    lab_in = np.empty((len(Pxy), nz), np.int32)
    for a in range(0, len(Pxy), chunk):
        Q = Pxy[a:a + chunk]
        d_xy = ((Q[:, None, :] - seeds[None, :, :2]) ** 2).sum(-1) - weights[None, :]
        for iz in range(nz):
            d2 = d_xy + (z[iz] - seeds[:, 2])[None, :] ** 2
            lab_in[a:a + chunk, iz] = d2.argmin(axis=1)
    raw[inside] = lab_in
    return raw


def _adjacency(labels):
    """Grain neighbour sets from face contacts (sim/core/specimen.py:353-363)."""
    adj = {}
    for ax in range(3):
        a = np.take(labels, np.arange(labels.shape[ax] - 1), axis=ax)
        b = np.take(labels, np.arange(1, labels.shape[ax]), axis=ax)
        m = (a != b) & (a >= 0) & (b >= 0)
        for u, v in np.unique(np.stack([a[m], b[m]], 1), axis=0):
            adj.setdefault(int(u), set()).add(int(v))
            adj.setdefault(int(v), set()).add(int(u))
    return adj


@dataclass
class DiskSpecimen:
    """The upstream disk specimen (sim/core/specimen.py:150-164 fields)."""
    diameter_m: float = 100e-3
    thickness_m: float = 35e-3
    n_grains: int = 400
    size_cv: float = 0.35
    concentration: float = 4.0
    spatial_corr: float = 0.0
    fabric_axis: tuple = (0.0, 0.0, 1.0)
    seed: int = 7

    def build(self, h, labeller="all-seeds", adjacency=True):
        """Grains on a grid of spacing h: dict(labels (nx, nx, nz) int16,
        axes, seeds, weights, adjacency, h, inside, shape)."""
        rng = np.random.default_rng(self.seed)
        R = self.diameter_m / 2
        nx = int(np.ceil(self.diameter_m / h))
        nz = int(np.ceil(self.thickness_m / h))
        mc = _mc_points(R, self.thickness_m, 400000, rng)
        base = (R / self.n_grains ** (1 / 3)) ** 2
        sigma = np.sqrt(np.log(1 + self.size_cv ** 2))
        n_seed = self.n_grains
        for outer in range(4):
            seeds = _seed_points(R, self.thickness_m, n_seed, rng)
            radii = rng.lognormal(-0.5 * sigma ** 2, sigma, n_seed)
            wshape = base * (radii ** 2 - 1.0)
            lo, hi, scale = 0.0, 8.0, 0.0
            for _ in range(20):
                trial = 0.5 * (lo + hi)
                cnt = np.bincount(_assign(mc, seeds, trial * wshape),
                                  minlength=n_seed).astype(float)
                live = cnt[cnt > 0]
                cv = (live ** (1 / 3)).std() / (live ** (1 / 3)).mean()
                if cv < self.size_cv:
                    lo = trial
                else:
                    hi = trial
                scale = trial
            n_live = int((np.bincount(_assign(mc, seeds, scale * wshape),
                                      minlength=n_seed) > 0).sum())
            if n_live >= self.n_grains or outer == 3:
                break
            n_seed = int(np.ceil(n_seed * self.n_grains / max(n_live, 1)))
        weights = scale * wshape
        self._weight_scale, self._n_seed = scale, n_seed

        x = (np.arange(nx) - (nx - 1) / 2) * h
        z = (np.arange(nz) - (nz - 1) / 2) * h
        if labeller == "all-seeds":
            raw = _label_all_seeds(x, z, R, seeds, weights)
        elif labeller == "shortlist":
            X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
            inside = (X ** 2 + Y ** 2) <= R ** 2
            raw = np.full(X.shape, -1, np.int32)
            P = np.stack([X[inside], Y[inside], Z[inside]], axis=1)
            raw[inside] = _assign(P, seeds, weights).astype(np.int32)
        else:
            raise ValueError("labeller must be 'all-seeds' or 'shortlist'")

        inside = raw >= 0
        keep = np.unique(raw[inside])
        remap = np.full(len(seeds), -1, np.int32)
        remap[keep] = np.arange(len(keep))
        labels = np.full(raw.shape, -1, np.int16)
        labels[inside] = remap[raw[inside]].astype(np.int16)
        seeds = seeds[keep]
        n_real = len(keep)

        axes = sample_watson(n_real, self.fabric_axis, self.concentration, rng)
        adj = _adjacency(labels) if (adjacency or self.spatial_corr > 0) else None
        if self.spatial_corr > 0:
            target_a1 = orientation_tensor(axes)[0][0]
            for _ in range(3):
                nb = np.zeros_like(axes)
                for g, ns in adj.items():
                    if ns:
                        v = axes[list(ns)]
                        v = v * np.sign(v @ axes[g])[:, None]
                        nb[g] = v.mean(axis=0)
                    else:
                        nb[g] = axes[g]
                axes = (1 - self.spatial_corr) * axes + self.spatial_corr * nb
                axes /= np.linalg.norm(axes, axis=1, keepdims=True)
            axes = renormalise_fabric(axes, target_a1)
        return dict(labels=labels, axes=axes, seeds=seeds, weights=weights[keep],
                    adjacency=adj, h=h, inside=inside, shape=labels.shape)


def solver_grid(build, h, pad):
    """Nearest-neighbour re-grid of a build at spacing h into a (z, y, x)
    int32 solver grid with a ``pad``-cell couplant frame (-1) on every side
    (sim/fw_checks/coda_convergence.py:40-53, pad = sponge + 3).  Returns
    (lab, nd, pad)."""
    labels, h_spec = build["labels"], build["h"]
    nx_s, _, nz_s = labels.shape
    D = nx_s * h_spec
    nd = int(round(D / h))
    ndz = int(round(nz_s * h_spec / h))
    gx = np.clip((np.arange(nd) + 0.5) * h / h_spec, 0, nx_s - 1).astype(int)
    gz = np.clip((np.arange(ndz) + 0.5) * h / h_spec, 0, nz_s - 1).astype(int)
    core = labels[np.ix_(gx, gx, gz)].transpose(2, 1, 0).astype(np.int32)
    nz, n = ndz + 2 * pad, nd + 2 * pad
    lab = np.full((nz, n, n), -1, np.int32)
    lab[pad:pad + ndz, pad:pad + nd, pad:pad + nd] = core
    return lab, nd, pad
