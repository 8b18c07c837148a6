"""Full-waveform fabric inversion over an azimuthal pulse-echo sweep.

The upstream fits its fabric (mean c-axis azimuth and Watson concentration)
by Nelder-Mead on trace observables through a fast forward model
(sim/model/fit_fabric.py:224-320, fit_sweep.py:356-560).  Here the fabric
parameters are fitted to the full traces of every azimuth by a
quasi-Newton method whose gradient is exact: the discrete adjoint of the
GPU solver gives dJ/dC for every table row, and the chain to the fabric
parameters runs through a deterministic reparameterisation of the Watson
sample.

Reparameterisation.  Grain g keeps fixed latent numbers (u_g, psi_g) drawn
once from U(0, 1) x U(0, 2 pi).  Its c-axis for the fabric (alpha, kappa)
(cluster axis mu = (cos alpha, sin alpha, 0) in the disk plane,
concentration kappa; kappa > 0 clusters, kappa < 0 girdles) is

    a_g = t_g mu + sqrt(1 - t_g^2) (cos psi_g e_z + sin psi_g (mu x e_z))

with t_g the u_g-quantile of |mu . a|, whose density on [0, 1] under the
Watson law p(a) ~ exp(kappa (mu . a)^2) is exp(kappa t^2) / Z(kappa): so the
a_g are an exact Watson sample for every (alpha, kappa), and smooth in both.
The quantile is found by safeguarded Newton iteration on the CDF, itself
evaluated by 64-point Gauss-Legendre quadrature (exact to rounding for
these smooth integrands).  The eigenvalue reported beside kappa is
E[(mu . a)^2].

Per azimuth the solver grid, footprint, time step and step count are fixed
(the time step is the stable step of a fabric whose every grain has its
largest diagonal stiffness C33 along a grid axis, which bounds every
fabric), so only the stiffness table changes with the fabric.  The misfit
J = 1/2 sum over azimuths, steps and receivers of (y - d)^2 is accumulated
in float64; dJ/d(alpha, kappa) = sum over azimuths of <dJ/dC_az, dC_az/dp>
with dC_az/dp by central differences of the float64 host table map.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import materials

_GLX, _GLW = np.polynomial.legendre.leggauss(64)
_GLX = 0.5 * (_GLX + 1.0)           # nodes on [0, 1]
_GLW = 0.5 * _GLW


def _partial(t, kappa):
    """int_0^t exp(kappa s^2) ds for an array t in [0, 1]."""
    t = np.asarray(t, float)
    s = t[..., None] * _GLX
    return t * (np.exp(kappa * s * s) @ _GLW)


def watson_cos(u, kappa):
    """t with P(|mu . a| <= t) = u under Watson(kappa), for an array u."""
    u = np.clip(np.asarray(u, float), 0.0, 1.0)
    Z = float(_partial(np.array([1.0]), kappa)[0])
    lo = np.zeros_like(u)
    hi = np.ones_like(u)
    t = u.copy()
    for _ in range(60):
        F = _partial(t, kappa) / Z - u
        lo = np.where(F < 0, t, lo)
        hi = np.where(F >= 0, t, hi)
        dF = np.exp(kappa * t * t) / Z
        tn = t - F / dF
        bad = (tn <= lo) | (tn >= hi) | ~np.isfinite(tn)
        tn = np.where(bad, 0.5 * (lo + hi), tn)
        if np.max(np.abs(tn - t)) < 1e-16:
            t = tn
            break
        t = tn
    return t


def watson_eigenvalue(kappa) -> float:
    """E[(mu . a)^2] under Watson(kappa)."""
    Z = float(_partial(np.array([1.0]), kappa)[0])
    return float((_GLX ** 2 * np.exp(kappa * _GLX ** 2)) @ _GLW / Z)


def watson_axes(alpha, kappa, u, psi):
    """(G, 3) c-axes of the reparameterised Watson sample (module docstring)."""
    t = watson_cos(u, kappa)
    st = np.sqrt(np.maximum(0.0, 1.0 - t * t))
    mu = np.array([math.cos(alpha), math.sin(alpha), 0.0])
    ez = np.array([0.0, 0.0, 1.0])
    e2 = np.cross(mu, ez)
    psi = np.asarray(psi, float)
    return (t[:, None] * mu + st[:, None] * (np.cos(psi)[:, None] * ez
                                              + np.sin(psi)[:, None] * e2))


def stiffness_tables64(axes, c_couplant=1500.0, rho_couplant=300.0, material=materials.ICE):
    """float64 (G+1, 36) table of grains with c-axes ``axes`` (G, 3), all grains
    at once: the rotation of materials.material_tables (Rodrigues, taking +z
    onto the axis; the antiparallel case diag(1, -1, -1)) applied through the
    Bond matrix, C' = M C M^T, instead of the per-grain 4-tensor einsum.  The
    two agree to rounding: 2.5e-16 of the largest entry at the median grain,
    up to 1e-13 for axes near -z, where the 1/(1 + c) of the Rodrigues
    rotation amplifies the rounding of both.  This one is about a hundred
    times faster on 3000 grains, which matters because the inversion builds
    five tables per azimuth and evaluation."""
    a = np.asarray(axes, float).reshape(-1, 3)
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    G = len(a)
    v0, v1, c = -a[:, 1], a[:, 0], a[:, 2]
    K = np.zeros((G, 3, 3))
    K[:, 0, 2], K[:, 1, 2] = v1, -v0
    K[:, 2, 0], K[:, 2, 1] = -v1, v0
    vn = np.sqrt(v0 * v0 + v1 * v1)
    deg = vn < 1e-12
    den = np.where(deg, 1.0, 1.0 + c)
    R = np.eye(3) + K + (K @ K) / den[:, None, None]
    R[deg] = np.eye(3)
    R[deg & (c <= 0)] = np.diag([1.0, -1.0, -1.0])
    pairs = materials._VOIGT_PAIRS
    M = np.empty((G, 6, 6))
    for x, (i, j) in enumerate(pairs):
        for y, (k, l) in enumerate(pairs):
            M[:, x, y] = (R[:, i, k] * R[:, j, k] if y < 3
                          else R[:, i, k] * R[:, j, l] + R[:, i, l] * R[:, j, k])
    C0 = materials.ti_stiffness_6(**material)
    out = np.empty((G + 1, 36))
    out[0] = materials.fluid_stiffness_6(c_couplant, rho_couplant).reshape(36)
    out[1:] = (M @ C0 @ M.transpose(0, 2, 1)).reshape(G, 36)
    return out


def latents(n_grains, seed):
    """Fixed per-grain latent numbers (u, psi)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, n_grains), rng.uniform(0.0, 2.0 * math.pi, n_grains)


def _rz(az_deg):
    a = math.radians(az_deg)
    ca, sa = math.cos(a), math.sin(a)
    return np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])


class FabricInversion:
    """Forward and gradient of the sweep misfit in the fabric parameters.

    ``build``: a disk in the upstream build layout (sweep.disk_build);
    ``azimuths`` in degrees; the shots use the sweep preparation of
    ``sweep.GpuAzimuthPrep`` (rim transducer, sponge, couplant mask).
    """

    def __init__(self, build, azimuths, h, coeffs, f0, nt, lat, sponge=10,
                 checkpoint_every=100, precision="f32", element_d_mm=6.35, history_gb=None,
                 hist="strain"):
        from . import adjoint, sweep
        self.azimuths = [float(a) for a in azimuths]
        self.lat = lat
        self.nt = int(nt)
        self.checkpoint_every = int(checkpoint_every)
        # first-pass history kept in HBM per azimuth (segments that do not
        # fit are recomputed from their checkpoints); None: the free device
        # memory less 16 GiB at each evaluation
        self.history_gb = history_gb
        prep = sweep.GpuAzimuthPrep(build, h, coeffs, f0, sponge=sponge,
                                    element_d_mm=element_d_mm)
        G = len(lat[0])
        zaxes = np.tile([0.0, 0.0, 1.0], (G, 1))
        self.shots, self.wavs, self.rot, self.adj = [], [], [], []
        for az in self.azimuths:
            R = _rz(az)
            bound = materials.material_tables(zaxes @ R)
            sh, wav, _, _ = prep.shot(az, precision=precision, nt=self.nt, tables=bound)
            # every stiffness term active: the bound tables have zeros that
            # a general fabric fills
            sh.desc.act_mask = (1 << 36) - 1
            self.shots.append(sh)
            self.wavs.append(wav)
            self.rot.append(R)
            # the gradient reduced on the device onto the two fabric
            # parameters (projection rows dC/dalpha, dC/dkappa per grain,
            # set at each evaluation); strain history (the reverse pass reads
            # the strains the forward stored instead of recomputing them)
            self.adj.append(adjoint.DeviceAdjoint(sh, n_proj=2,
                                                  proj=np.zeros((G + 1, 2, 36)), hist=hist))
        self.torch = self.shots[0].torch
        self.evals = []
        self._axes = {}

    def tables64(self, p, k):
        """float64 (n_mat, 36) table of azimuth k for fabric p = (alpha, kappa).
        The grain axes depend on p only: they are kept for the last few p, as
        an evaluation asks for the same five fabrics at every azimuth."""
        key = (float(p[0]), float(p[1]))
        axes = self._axes.get(key)
        if axes is None:
            if len(self._axes) >= 8:
                self._axes.clear()
            axes = self._axes[key] = watson_axes(key[0], key[1], *self.lat)
        return stiffness_tables64(axes @ self.rot[k])

    def _set(self, k, C):
        sh = self.shots[k]
        sh.Ct.copy_(self.torch.as_tensor(np.asarray(C).reshape(sh.Ct.shape), device=sh.device))

    def traces(self, p):
        """Device traces of every azimuth for fabric p."""
        out = []
        for k, sh in enumerate(self.shots):
            self._set(k, self.tables64(p, k))
            sh.reset()
            tr = sh.new_traces(self.nt)
            sh.run(self.wavs[k], 0, self.nt, tr)
            out.append(tr)
        return out

    def misfit_gradient(self, p, data, h_alpha=1e-6, h_kappa=1e-6, log=None):
        """(J, dJ/d(alpha, kappa)) in float64 for the fabric p; ``log``, if
        given, receives a line per azimuth."""
        from . import adjoint
        torch = self.torch
        J = 0.0
        g = np.zeros(2)
        steps = (np.array([h_alpha, 0.0]), np.array([0.0, h_kappa * max(1.0, abs(p[1]))]))
        for k, sh in enumerate(self.shots):
            t0 = time.perf_counter()
            self._set(k, self.tables64(p, k))
            sh.reset()
            d = data[k]

            def resid(tr, d=d):
                r = tr - d
                return 0.5 * float((r.double() ** 2).sum()), r

            if self.history_gb is None:
                # (the caching allocator's free blocks, the previous
                # azimuth's history among them, count as free)
                free = torch.cuda.mem_get_info(sh.device)[0] + \
                    torch.cuda.memory_reserved(sh.device) - torch.cuda.memory_allocated(sh.device)
                hist_bytes = max(0, free - (16 << 30))
            else:
                hist_bytes = int(self.history_gb * (1 << 30))
            # dC/dp by central differences of the float64 table map, the
            # projection rows of the device reduction
            dC = np.stack([(self.tables64(np.asarray(p, float) + dp, k)
                            - self.tables64(np.asarray(p, float) - dp, k)) / (2.0 * dp[j])
                           for j, dp in enumerate(steps)], axis=1)
            self.adj[k].proj.copy_(torch.as_tensor(dC, dtype=self.adj[k].proj.dtype))
            t1 = time.perf_counter()
            Jk, gp, _ = adjoint.misfit_gradient(sh, self.wavs[k], self.nt, resid,
                                                checkpoint_every=self.checkpoint_every,
                                                adjoint=self.adj[k], history_bytes=hist_bytes)
            g += gp.sum(dim=0).cpu().numpy()
            t2 = time.perf_counter()
            J += Jk
            if log:
                log("azimuth %d/%d (%.1f deg): J_k %.4e; tables %.2f s, forward + adjoint %.2f s"
                    % (k + 1, len(self.shots), self.azimuths[k], Jk, t1 - t0, t2 - t1))
        torch.cuda.synchronize()
        return J, g

    def fit(self, p0, data, max_iter=30, bounds=((None, None), (-20.0, 30.0)), log=None):
        """L-BFGS-B from p0; returns (p, J, history)."""
        from scipy.optimize import minimize
        hist = []
        t0 = time.perf_counter()

        def fun(x):
            J, g = self.misfit_gradient(x, data)
            hist.append(dict(alpha=float(x[0]), kappa=float(x[1]), J=J,
                             g=[float(g[0]), float(g[1])], t=time.perf_counter() - t0))
            if log:
                log("eval %d: alpha %.3f deg, kappa %.3f, J %.6e"
                    % (len(hist), math.degrees(x[0]), x[1], J))
            return J, g

        res = minimize(fun, np.asarray(p0, float), jac=True, method="L-BFGS-B",
                       bounds=bounds, options=dict(maxiter=int(max_iter), ftol=1e-14,
                                                   gtol=1e-12))
        return res.x, float(res.fun), hist, res
