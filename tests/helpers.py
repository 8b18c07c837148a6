"""Shared helpers of the test suite: fixture inputs and synthetic media."""
from __future__ import annotations

import numpy as np

from pulse_echo_azimuthal_simulation_amd import host, materials
from tests.golden import recipes as R


def csr(groups, shape):
    ptr, idx, w = [0], [], []
    for pts, wts in groups:
        idx += [int(np.ravel_multi_index(tuple(q), shape)) for q in pts]
        w += [float(x) for x in np.ravel(wts)]
        ptr.append(len(idx))
    return np.array(ptr, np.int32), np.array(idx, np.int64), np.array(w)


def fixture_inputs(case):
    """Host arrays of a golden-archive run: dict usable by the oracle and the solver."""
    N = case["N"]
    shape = (N, N, N)
    co = R.coeffs_kh2()
    dt = case["dt_ref"]
    nt = int(7.0e-6 / dt)
    src_idx = np.array([np.ravel_multi_index(p, shape) for p, _ in case["src"]], np.int64)
    src_w = np.array([w for _, w in case["src"]], float)
    return dict(mats=case["lab"] + 1, Ct=case["Ct"], rho_t=case["rho_t"],
                damp=host.sponge_profile(shape, R.SPONGE, 0.09), coeffs=co, h=case["h"],
                dt=dt, nt=nt, wavelet=host.ricker(R.F0, dt, nt), src=(src_idx, src_w),
                rec=csr(case["groups"], shape), shape=shape)


def random_axes(n, rng):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_medium(shape, n_grain=11, seed=0, order=8, sponge=3, fluid_pad=1,
                  n_src=5, n_rec_groups=3, with_mask=True):
    """A small heterogeneous polycrystal with couplant, damping and footprints.

    Voronoi grains from seeded centres fill an inner box, a couplant layer
    (label -1) surrounds it, the damp volume is a sponge times an
    exponential fluid mask, and source / receiver footprints sit on ice
    cells where damp == 1 (as the solver requires).
    """
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([zz, yy, xx], -1).reshape(-1, 3).astype(float)
    seeds = rng.uniform(0, 1, (n_grain, 3)) * np.array(shape)
    d2 = ((cells[:, None, :] - seeds[None]) ** 2).sum(-1)
    lab = d2.argmin(1).reshape(shape).astype(np.int32)
    p = fluid_pad
    inner = np.zeros(shape, bool)
    sl = tuple(slice(p, n - p) if n > 2 * p + 2 else slice(None) for n in shape)
    inner[sl] = True
    lab[~inner] = -1
    axes = random_axes(n_grain, rng)
    Ct, rho_t = materials.material_tables(axes)
    co = host.optimised_coeffs(order) if order > 2 else np.array([1.0])
    damp = host.sponge_profile(shape, sponge, 0.09)
    if with_mask:
        damp = damp * host.fluid_damp_mask(lab, 0.02)
    ok = np.argwhere((damp >= 0.999) & (lab >= 0))
    pick = rng.choice(len(ok), size=min(len(ok), n_src + 4 * n_rec_groups), replace=False)
    pts = [tuple(int(a) for a in ok[k]) for k in pick]
    src_pts = [(q, float(rng.uniform(0.2, 1.0))) for q in pts[:n_src]]
    groups = []
    rest = pts[n_src:]
    for gi in range(n_rec_groups):
        g = rest[4 * gi:4 * gi + 4] or pts[:1]
        groups.append((g, rng.uniform(0.1, 1.0, len(g))))
    h = 1e-4
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, co, safety=0.5)
    nt = 40
    wav = host.ricker(2.0e6, dt, nt, t0=4 * dt)
    src_idx = np.array([np.ravel_multi_index(q, shape) for q, _ in src_pts], np.int64)
    src_w = np.array([w for _, w in src_pts])
    return dict(lab=lab, axes=axes, mats=lab + 1, Ct=Ct, rho_t=rho_t, damp=damp, coeffs=co,
                h=h, dt=dt, nt=nt, wavelet=wav, src=(src_idx, src_w),
                rec=csr(groups, shape), shape=shape, src_pts=src_pts, groups=groups)


def rel_l2(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def random_state(torch, sh, osh, seed):
    """The same random v, s (and both buffers of every CPML memory variable)
    on a DeviceShot and its OracleShot: every cell, tile, z chunk and band
    path is then at work from the first step."""
    rng = np.random.default_rng(seed)
    for dev, hst in zip(sh.v + sh.s, osh.v + osh.s):
        a = rng.standard_normal(hst.shape, dtype=np.float32)
        hst[...] = a
        dev.view(-1).copy_(torch.from_numpy(a.ravel()))
    for dev, hst in zip(sh.psi, osh.psi):
        a = (0.1 * rng.standard_normal(hst.shape)).astype(np.float32)
        hst[...] = a
        dev.view(-1).copy_(torch.from_numpy(a.ravel()))
    torch.cuda.synchronize()


def compare_fields(torch, sh, osh, tr, otr, trace_tol=1e-6):
    """Fields and memory variables bit for bit, traces within trace_tol rel-L2.

    From a random state a pressure trace sums hundreds of random-signed
    terms, so its relative rounding error is amplified by the cancellation;
    the bar is then trace_tol times the conditioning of the last row's sum,
    sum |w p| / |sum w p| (computed in float64 from the final stress, which
    the velocity half-step leaves as it was when the last row was recorded).
    """
    torch.cuda.synchronize()
    v, s = sh.fields()
    for k in range(3):
        assert np.array_equal(v[k], osh.v[k]), "v%d" % k
    for k in range(6):
        assert np.array_equal(s[k], osh.s[k]), "s%d" % (k + 1)
    for a, p in enumerate(sh.psi):
        assert np.array_equal(p.cpu().numpy().reshape(osh.psi[a].shape), osh.psi[a]), \
            "psi axis %d" % a
    assert np.isfinite(otr).all() and np.abs(otr).max() > 0
    cond = 1.0
    if sh.desc.record_mode == 0:
        ptr = sh.rec_ptr.cpu().numpy()
        idx = sh.rec_idx.cpu().numpy()
        w = sh.rec_w.cpu().numpy().astype(np.float64)
        p = -(osh.s[0].ravel()[idx].astype(np.float64) + osh.s[1].ravel()[idx]
              + osh.s[2].ravel()[idx]) / 3.0
        for g in range(len(ptr) - 1):
            terms = w[ptr[g]:ptr[g + 1]] * p[ptr[g]:ptr[g + 1]]
            tot = abs(terms.sum())
            if tot > 0:
                cond = max(cond, float(np.abs(terms).sum() / tot))
    assert rel_l2(tr.cpu().numpy(), otr) <= trace_tol * cond
