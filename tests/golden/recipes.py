"""Inputs of the upstream float32 solver archives kept under tests/golden/reference/.

Each archive is the raw output of the upstream ``forward_fused_labels`` on a
closed-form medium.  This module rebuilds the solver inputs those runs used,
from the recipes in the upstream validation scripts (cited per function), so
that the archives can serve as golden vectors for the oracle and the HIP path.

Common settings (sim/validation/tilt_testbed.py:49-55): C_REF 3850 m/s,
F0 2 MHz, order 8, multistart coefficients at kh_max 2.0, L 36 mm,
sponge 20 cells (strength 0.09), no damp mask, Ricker with t0 = 1.2/F0,
nt = int(7 us / dt), source weight 1 at the centre cell (c, c, c), c = N // 2.
"""
from __future__ import annotations

import functools
import os

import numpy as np

from pulse_echo_azimuthal_simulation_amd import host, materials

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "reference")

C_REF, F0, ORDER, KHM = 3850.0, 2.0e6, 8, 2.0
L, DEPTH, SPONGE, NBIN, SUP = 0.036, 0.008, 20, 21, 4
RHO_ICE = 917.0


def load(name):
    return np.load(os.path.join(REF_DIR, name), allow_pickle=False)


@functools.lru_cache(maxsize=None)
def coeffs_kh2():
    return host.optimised_coeffs(ORDER, kh_max=KHM, multistart=True)


def grid(ppw):
    """h, N, centre c for a testbed at ``ppw`` (tilt_testbed.py:209-212)."""
    h = C_REF / F0 / ppw
    n = int(round(L / h))
    N = n + 2 * (SPONGE + 3)
    return h, N, N // 2


def grain_C(axis):
    """6x6 stiffness of an ice grain with c-axis ``axis`` (tilt_testbed.py:71-75)."""
    return materials.ti_stiffness_3d(np.asarray(axis, float), materials.ICE)


def iso_C(vp):
    """Isotropic ice with qP = vp and qS = 1950 m/s (rsg_green_probe.py:92-106)."""
    rho = RHO_ICE
    mu = rho * 1950.0 ** 2
    lam = rho * vp ** 2 - 2.0 * mu
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    for i in range(3):
        C[i, i] = lam + 2 * mu
    for i in range(3, 6):
        C[i, i] = mu
    return C


def green_receivers(N, h, c):
    """Arc, ray and neighbour receivers of rsg_green_probe.py:119-155."""
    R_ARC, ARC_STEP = 0.016, 5.0
    RAY_DIRS = (0.0, 30.0, 45.0, 54.7356, 60.0, 90.0)
    RAY_FRACS = (0.25, 0.5, 0.75, 1.0)
    p_src = np.array([(c + 0.5) * h] * 3)
    idx, rad = [], []

    def add(p):
        q = np.rint(np.asarray(p) / h - 0.5).astype(int)       # (x, y, z)
        pc = (q + 0.5) * h
        idx.append((int(q[2]), int(q[1]), int(q[0])))
        rad.append(float(np.linalg.norm(pc - p_src)))

    for a in np.arange(0.0, 90.0 + 1e-9, ARC_STEP):
        t = np.radians(a)
        add(p_src + R_ARC * np.array([np.cos(t), 0.0, np.sin(t)]))
    for a in RAY_DIRS:
        t = np.radians(a)
        for f in RAY_FRACS[:-1]:
            add(p_src + f * R_ARC * np.array([np.cos(t), 0.0, np.sin(t)]))
    for o in (-2, -1, 1, 2):
        add(p_src + np.array([R_ARC, 0.0, o * h]))
    for o in (-2, -1, 1, 2):
        add(p_src + np.array([R_ARC, o * h, 0.0]))
    return idx, np.asarray(rad)


def lattice_receivers(N, h, c):
    """Cells along +x and the +x+z diagonal (rsg_lattice_probe.py:68-83)."""
    R_MAX = 0.016
    idx = []
    for m in range(1, int(R_MAX / h) + 1):
        if c + m >= N - SPONGE:
            break
        idx.append((c, c, c + m))
    for m in range(1, int(R_MAX / (h * np.sqrt(2.0))) + 1):
        if c + m >= N - SPONGE:
            break
        idx.append((c + m, c, c + m))
    return idx


def homogeneous_case(C6, ppw):
    """Solver inputs of a homogeneous testbed run with medium ``C6``."""
    h, N, c = grid(ppw)
    Ct = np.tile(np.asarray(C6, np.float32).reshape(1, 36), (3, 1))
    rho_t = np.full(3, np.float32(RHO_ICE))
    lab = np.zeros((N, N, N), np.int32)
    return dict(h=h, N=N, c=c, lab=lab, Ct=Ct, rho_t=rho_t)


def green_case(med, ppw):
    """Inputs and expected traces of rsg_green_ssg_<med>_ppw<ppw>.npz
    (rsg_green_probe.py:157-196)."""
    ref = load("rsg_green_ssg_%s_ppw%d.npz" % (med, ppw))
    vp = float(ref["vp"])
    C6 = iso_C(vp) if med == "iso" else grain_C([1.0, 0.0, 0.0])
    case = homogeneous_case(C6, ppw)
    h, N, c = case["h"], case["N"], case["c"]
    idx, rad = green_receivers(N, h, c)
    groups = [([(c, c, c)], [1.0])] + [([q], [1.0]) for q in idx]
    case.update(src=[((c, c, c), 1.0)], groups=groups, rad=rad, ref=ref,
                expected=np.asarray(ref["rec"]), dt_ref=float(ref["dt"]))
    return case


def lattice_case(ppw=6):
    """Inputs and expected traces of rsg_lattice_ssg_ppw6.npz (rsg_lattice_probe.py:86-126)."""
    ref = load("rsg_lattice_ssg_ppw%d.npz" % ppw)
    case = homogeneous_case(iso_C(float(ref["vp"])), ppw)
    h, N, c = case["h"], case["N"], case["c"]
    idx = lattice_receivers(N, h, c)
    groups = [([(c, c, c)], [1.0])] + [([q], [1.0]) for q in idx]
    case.update(src=[((c, c, c), 1.0)], groups=groups, idx=idx, ref=ref,
                expected=np.asarray(ref["rec"]), dt_ref=float(ref["dt"]))
    return case


# --------------------------------------------------------------- bicrystal
def rot_y(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def frac_field(shape, h, nrm, x0, sup=SUP):
    """Volume fraction of the half-space n.(x - x0) < 0 per cell, by
    sup^3 supersampling (tilt_testbed.py:127-146)."""
    nz, ny, nx = shape
    o = (np.arange(sup) + 0.5) / sup - 0.5
    acc = np.zeros(shape, np.float32)
    for dz in o:
        for dy in o:
            for dx in o:
                z = (np.arange(nz) + 0.5 + dz) * h
                y = (np.arange(ny) + 0.5 + dy) * h
                x = (np.arange(nx) + 0.5 + dx) * h
                s = (nrm[2] * (z - x0[2])[:, None, None]
                     + nrm[1] * (y - x0[1])[None, :, None]
                     + nrm[0] * (x - x0[0])[None, None, :])
                acc += (s < 0).astype(np.float32)
    return acc / sup ** 3


def naive_tables(CA, CB):
    """NBIN-bin table of the 'naive' treatment (bin fraction of A > 0.5 -> A),
    with the row for label -1 prepended (tilt_testbed.py:149-180)."""
    f = np.linspace(0.0, 1.0, NBIN)
    out = np.zeros((NBIN, 36), np.float32)
    for i, fa in enumerate(f):
        out[i] = np.asarray(CA if fa > 0.5 else CB, np.float32).reshape(-1)
    out = np.vstack([out[:1], out])
    rho = np.full(NBIN + 1, RHO_ICE, np.float32)
    return out, rho


def twoway_geometry(theta_deg, ppw):
    """twoway_ref.py:79-98."""
    th = np.radians(theta_deg)
    R = rot_y(th)
    nrm = R @ np.array([1.0, 0.0, 0.0])
    h, N, c = grid(ppw)
    ctr = np.array([N * h / 2] * 3)
    d = round((ctr[0] + DEPTH) / h) * h - ctr[0]
    x0 = ctr + d * nrm
    p_src = np.array([(c + 0.5) * h] * 3)
    return dict(R=R, nrm=nrm, h=h, N=N, c=c, d=d, x0=x0, p_src=p_src)


@functools.lru_cache(maxsize=4)
def twoway_labels(theta, ppw):
    """Naive-treatment bin labels of the tilted plane (twoway_ref.py:157-158)."""
    g = twoway_geometry(theta, ppw)
    fr = frac_field((g["N"],) * 3, g["h"], g["nrm"], g["x0"])
    return np.rint(fr * (NBIN - 1)).astype(np.int32)


def twoway_case(theta, ppw, homog, psi_a=0.0, psi_b=51.0):
    """Inputs of the bicrystal (homog=False) or homogeneous twoway run; the
    homogeneous run reuses the bicrystal's dt (rsg_twoway_run.py:95-124,
    twoway_ref.py:141-175)."""
    g = twoway_geometry(theta, ppw)
    R, nrm, h, N, c = g["R"], g["nrm"], g["h"], g["N"], g["c"]
    ca = R @ np.array([np.cos(np.radians(psi_a)), 0.0, np.sin(np.radians(psi_a))])
    cb = R @ np.array([np.cos(np.radians(psi_b)), 0.0, np.sin(np.radians(psi_b))])
    CA, CB = grain_C(ca), grain_C(cb if not homog else ca)
    lab = twoway_labels(theta, ppw)
    Ct, rho_t = naive_tables(CA, CB)
    ref = load("rsg_twoway_ssg_ppw%d.npz" % ppw)
    key = "%d" % theta
    idx = [tuple(int(a) for a in q) for q in ref["idx_" + key]]
    groups = [([(c, c, c)], [1.0])] + [([q], [1.0]) for q in idx]
    expected = np.asarray(ref[("ho_" if homog else "bi_") + key])
    # dt of the bicrystal, computed from the bicrystal tables either way
    if homog:
        Cbi, rbi = naive_tables(CA, grain_C(cb))
        mats_dt, Ct_dt, rho_dt = lab + 1, Cbi, rbi
    else:
        mats_dt, Ct_dt, rho_dt = lab + 1, Ct, rho_t
    return dict(h=h, N=N, c=c, lab=lab, Ct=Ct, rho_t=rho_t, src=[((c, c, c), 1.0)],
                groups=groups, expected=expected, dt_ref=float(ref["dt_" + key]),
                dt_inputs=(mats_dt, Ct_dt, rho_dt), ref=ref)


# ------------------------------------------------------------ full disk
FULLDISK_FRACS = {"00": 0.0, "25": 0.25, "50": 0.5, "100": 1.0}


@functools.lru_cache(maxsize=1)
def fulldisk_setup():
    """Grid and shot of the upstream 2 MHz full-disk reference traces
    (sim/fw_checks/fw_reference_2mhz.py:52-78 with ladder.standard_build,
    sim/model/ladder.py:68-73, and coda_convergence.build_grid,
    sim/fw_checks/coda_convergence.py:40-53): the 100 mm x 35 mm disk of
    100 grains (size_cv 0.35, Watson kappa 3.93 about x, seed 7) at
    h = 3850 / 2 MHz / 6, padded by sponge 10 + 3 couplant cells; the
    legacy single-start kh 1.6 coefficients (the archives' dt); dt from the
    base fabric; couplant mask exp(-0.02 EDT); the +x rim probe of 6.35 mm."""
    from pulse_echo_azimuthal_simulation_amd.disk_specimen import DiskSpecimen, solver_grid
    h = C_REF / 2e6 / 6.0
    build = DiskSpecimen(diameter_m=0.100, thickness_m=0.035, n_grains=100, size_cv=0.35,
                         concentration=3.93, spatial_corr=0.0, fabric_axis=(1, 0, 0),
                         seed=7).build(h, adjacency=False)
    lab, nd, m = solver_grid(build, h, 10 + 3)
    nz, nx = lab.shape[0], lab.shape[1]
    co = host.optimised_coeffs(8, kh_max=1.6, multistart=False)
    Ct, rho_t = materials.material_tables(build["axes"])
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, co, safety=0.5)
    nt = int(2.6 * nd * h / C_REF / dt)
    dm = host.fluid_damp_mask(lab, 0.02)
    cz, cy, cx = nz // 2, nx // 2, nx // 2
    Rc = nd // 2
    ixp = next(cx + k - 1 for k in range(Rc - 1, 0, -1) if lab[cz, cy, cx + k] >= 0)
    er = max(int(6.35e-3 / 2 / h), 1)
    pts = [(cz + dz, cy + dy, ixp) for dy in range(-er, er + 1) for dz in range(-er, er + 1)
           if dy * dy + dz * dz <= er * er and lab[cz + dz, cy + dy, ixp] >= 0]
    return dict(build=build, lab=lab, nd=nd, h=h, coeffs=co, dt=dt, nt=nt, damp_mask=dm,
                pts=pts, w=1.0 / len(pts), wavelet=host.ricker(2e6, dt, nt))


def fulldisk_fabrics(base_axes):
    """The four fabric sets of ladder.standard_fabrics (sim/model/ladder.py:
    88-101): a fraction 0 / 0.25 / 0.5 / 1 of the grains given random axes
    from one default_rng(11) stream."""
    rng = np.random.default_rng(11)
    out = {}
    for tag, frac in FULLDISK_FRACS.items():
        ax = np.array(base_axes, float, copy=True)
        if frac > 0:
            kk = int(frac * len(ax))
            idx = rng.choice(len(ax), kk, replace=False)
            v = rng.normal(size=(kk, 3))
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            ax[idx] = v
        out[tag] = ax
    return out


def coda_level(x, dt):
    """The ladder's coda metric (sim/fw_checks/fw_reference_2mhz.py:38-46):
    envelope maximum over 22-42 us relative to the first-echo envelope E1
    (within 2 us of the 200 mm round trip), in dB; returns (dB, E1)."""
    from scipy.signal import hilbert
    t1 = 2 * 0.100 / C_REF
    e = np.abs(hilbert(x))
    k0 = int(t1 / dt)
    E1 = e[max(k0 - int(2e-6 / dt), 0):min(k0 + int(2e-6 / dt), len(e) - 1)].max()
    lo, hi = int(22e-6 / dt), min(int(42e-6 / dt), len(e))
    return 20 * np.log10(e[lo:hi].max() / E1 + 1e-30), E1


def validate_setup(f0, build, ppw=6.0, sponge=10, fluid_damp=0.02, coeffs=None):
    """sim/core/validate_labels.py:34-63: solver grid of a specimen build
    (pad sponge + 3), material tables, stable dt (safety 0.5), record length
    2.6 D / c, Ricker, couplant mask exp(-fluid_damp EDT) and the +x rim
    probe footprint of a 6.35 mm element."""
    from pulse_echo_azimuthal_simulation_amd.disk_specimen import solver_grid
    h = C_REF / f0 / ppw
    lab, nd, m = solver_grid(build, h, sponge + 3)
    nz, nx = lab.shape[0], lab.shape[1]
    co = host.optimised_coeffs(8) if coeffs is None else coeffs
    Ct, rho_t = materials.material_tables(build["axes"])
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, co, safety=0.5)
    nt = int(2.6 * nd * h / C_REF / dt)
    dm = host.fluid_damp_mask(lab, fluid_damp)
    cz, cy, cx = nz // 2, nx // 2, nx // 2
    Rc = nd // 2
    ixp = next(cx + k - 1 for k in range(Rc - 1, 0, -1) if lab[cz, cy, cx + k] >= 0)
    er = max(int(6.35e-3 / 2 / h), 1)
    pts = [(cz + dz, cy + dy, ixp) for dy in range(-er, er + 1) for dz in range(-er, er + 1)
           if dy * dy + dz * dz <= er * er and lab[cz + dz, cy + dy, ixp] >= 0]
    return dict(lab=lab, nd=nd, h=h, coeffs=co, dt=dt, nt=nt, damp_mask=dm, pts=pts,
                w=1.0 / len(pts), wavelet=host.ricker(f0, dt, nt), sponge=sponge)


def azimuth_setup(build, rot, h=C_REF / 2e6 / 6.0, sponge=10):
    """The upstream's rotated-disk reference traces sim/ref/fw_az{rot}.npz
    (sim/pipeline/azimuth_sweep.py:60-95): the build resampled by
    rotation_test.rotated_grid at rot degrees, the legacy single-start kh 1.6
    coefficients and the pre-'rigid2' axes convention axes @ R.T (c-axes
    turned by +rot; rotation_test.py:68-76 records the later fix) -- both
    fixed by the archives' stored dt, which this recipe reproduces -- then
    the same dt, record, Ricker, couplant mask and rim probe as the
    full-disk recipe."""
    from pulse_echo_azimuthal_simulation_amd import sweep
    lab, nd, m, _ = sweep.rotated_grid(build, h, sponge, rot)
    a = np.radians(rot)
    R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    axes = np.asarray(build["axes"], float) @ R.T
    nz, nx = lab.shape[0], lab.shape[1]
    co = host.optimised_coeffs(8, kh_max=1.6, multistart=False)
    Ct, rho_t = materials.material_tables(axes)
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, co, safety=0.5)
    nt = int(2.6 * nd * h / C_REF / dt)
    cz, cy, cx = nz // 2, nx // 2, nx // 2
    Rc = nd // 2
    ixp = next(cx + k - 1 for k in range(Rc - 1, 0, -1) if lab[cz, cy, cx + k] >= 0)
    er = max(int(6.35e-3 / 2 / h), 1)
    pts = [(cz + dz, cy + dy, ixp) for dy in range(-er, er + 1) for dz in range(-er, er + 1)
           if dy * dy + dz * dz <= er * er and lab[cz + dz, cy + dy, ixp] >= 0]
    return dict(lab=lab, axes=axes, nd=nd, h=h, coeffs=co, dt=dt, nt=nt,
                damp_mask=host.fluid_damp_mask(lab, 0.02), pts=pts, w=1.0 / len(pts),
                wavelet=host.ricker(2e6, dt, nt), sponge=sponge)


# ------------------------------------------------------------ tilt table
TILTS = (0.0, 15.0, 22.5, 30.0, 45.0)


def qp_speed_ice(psi):
    """qP phase speed of ice at angle psi from the c-axis: the transversely
    isotropic Christoffel root (the closed form of anisotropy.py:265-275)
    rho v^2 = (A + sqrt(B^2 + 4 (C13 + C44)^2 s^2 c^2)) / 2 with
    A = (C11 + C44) s^2 + (C33 + C44) c^2, B = (C11 - C44) s^2 - (C33 - C44) c^2."""
    m = materials.ICE
    s2, c2 = np.sin(psi) ** 2, np.cos(psi) ** 2
    A = (m["C11"] + m["C44"]) * s2 + (m["C33"] + m["C44"]) * c2
    B = (m["C11"] - m["C44"]) * s2 - (m["C33"] - m["C44"]) * c2
    return float(np.sqrt(0.5 * (A + np.sqrt(B * B + 4 * (m["C13"] + m["C44"]) ** 2 * s2 * c2))
                         / m["rho"]))


def echo_amp(tr_bi, tr_hom, dt, d, v, half_us=0.5):
    """Interface-echo amplitude of the A-B difference (tilt_testbed.py:
    237-243): envelope maximum of bi - hom within +-half_us of the two-way
    time 2 d / v plus the Ricker delay 1.2 / f0."""
    from scipy.signal import hilbert
    e = np.abs(hilbert(np.asarray(tr_bi, float) - np.asarray(tr_hom, float)))
    fs = 1.0 / dt
    i0 = int((2 * d / v + 1.2 / F0) * fs)
    w = int(half_us * 1e-6 * fs)
    return float(e[max(i0 - w, 0):i0 + w].max())


def tilt_case(theta, ppw, treatment, homog, psi_a=0.0, psi_b=51.0):
    """Inputs of one tilted-bicrystal run of the Table-2 testbed
    (tilt_testbed.py:183-226, sim/runs/tilt_table_run.py:74-98) built with
    the product's interface treatments: the twoway geometry, both c-axes
    turned with the plane, the near-half volume fraction binned to 21
    labels, the treatment's table (the homogeneous reference has grain A on
    both sides) and the bicrystal's own dt for both runs."""
    from pulse_echo_azimuthal_simulation_amd import treatments as T
    g = twoway_geometry(theta, ppw)
    R, nrm = g["R"], g["nrm"]
    ca = R @ np.array([np.cos(np.radians(psi_a)), 0.0, np.sin(np.radians(psi_a))])
    cb = R @ np.array([np.cos(np.radians(psi_b)), 0.0, np.sin(np.radians(psi_b))])
    CA, CB = grain_C(ca), grain_C(cb)
    lab = tilt_labels(theta, ppw)
    Ct, rho_t = T.interface_tables(CA, CA if homog else CB, nrm, treatment, nbin=NBIN)
    Cbi, rbi = T.interface_tables(CA, CB, nrm, treatment, nbin=NBIN)
    dt = host.safe_dt_labels(lab + 1, Cbi, rbi, g["h"], coeffs_kh2(), safety=0.5)
    nt = int(7.0e-6 / dt)
    c = g["c"]
    return dict(lab=lab, Ct=Ct, rho_t=rho_t, h=g["h"], dt=dt, nt=nt, d=g["d"],
                wavelet=host.ricker(F0, dt, nt), src=[((c, c, c), 1.0)],
                groups=[([(c, c, c)], np.array([1.0]))])


@functools.lru_cache(maxsize=5)
def tilt_labels(theta, ppw):
    from pulse_echo_azimuthal_simulation_amd import treatments as T
    g = twoway_geometry(theta, ppw)
    return T.fraction_labels(T.fraction_field((g["N"],) * 3, g["h"], g["nrm"], g["x0"]), NBIN)
