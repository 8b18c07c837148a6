"""Per-azimuth observables of a sweep directory, in the rows the fabric fit
consumes (upstream ``fit_sweep.sweep_data``, sim/model/fit_sweep.py:356-504,
with its estimator helpers :225-275 and settings :44-352).

For every completed ``az{ddd}.npz`` of a sweep the trace gives

* ``e1``       -- first-echo envelope peak within +-2 us of the 200 mm round
                  trip (the fixed 100 mm disk of the upstream, ``D``);
* ``tof_us``   -- the leading edge of that echo: the first envelope sample at
                  or above 25 % of E1 before the peak, linearly interpolated
                  between samples (not the envelope peak);
* ``coda``     -- the frequency-diversity coda amplitude: the record gated to
                  the clean window (24-36 us) +-3 us with 2 us raised-cosine
                  ramps, split by a COLA Hann bank of ``FD_BANDS`` bands over
                  [0.5 f0, f0 ppw / 2], each band's mean squared envelope over
                  the clean window, combined by the geometric mean (linear
                  amplitude); with one band, the plain clean-window envelope
                  RMS;
* ``tail_db``  -- envelope RMS over 54-66 us relative to E1 (NaN when less
                  than 2 us of that window was recorded);
* ``cent_mhz`` -- magnitude-weighted spectral centroid of the trace within
                  +-2 us of the echo peak, over 0.25-2 f0;
* ``speckle_contrast`` -- std / mean of the squared envelope over the clean
                  window;

plus ``e1_db``.  Rows are sorted by azimuth and, with more than four of them,
the coda (in dB) and ToF are averaged over every azimuth within +-1.5 deg
modulo 180 deg.  Observables are cached per azimuth in ``obs_cache.npz``
(version-tagged; rewritten atomically when new azimuths were read), so a
refit after more azimuths arrive reads only the new files.

These are host computations over a few thousand samples per azimuth; the
estimator definitions have to match the upstream's for a fit on a sweep made
here to read the same numbers.
"""
from __future__ import annotations

import os

import numpy as np

from . import sweep_runner as SR

C_REF = 3850.0
D = 0.100                               # fit_sweep.py:42 (fixed, not the config's)
W_CLEAN = (24e-6, 36e-6)                # ladder.py:65
FD_BANDS = 3                            # fit_sweep.py:82
FD_SPAN = (0.5, None)                   # x f0; None = f0 ppw / 2 (grid Nyquist)
FD_GUARD = 3e-6
FD_RAMP = 2e-6
FD_COMBINE = "geo"
AZ_SMOOTH_DEG = 1.5
TAIL_W = (54e-6, 66e-6)
OBS_VER = "fd%d%s-le25-t4" % (FD_BANDS, FD_COMBINE) if FD_BANDS > 1 else "fd1-le25-t4"
LEADING_EDGE = 0.25

__all__ = ["fd_gate", "fd_coda", "trace_observables", "smooth_azimuths", "sweep_data",
           "OBS_VER", "W_CLEAN"]


def fd_gate(n, dt):
    """Time gate applied before the band split (fit_sweep.py:232-243): 1 on
    [W0 - guard, W1 + guard], a raised-cosine ramp of FD_RAMP to exactly 0
    on either side, 0 beyond."""
    t = np.arange(int(n)) * dt
    a, b = W_CLEAN[0] - FD_GUARD, W_CLEAN[1] + FD_GUARD
    g = np.where((t >= a) & (t <= b), 1.0, 0.0)
    up = (t > a - FD_RAMP) & (t < a)
    dn = (t > b) & (t < b + FD_RAMP)
    g[up] = 0.5 - 0.5 * np.cos(np.pi * (t[up] - a + FD_RAMP) / FD_RAMP)
    g[dn] = 0.5 + 0.5 * np.cos(np.pi * (t[dn] - b) / FD_RAMP)
    return g


def _band_masks(fr, f0, ppw, nb):
    """COLA Hann bank: nb bands with centres on equal splits of the span and
    full width twice the spacing, so adjacent masks sum to 1 between
    centres."""
    lo = FD_SPAN[0] * f0
    hi = FD_SPAN[1] * f0 if FD_SPAN[1] else 0.5 * f0 * ppw
    edges = np.linspace(lo, hi, nb + 1)
    w = edges[1] - edges[0]
    out = np.zeros((nb, len(fr)))
    for b in range(nb):
        x = fr - 0.5 * (edges[b] + edges[b + 1])
        inside = np.abs(x) <= w
        out[b, inside] = 0.5 + 0.5 * np.cos(np.pi * x[inside] / w)
    return out


def fd_coda(tr, dt, f0, ppw, nb=FD_BANDS):
    """Frequency-diversity coda amplitude (fit_sweep.py:246-275): gate, split
    into nb Hann bands, mean squared Hilbert envelope of each band over the
    clean window, geometric mean of those energies (FD_COMBINE "geo"), as a
    linear amplitude."""
    from scipy.signal import hilbert
    tr = np.asarray(tr, float)
    n = len(tr)
    spec = np.fft.rfft(tr * fd_gate(n, dt))
    masks = _band_masks(np.fft.rfftfreq(n, dt), f0, ppw, nb)
    i0, i1 = int(W_CLEAN[0] / dt), int(W_CLEAN[1] / dt)
    env = np.abs(hilbert(np.fft.irfft(spec[None, :] * masks, n, axis=1), axis=1))
    energy = np.maximum((env[:, i0:i1] ** 2).mean(axis=1), 1e-300)
    if FD_COMBINE == "geo":
        return float(10.0 ** (np.mean(10.0 * np.log10(energy)) / 20.0))
    if FD_COMBINE == "med":
        return float(np.sqrt(np.median(energy)))
    return float(np.sqrt(energy.mean()))


def _leading_edge(seg, ipk, level):
    """First sample at or above ``level`` up to the peak, with the crossing
    placed by linear interpolation from the sample before it; returns the
    position in samples relative to seg[0] (fit_sweep.py:415-429)."""
    hit = np.flatnonzero(seg[:ipk + 1] >= level)
    j = int(hit[0]) if len(hit) else int(ipk)
    frac = 0.0
    if j > 0 and seg[j] > seg[j - 1]:
        frac = (level - seg[j - 1]) / (seg[j] - seg[j - 1])
    return j - 1 + frac


def trace_observables(tr, dt, f0, ppw, nb=FD_BANDS):
    """The per-azimuth observables of one trace (fit_sweep.py:402-475)."""
    from scipy.signal import hilbert
    tr = np.asarray(tr, float).ravel()
    n = len(tr)
    fs = 1.0 / dt
    env = np.abs(hilbert(tr))
    w2 = int(2e-6 * fs)
    k0 = int(2 * D / C_REF * fs)
    a = max(k0 - w2, 0)
    seg = env[a:k0 + w2]
    ipk = int(np.argmax(seg))
    E1 = float(seg[ipk])
    tof = (a + _leading_edge(seg, ipk, LEADING_EDGE * E1)) / fs
    c0, c1 = int(W_CLEAN[0] * fs), int(W_CLEAN[1] * fs)
    clean = env[c0:c1] ** 2
    coda = fd_coda(tr, dt, f0, ppw, nb) if nb > 1 else float(np.sqrt(clean.mean()))
    t0, t1 = int(TAIL_W[0] * fs), min(int(TAIL_W[1] * fs), n)
    if t1 - t0 >= w2:
        tail_db = float(20 * np.log10(np.sqrt((env[t0:t1] ** 2).mean()) / max(E1, 1e-30)))
    else:
        tail_db = float("nan")
    kpk = a + ipk
    piece = tr[max(kpk - w2, 0):min(kpk + w2, n)]
    mag = np.abs(np.fft.rfft(piece))
    frq = np.fft.rfftfreq(len(piece), dt)
    band = (frq >= 0.25 * f0) & (frq <= 2.0 * f0)
    cent = float((frq[band] * mag[band]).sum() / max(mag[band].sum(), 1e-30) / 1e6)
    speckle = float(clean.std() / max(clean.mean(), 1e-30))
    return dict(coda=float(coda), e1=E1, tof_us=tof * 1e6, tail_db=tail_db, cent_mhz=cent,
                speckle_contrast=speckle)


def smooth_azimuths(rows, deg=AZ_SMOOTH_DEG):
    """Average coda (in dB) and ToF over the azimuths within +-deg modulo
    180 deg of each row (fit_sweep.py:488-503); rows sorted by "rot".  The
    averages use the unsmoothed values of every row."""
    if deg <= 0 or len(rows) <= 4:
        return rows
    folded = np.array([r["rot"] for r in rows], float) % 180.0
    coda_db = 20 * np.log10([r["coda"] for r in rows])
    tof = np.array([r["tof_us"] for r in rows], float)
    dist = np.abs(folded[:, None] - folded[None, :])
    near = np.minimum(dist, 180.0 - dist) <= deg
    for i, r in enumerate(rows):
        r["coda"] = 10 ** (coda_db[near[i]].mean() / 20)
        r["tof_us"] = float(tof[near[i]].mean())
    return rows


_CACHE_KEYS = ("coda", "e1", "tof_us", "tail_db", "cent_mhz", "speckle")
_ROW_KEYS = ("coda", "e1", "tof_us", "tail_db", "cent_mhz", "speckle_contrast")


def _read_cache(path, ver):
    if not os.path.exists(path):
        return {}
    z = np.load(path, allow_pickle=False)
    if str(z["ver"]) != ver:
        return {}
    cols = [z[k] for k in _CACHE_KEYS]
    return {int(az): tuple(float(c[i]) for c in cols) for i, az in enumerate(z["az"])}


def _write_cache(path, ver, cached):
    ks = sorted(cached)
    tmp = path[:-4] + ".tmp%d.npz" % os.getpid()
    try:
        np.savez(tmp, ver=ver, az=np.asarray(ks),
                 **{k: np.asarray([cached[a][i] for a in ks]) for i, k in enumerate(_CACHE_KEYS)})
        SR.atomic_replace(tmp, path, critical=False)
    except OSError:
        pass                                       # the cache is best-effort


def sweep_data(root, cfg):
    """Fit rows of the sweep ``cfg["name"]`` under ``root``: one dict per
    completed azimuth with rot, coda, e1, tof_us, tail_db, e1_db, cent_mhz,
    speckle_contrast (fit_sweep.py:356-504)."""
    d = SR.sweep_dir(root, cfg["name"])
    cache_p = os.path.join(d, "obs_cache.npz")
    nb = FD_BANDS
    ver = OBS_VER
    cached = _read_cache(cache_p, ver)
    fresh = False
    for az in SR.done_azimuths(root, cfg):
        if az in cached:
            continue
        z = np.load(os.path.join(d, "az%03d.npz" % az), allow_pickle=False)
        f0 = float(z["f0"]) if "f0" in z.files else cfg["f0_mhz"] * 1e6
        ppw = float(z["ppw"]) if "ppw" in z.files else cfg["ppw"]
        o = trace_observables(z["trace"], float(z["dt"]), f0, ppw, nb)
        cached[az] = tuple(o[k] for k in _ROW_KEYS)
        fresh = True
    if fresh:
        _write_cache(cache_p, ver, cached)
    rows = []
    for az in SR.done_azimuths(root, cfg):
        r = dict(rot=az, **dict(zip(_ROW_KEYS, cached[az])))
        r["e1_db"] = 20 * np.log10(max(r["e1"], 1e-30))
        rows.append(r)
    return smooth_azimuths(rows)
