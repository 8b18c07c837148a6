"""fit_sweep.sweep_data and its estimators (sim/model/fit_sweep.py:225-504).

Pinned by the upstream's stored numbers where it has them -- the clean-window
level (envelope RMS over 24-36 us re E1, dB) stored with each rotated-disk
trace (sim/ref/fw_az*.npz, sim/pipeline/azimuth_sweep.py:92-98) and the E1
stored with each fabric trace (sim/ref/fw_fabric*.npz,
sim/fw_checks/fw_reference_2mhz.py:38-46) -- by the upstream's second,
independently written leading-edge pick (sim/model/fit_fabric.py:146-178)
on the same traces, and by closed forms of the band bank, gate, combiner,
centroid and azimuth folding.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from pulse_echo_azimuthal_simulation_amd import fit_sweep as FS
from pulse_echo_azimuthal_simulation_amd import sweep_runner as SR

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "reference")
AZ = (17, 31, 47, 62, 76, 90)
FAB = ("00", "25", "50", "100")


def _load(name):
    return np.load(os.path.join(REF, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("rot", AZ)
def test_clean_level_matches_stored(rot):
    """E1 and the FD1 clean-window RMS give the stored level_db."""
    z = _load("fw_az%02d" % rot)
    o = FS.trace_observables(z["trace"], float(z["dt"]), 2e6, 6.0, nb=1)
    lvl = 20 * np.log10(o["coda"] / o["e1"] + 1e-30)
    assert abs(lvl - float(z["level_db"])) < 1e-9


@pytest.mark.parametrize("tag", FAB)
def test_e1_matches_stored(tag):
    z = _load("fw_fabric%s" % tag)
    o = FS.trace_observables(z["trace"], float(z["dt"]), 2e6, 6.0)
    assert abs(o["e1"] / float(z["E1"]) - 1) < 1e-12


def _tof_walk_back(tr, dt, f0=2e6):
    """The upstream's other leading-edge pick (fit_fabric.py:151-176): walk
    back from the peak to the last sample at or below 25 % of E1, then
    interpolate forward; it reports the time minus the 1.2/f0 Ricker
    delay."""
    from scipy.signal import hilbert
    e = np.abs(hilbert(np.asarray(tr, float)))
    fs = 1.0 / dt
    k0 = int(2 * 0.100 / 3850.0 * fs)
    a0 = max(k0 - int(2e-6 * fs), 0)
    seg = e[a0:k0 + int(2e-6 * fs)]
    thr = 0.25 * seg.max()
    j = int(np.argmax(seg))
    while j > 0 and seg[j] > thr:
        j -= 1
    frac = (thr - seg[j]) / (seg[j + 1] - seg[j]) if seg[j + 1] > seg[j] else 0.0
    return ((a0 + j + frac) / fs - 1.2 / f0) * 1e6


@pytest.mark.parametrize("name", ["fw_az%02d" % r for r in AZ] + ["fw_fabric%s" % t for t in FAB])
def test_leading_edge_agrees_with_walk_back(name):
    """Where the envelope stays above 25 % of E1 from the first crossing to
    the peak the two picks coincide (after the 0.6 us delay the walk-back removes);
    where it dips below 25 % again on the way (fw_fabric50), the
    first-crossing pick is the earlier one."""
    from scipy.signal import hilbert
    z = _load(name)
    tr, dt = np.asarray(z["trace"], float), float(z["dt"])
    o = FS.trace_observables(tr, dt, 2e6, 6.0)
    back = _tof_walk_back(tr, dt) + 0.6
    e = np.abs(hilbert(tr))
    k0 = int(2 * 0.100 / 3850.0 / dt)
    seg = e[k0 - int(2e-6 / dt):k0 + int(2e-6 / dt)]
    ipk = int(np.argmax(seg))
    j = int(np.flatnonzero(seg[:ipk + 1] >= 0.25 * seg[ipk])[0])
    if np.all(seg[j:ipk + 1] > 0.25 * seg[ipk]):
        assert abs(o["tof_us"] - back) < 1e-9
    else:
        assert name == "fw_fabric50" and o["tof_us"] < back - 0.1
    # the back-wall echo of the 100 mm disk at ~3850 m/s, plus the Ricker delay
    assert 50.0 < o["tof_us"] < 54.5


def test_leading_edge_linear_ramp():
    """A linear ramp crossing the level between samples: the exact crossing."""
    seg = np.array([0.0, 0.1, 0.2, 0.3, 0.5, 1.0, 0.7])
    # samples 2 (0.2) and 3 (0.3) bracket 0.25: crossing at 2.5 = j - 1 + frac = 3 - 1 + 0.5
    assert FS._leading_edge(seg, 5, 0.25) == pytest.approx(2.5, abs=1e-12)
    assert FS._leading_edge(seg, 5, 0.0) == pytest.approx(-1.0)     # first sample qualifies


def test_gate_shape():
    dt = 1e-8
    n = 6000
    g = FS.fd_gate(n, dt)
    t = np.arange(n) * dt
    a, b = 24e-6 - 3e-6, 36e-6 + 3e-6
    assert (g[(t >= a) & (t <= b)] == 1).all()
    assert (g[(t <= a - 2e-6) | (t >= b + 2e-6)] == 0).all()
    assert np.all(np.diff(g[(t > a - 2e-6) & (t < a)]) > 0)
    # continuity at the ramp joins: the largest step is one ramp increment
    assert np.abs(np.diff(g)).max() < 2 * np.pi / 2 * dt / 2e-6


def test_band_bank_partition_of_unity():
    fr = np.linspace(0, 8e6, 4001)
    m = FS._band_masks(fr, 2e6, 6.0, 3)
    lo, hi = 1e6, 6e6
    w = (hi - lo) / 3
    inner = (fr >= lo + w / 2) & (fr <= hi - w / 2)
    assert np.allclose(m.sum(0)[inner], 1.0, atol=1e-12)
    assert (m[:, fr < lo - w / 2] == 0).all() and (m[:, fr > hi + w / 2] == 0).all()
    cen = lo + w * (np.arange(3) + 0.5)
    for b in range(3):
        k = int(np.argmin(np.abs(fr - cen[b])))
        assert m[b, k] == pytest.approx(1.0, abs=1e-6)


def test_fd_coda_geometric_mean_of_band_tones():
    """Three tones at the band centres with amplitudes 1, 2, 4: each band
    sees one tone whose envelope is its amplitude over the clean window, so
    the geometric combiner gives (1*2*4)^(1/3) = 2 (the arithmetic one
    would give sqrt(7)); a tone below the span adds nothing."""
    dt = 1e-8
    n = 8000
    t = np.arange(n) * dt
    cen = 1e6 + 5e6 / 3 * (np.arange(3) + 0.5)
    tr = sum(a * np.sin(2 * np.pi * f * t + 0.3 * k)
             for k, (a, f) in enumerate(zip((1.0, 2.0, 4.0), cen)))
    assert FS.fd_coda(tr, dt, 2e6, 6.0) == pytest.approx(2.0, rel=5e-3)
    low = tr + 3.0 * np.sin(2 * np.pi * 0.3e6 * t)
    assert FS.fd_coda(low, dt, 2e6, 6.0) == pytest.approx(2.0, rel=5e-3)


def test_fd_coda_batched_bands_equal_band_loop():
    """The bank evaluated as one 2-D transform equals band-by-band
    transforms."""
    from scipy.signal import hilbert
    z = _load("fw_az31")
    tr, dt = np.asarray(z["trace"], float), float(z["dt"])
    n = len(tr)
    spec = np.fft.rfft(tr * FS.fd_gate(n, dt))
    masks = FS._band_masks(np.fft.rfftfreq(n, dt), 2e6, 6.0, 3)
    i0, i1 = int(24e-6 / dt), int(36e-6 / dt)
    be = [(np.abs(hilbert(np.fft.irfft(spec * m, n)))[i0:i1] ** 2).mean() for m in masks]
    ref = 10 ** (np.mean(10 * np.log10(be)) / 20)
    assert FS.fd_coda(tr, dt, 2e6, 6.0) == pytest.approx(ref, rel=1e-12)


def test_centroid_and_speckle_of_a_tone():
    """A 2.25 MHz tone: 9 whole cycles in the 4 us centroid window, so its
    spectrum is one bin and the band-limited centroid is 2.25 MHz; a tone
    at 1.7 MHz leaks asymmetrically into the 0.5-4 MHz band and reads high.
    Its constant envelope has zero speckle contrast."""
    dt = 1e-8
    n = 7000
    t = np.arange(n) * dt
    tr = np.sin(2 * np.pi * 2.25e6 * t)
    o = FS.trace_observables(tr, dt, 2e6, 6.0)
    assert o["cent_mhz"] == pytest.approx(2.25, abs=1e-6)
    off = FS.trace_observables(np.sin(2 * np.pi * 1.7e6 * t), dt, 2e6, 6.0)["cent_mhz"]
    assert 1.7 < off < 2.0
    assert o["speckle_contrast"] < 1e-3
    assert o["tail_db"] == pytest.approx(0.0, abs=0.05)      # flat envelope: tail = E1
    short = FS.trace_observables(tr[:5500], dt, 2e6, 6.0)
    assert np.isnan(short["tail_db"])                         # < 2 us of 54-66 us recorded


def _rows(rots, coda_db, tof):
    return [dict(rot=r, coda=10 ** (c / 20), tof_us=t) for r, c, t in zip(rots, coda_db, tof)]


def test_smoothing_folds_mod_180():
    rots = [10, 11, 12, 190, 191, 192, 50]
    db = [0.0, 3.0, 6.0, 2.0, 5.0, 8.0, 1.0]
    out = FS.smooth_azimuths(_rows(rots, db, db))
    # rot 11 averages 10, 11, 12 and 190, 191, 192
    assert 20 * np.log10(out[1]["coda"]) == pytest.approx(np.mean(db[:6]))
    assert out[1]["tof_us"] == pytest.approx(np.mean(db[:6]))
    # rot 10 sees 10, 11 and 190, 191 (12 and 192 are 2 deg away)
    assert out[0]["tof_us"] == pytest.approx(np.mean([0, 3, 2, 5]))
    assert out[6]["tof_us"] == 1.0                              # alone
    few = FS.smooth_azimuths(_rows(rots[:4], db[:4], db[:4]))
    assert [r["tof_us"] for r in few] == db[:4]                  # <= 4 rows: untouched


def test_smoothing_keeps_linear_pattern():
    rots = list(range(20, 41))
    out = FS.smooth_azimuths(_rows(rots, [0.1 * r for r in rots], [0.2 * r for r in rots]))
    for r in out[1:-1]:
        assert r["tof_us"] == pytest.approx(0.2 * r["rot"])
        assert 20 * np.log10(r["coda"]) == pytest.approx(0.1 * r["rot"])


def _fake_sweep(root, rots):
    SR.create(str(root), "s", f0_mhz=2.0, ppw=6.0)
    cfg = SR.load(str(root), "s")
    for k, rot in enumerate(rots):
        z = _load("fw_az%02d" % AZ[k % len(AZ)])
        SR.save_azimuth(str(root), cfg, rot, z["trace"], float(z["dt"]),
                        3850.0 / 2e6 / 6.0)
    return cfg


def test_sweep_data_rows_and_cache(tmp_path):
    cfg = _fake_sweep(tmp_path, [0, 1, 2, 3, 4, 5])
    rows = FS.sweep_data(str(tmp_path), cfg)
    assert [r["rot"] for r in rows] == [0, 1, 2, 3, 4, 5]
    z = _load("fw_az17")
    o = FS.trace_observables(z["trace"], float(z["dt"]), 2e6, 6.0)
    assert rows[0]["e1"] == o["e1"] and rows[0]["cent_mhz"] == o["cent_mhz"]
    assert rows[0]["e1_db"] == pytest.approx(20 * np.log10(o["e1"]))
    cache = np.load(os.path.join(tmp_path, "s", "obs_cache.npz"), allow_pickle=False)
    assert str(cache["ver"]) == FS.OBS_VER == "fd3geo-le25-t4"
    assert list(cache["az"]) == [0, 1, 2, 3, 4, 5]
    assert cache["coda"][0] == o["coda"] and cache["speckle"][0] == o["speckle_contrast"]
    # a second call reads the cache: poison a trace file, the rows are unchanged
    again = FS.sweep_data(str(tmp_path), cfg)
    p = os.path.join(tmp_path, "s", "az000.npz")
    os.remove(p)
    np.savez(p, trace=np.zeros(10, np.float32), dt=1e-8)
    assert FS.sweep_data(str(tmp_path), cfg) == again
    # a new azimuth extends the cache; a stale version recomputes everything
    SR.save_azimuth(str(tmp_path), cfg, 7, z["trace"], float(z["dt"]), 3850.0 / 2e6 / 6.0)
    rows = FS.sweep_data(str(tmp_path), cfg)
    assert [r["rot"] for r in rows] == [0, 1, 2, 3, 4, 5, 7]
    cp = os.path.join(tmp_path, "s", "obs_cache.npz")
    c = dict(np.load(cp, allow_pickle=False))
    c["ver"] = np.array("old")
    np.savez(cp, **c)
    with pytest.raises(Exception):
        FS.sweep_data(str(tmp_path), cfg)                     # reads the poisoned az000


def test_sweep_data_uses_config_f0_when_missing(tmp_path):
    cfg = _fake_sweep(tmp_path, [0])
    d = os.path.join(tmp_path, "s")
    z = dict(np.load(os.path.join(d, "az000.npz"), allow_pickle=False))
    del z["f0"], z["ppw"]
    np.savez(os.path.join(d, "az000.npz"), **z)
    with open(os.path.join(d, "config.json")) as f:
        json.load(f)
    r = FS.sweep_data(str(tmp_path), cfg)[0]
    ref = FS.trace_observables(z["trace"], float(z["dt"]), 2e6, 6.0)
    assert r["coda"] == ref["coda"] and r["cent_mhz"] == ref["cent_mhz"]
