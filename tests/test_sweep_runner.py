"""Sweep runner state machine (host side): visiting order, resume from
files, atomic writes with retries, status merging, trace observables and
the refusals the upstream runner makes."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from pulse_echo_azimuthal_simulation_amd import sweep_runner as SR


def test_bit_reversed_order():
    assert SR.bit_reversed_order(8) == [0, 4, 2, 6, 1, 5, 3, 7]
    for n in (1, 5, 13, 180):
        o = SR.bit_reversed_order(n)
        assert sorted(o) == list(range(n))


def test_progressive_azimuths_cover_half_circle_first():
    cfg = dict(SR.DEFAULTS, az_start=0, az_stop=360, az_step=5)
    order = SR.azimuths(cfg)
    assert sorted(order) == list(range(0, 360, 5))
    assert all(a < 180 for a in order[:36]) and all(a >= 180 for a in order[36:])
    # van der Corput over the 36 half-circle slots (6 bits; 32, 16, 8, 24, 4, ...)
    assert order[:6] == [0, 160, 80, 40, 120, 20]
    assert SR.azimuths(dict(cfg, sweep_order="sequential")) == list(range(0, 360, 5))


def test_create_load_resume(tmp_path):
    root = str(tmp_path)
    SR.create(root, "s1", az_step=90)
    cfg = SR.load(root, "s1")
    assert cfg["name"] == "s1" and cfg["az_step"] == 90
    SR.create(root, "s1", az_step=45)                       # never overwritten
    assert SR.load(root, "s1")["az_step"] == 90
    d = os.path.join(root, "s1")
    np.savez(os.path.join(d, "az090.npz"), trace=np.zeros(3))
    np.savez(os.path.join(d, "az180.tmp.npz"), trace=np.zeros(3))   # a killed write
    assert SR.done_azimuths(root, cfg) == [90]
    assert SR.pending(root, cfg) == [0, 180, 270]


def test_atomic_replace_retries(tmp_path, monkeypatch):
    src = tmp_path / "a.tmp"
    dst = tmp_path / "a"
    src.write_text("x")
    real = os.replace
    calls = {"n": 0}

    def flaky(a, b):
        calls["n"] += 1
        if calls["n"] < 3:
            raise PermissionError("held open")
        return real(a, b)

    monkeypatch.setattr(SR.os, "replace", flaky)
    assert SR.atomic_replace(str(src), str(dst), critical=True, wait=0.0)
    assert dst.read_text() == "x" and calls["n"] == 3

    def never(a, b):
        raise PermissionError("held open")

    monkeypatch.setattr(SR.os, "replace", never)
    src.write_text("y")
    with pytest.raises(PermissionError):
        SR.atomic_replace(str(src), str(dst), critical=True, retries=3, wait=0.0)
    assert not SR.atomic_replace(str(src), str(dst), critical=False, retries=2, wait=0.0)
    assert not src.exists()                                  # cosmetic files give up


def test_status_merges(tmp_path):
    root = str(tmp_path)
    SR.create(root, "s", az_step=120)
    cfg = SR.load(root, "s")
    SR.status(root, cfg, phase="az 0", sec_per_az=12.5)
    s = SR.status(root, cfg, phase="stopped")
    assert s["sec_per_az"] == 12.5 and s["phase"] == "stopped" and s["total"] == 3


def test_analyse_observables():
    dt = 2e-8
    n = 3000
    t = np.arange(n) * dt
    t_echo = 2 * 0.100 / SR.C_REF
    pulse = lambda t0, a: a * np.exp(-((t - t0) / 0.4e-6) ** 2) * np.cos(2 * np.pi * 2.25e6 * (t - t0))
    tr = pulse(t_echo, 1.0) + pulse(30e-6, 0.01) + pulse(t_echo + 5e-6, 0.02)
    E1, t1, coda, tail = SR.analyse(tr, dt)
    assert abs(t1 - t_echo) <= 2 * dt                         # back wall at 2 D / c
    assert E1 == pytest.approx(1.0, rel=2e-2)
    assert -60 < coda < -30                                   # the weak 30 us packet (24-36 us)
    assert tail == pytest.approx(0.02, rel=0.1)               # the packet 5 us after E1


def test_legacy_and_no_gpu_refusals(tmp_path, monkeypatch):
    root = str(tmp_path)
    SR.create(root, "old", az_step=90)
    cfg = SR.load(root, "old")
    cfg.pop("axes_convention")
    np.savez(os.path.join(root, "old", "az000.npz"), trace=np.zeros(2))
    assert SR.run(root, cfg, log=lambda *a: None) == "legacy"
    assert json.load(open(os.path.join(root, "old", "status.json")))["phase"].startswith("LEGACY")
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    SR.create(root, "new", az_step=90)
    assert SR.run(root, SR.load(root, "new"), log=lambda *a: None) == "no-gpu"
