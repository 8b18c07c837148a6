"""Host-side checks of the benchmark and of the launch guards (no GPU).

* ``bench.cpu_baseline`` runs on a configuration whose wavelet is shorter
  than its warm-up + timed steps (the benchmark is run with --steps 20:
  the configuration's record length then has 20 samples).
* ``bench.py --gpus 2`` refuses with a message (one GPU only).
* The run guards of DeviceShot and ShotBatch refuse wavelet and trace
  buffers shorter than the steps they would be read or written for (the
  kernels take raw pointers and no lengths).
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpu_baseline_short_configuration():
    import bench
    from pulse_echo_azimuthal_simulation_amd import specimen
    cfg = specimen.config_c0(nt=20)
    assert len(cfg["wavelet"]) == 20
    out = bench.cpu_baseline(cfg, nz_sample=4, warm=2, timed=20)
    assert out["value"] > 0 and out["unit"] == "Gcell-updates/s"
    assert out["cores"] == 1 and out["kind"] == "port"
    assert "20 timed steps" in out["sample"]


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


class _FakeShot:
    """The attributes DeviceShot._check_run reads, with CPU tensors."""

    def __init__(self, torch, n_rec=2):
        self.torch = torch
        self.device = torch.device("cpu")
        self.real = torch.float32
        self.n_rec = n_rec


def test_run_guard_refuses_short_buffers():
    torch = pytest.importorskip("torch")
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    sh = _FakeShot(torch)
    check = DeviceShot._check_run
    wav = torch.zeros(20, dtype=torch.float64)
    tr = torch.zeros((20, 2), dtype=torch.float32)
    check(sh, wav, 0, 20, tr)                      # fits
    check(sh, wav, 5, 20, None)
    with pytest.raises(ValueError, match="wavelet has 20 samples"):
        check(sh, wav, 0, 2000, None)              # the sponge leg's old 2000 steps
    with pytest.raises(ValueError, match="traces hold 20 rows"):
        check(sh, torch.zeros(40, dtype=torch.float64), 0, 40, tr)
    with pytest.raises(ValueError, match="float64"):
        check(sh, wav.float(), 0, 10, None)
    with pytest.raises(ValueError, match="n_rec"):
        check(sh, wav, 0, 10, torch.zeros((20, 3), dtype=torch.float32))
    with pytest.raises(ValueError, match="bad step range"):
        check(sh, wav, 5, 4, None)


def test_batch_guard_refuses_short_wavelet():
    torch = pytest.importorskip("torch")
    from pulse_echo_azimuthal_simulation_amd import sweep
    sh = _FakeShot(torch)
    with pytest.raises(ValueError, match="wavelet has 10 samples, nt = 11"):
        sweep.ShotBatch([sh], [torch.zeros(10, dtype=torch.float64)], [11])
    with pytest.raises(ValueError, match="one wavelet and one step count per shot"):
        sweep.ShotBatch([sh], [], [11])
    with pytest.raises(ValueError, match="float64"):
        sweep.ShotBatch([sh], [np.zeros(20)], [11])
