"""The upstream's full-disk gate on the GPU: the drop-in forward_fused_labels
on the 2 MHz, 100-grain, 100 mm x 35 mm disk of ladder.standard_build
(rebuilt by disk_specimen.DiskSpecimen, whose labels are pinned bit for bit
to the upstream's stored builds) with each of the four fabric sets, against
the upstream's stored traces sim/ref/fw_fabric{00,25,50,100}.npz.

Bar (sim/core/validate_labels.py:103-132): correlation > 0.999 over the
whole 4678-step record and the coda level (sim/fw_checks/
fw_reference_2mhz.py:38-46) within 0.2 dB of the stored trace's; the test
also holds the measured agreement (rel-L2 below 1e-5, coda within 0.01 dB).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests import helpers as H
from tests.golden import recipes as R

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def setup():
    return R.fulldisk_setup()


@pytest.mark.parametrize("tag", ["00", "25", "50", "100"])
def test_fulldisk_gate(torch_gpu, setup, tag):
    from pulse_echo_azimuthal_simulation_amd import solver
    print("\n[gpu case] full-disk gate fabric %s" % tag, flush=True)
    ref = np.load(os.path.join(HERE, "golden", "reference", "fw_fabric%s.npz" % tag),
                  allow_pickle=False)
    st = setup
    assert st["lab"].shape == (136, 338, 338)
    assert st["nt"] == len(ref["trace"])
    assert abs(st["dt"] / float(ref["dt"]) - 1.0) < 1e-7
    axes = R.fulldisk_fabrics(st["build"]["axes"])[tag]
    pts, w = st["pts"], st["w"]
    tr = solver.forward_fused_labels(st["lab"], axes, st["h"], st["dt"], st["nt"],
                                     [(p, w) for p in pts], st["wavelet"],
                                     [(pts, np.full(len(pts), w))], order=8,
                                     coeffs=st["coeffs"], sponge_width=10,
                                     damp_mask=st["damp_mask"])[:, 0].astype(float)
    refr = np.asarray(ref["trace"], float)
    corr = float(np.corrcoef(tr, refr)[0, 1])
    la, E1a = R.coda_level(tr, st["dt"])
    lb, E1b = R.coda_level(refr, float(ref["dt"]))
    print("corr %.7f, rel-L2 %.3e, coda %.3f vs %.3f dB (stored %.3f), E1 %.5e vs %.5e"
          % (corr, H.rel_l2(tr, refr), la, lb, float(ref["level_db"]), E1a, E1b), flush=True)
    assert np.isfinite(tr).all()
    assert corr > 0.999
    assert abs(la - lb) < 0.2
    # measured: rel-L2 0.6-1.4e-6 and coda within 0.001 dB on all four sets
    assert H.rel_l2(tr, refr) < 1e-5
    assert abs(la - lb) < 0.01


@pytest.mark.parametrize("rot", [17, 31, 47, 62, 76, 90])
def test_rotated_disk_gate(torch_gpu, setup, rot):
    """The same disk rotated by sweep.rotated_grid (rotation_test.py:60-137)
    against the upstream's sim/ref/fw_az{rot}.npz (azimuth_sweep.py:60-121):
    the stored dt reproduced, then the full record held to the gate's bar
    and its stored clean-window level (envelope RMS over 24-36 us re E1)."""
    from pulse_echo_azimuthal_simulation_amd import fit_sweep, solver
    print("\n[gpu case] rotated-disk gate rot %d" % rot, flush=True)
    ref = np.load(os.path.join(HERE, "golden", "reference", "fw_az%02d.npz" % rot),
                  allow_pickle=False)
    st = R.azimuth_setup(setup["build"], rot)
    assert st["nt"] == len(ref["trace"])
    assert abs(st["dt"] / float(ref["dt"]) - 1.0) < 1e-7
    pts, w = st["pts"], st["w"]
    tr = solver.forward_fused_labels(st["lab"], st["axes"], st["h"], st["dt"], st["nt"],
                                     [(p, w) for p in pts], st["wavelet"],
                                     [(pts, np.full(len(pts), w))], order=8,
                                     coeffs=st["coeffs"], sponge_width=st["sponge"],
                                     damp_mask=st["damp_mask"])[:, 0].astype(float)
    refr = np.asarray(ref["trace"], float)
    corr = float(np.corrcoef(tr, refr)[0, 1])

    def level(x, dt):
        o = fit_sweep.trace_observables(x, dt, 2e6, 6.0, nb=1)
        return 20 * np.log10(o["coda"] / o["e1"] + 1e-30)

    la = level(tr, st["dt"])
    print("corr %.7f, rel-L2 %.3e, clean level %.3f dB vs stored %.3f dB"
          % (corr, H.rel_l2(tr, refr), la, float(ref["level_db"])), flush=True)
    assert np.isfinite(tr).all()
    assert corr > 0.999
    assert abs(la - float(ref["level_db"])) < 0.2
    # measured: rel-L2 0.05-0.9e-6, levels equal to 0.001 dB
    assert H.rel_l2(tr, refr) < 1e-5
    assert abs(la - float(ref["level_db"])) < 0.01
