"""The upstream's bit-equivalence gate on the GPU (sim/core/validate_labels.py:
66-100, "quick"): the dense-table twin ``forward_fused`` (separate kernels:
36 per-cell stiffness planes, face-density volumes, damping after the
velocity update, post-damp recording) against the label path
``forward_fused_labels`` on the same specimen, over the FULL record; the
traces must be identical (max |a - b| / max |a| == 0).

Specimen: the gate's DiskSpecimen (60 mm x 25 mm, 40 grains, size_cv 0.35,
Watson kappa 3.93 about x, seed 7) at 2 MHz, ppw 6, 10-cell sponge + the
exp(-0.02 EDT) couplant mask, rim probe -- rebuilt by disk_specimen, whose
output is pinned bit for bit to the upstream's stored builds.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.golden import recipes as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def gate():
    from pulse_echo_azimuthal_simulation_amd.disk_specimen import DiskSpecimen
    sp = DiskSpecimen(diameter_m=0.060, thickness_m=0.025, n_grains=40, size_cv=0.35,
                      concentration=3.93, spatial_corr=0.0, fabric_axis=(1, 0, 0), seed=7)
    build = sp.build(R.C_REF / 2e6 / 6.0, adjacency=False)
    st = R.validate_setup(2e6, build)
    st["axes"] = build["axes"]
    return st


@pytest.mark.parametrize("precision,kernel", [("f32", "tiled"), ("f32", "direct"),
                                              ("f64", "tiled")])
def test_dense_twin_bit_identical_to_labels(torch_gpu, gate, precision, kernel):
    from pulse_echo_azimuthal_simulation_amd import materials, solver
    st = gate
    print("\n[gpu case] dense gate %s %s, grid %s, %d steps"
          % (precision, kernel, st["lab"].shape, st["nt"]), flush=True)
    srcs = [(p, st["w"]) for p in st["pts"]]
    recs = [(st["pts"], np.full(len(st["pts"]), st["w"]))]
    C, rho = materials.polycrystal_stiffness_3d(st["lab"], st["axes"], 1500.0, 300.0)
    a = solver.forward_fused(C, rho, st["h"], st["dt"], st["nt"], srcs, st["wavelet"], recs,
                             order=8, coeffs=st["coeffs"], sponge_width=st["sponge"],
                             damp_mask=st["damp_mask"], precision=precision).ravel()
    del C, rho
    b = solver.forward_fused_labels(st["lab"], st["axes"], st["h"], st["dt"], st["nt"], srcs,
                                    st["wavelet"], recs, order=8, coeffs=st["coeffs"],
                                    sponge_width=st["sponge"], damp_mask=st["damp_mask"],
                                    precision=precision, kernel=kernel).ravel()
    scale = np.abs(a).max()
    err = np.abs(a.astype(float) - b.astype(float)).max() / scale
    print("max |diff| / max |trace| = %.3e over %d steps" % (err, st["nt"]), flush=True)
    assert scale > 0 and np.isfinite(a).all()
    assert err == 0.0


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_dense_twin_random_medium(torch_gpu, precision):
    """Same gate on a random 11-grain polycrystal with couplant, sponge and
    fluid mask, every receiver group at once (vx record as well)."""
    from pulse_echo_azimuthal_simulation_amd import materials, solver
    from tests import helpers as H
    print("\n[gpu case] dense gate random medium %s" % precision, flush=True)
    inp = H.random_medium((30, 34, 38), n_grain=11, seed=5, sponge=4, with_mask=False)
    dm = H.host.fluid_damp_mask(inp["lab"], 0.02)
    C, rho = materials.polycrystal_stiffness_3d(inp["lab"], inp["axes"])
    wav = H.host.ricker(2.0e6, inp["dt"], 200, t0=4 * inp["dt"])
    for record in ("pressure", "vx"):
        kw = dict(order=8, coeffs=inp["coeffs"], sponge_width=4, damp_mask=dm,
                  record=record, precision=precision)
        a = solver.forward_fused(C, rho, inp["h"], inp["dt"], 200, inp["src_pts"],
                                 wav, inp["groups"], **kw)
        b = solver.forward_fused_labels(inp["lab"], inp["axes"], inp["h"], inp["dt"], 200,
                                        inp["src_pts"], wav,
                                        inp["groups"], **kw)
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b), record
