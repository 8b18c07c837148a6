"""The batched multi-shot kernels at the azimuth sweep's real shot shapes,
against the CPU oracle (not against single-shot HIP runs).

Batch: two azimuths (0 and 35 deg) of the 100 mm x 35 mm, 3000-grain
girdle-fabric disk prepared on the GPU exactly as the sweep prepares them
(169 x 397 x 397 each), plus the smallest and the largest of the sweep's
uneven disks (grids 128 x 128 x 64 and 256 x 256 x 128).  Every shot starts
from one random state of every field (and of both buffers of every CPML
memory variable), so every tile, z chunk and band path of the batched
kernels works from the first step; the batch advances 3 steps through
pe_batch (graph replays), the oracle advances each shot 3 steps from the
same state.  Run with the sweep's 20-cell sponge and with a 20-cell CPML.

Bars: every field and memory variable bit for bit; traces within 1e-6
rel-L2 (tree-ordered receiver sums).  Reference: the per-azimuth recipe of
sim/pipeline/sweep_runner.py:331-365 solved one azimuth at a time by
forward_fused_labels (sim/core/fdtd.py:1618-1816).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

NSTEP = 3


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def disks():
    from pulse_echo_azimuthal_simulation_amd import specimen, sweep
    f0 = 2.25e6
    h = specimen.C_REF / f0 / 6
    big = sweep.disk_build(100e-3, 35e-3, h, 3000, seed=11, kappa=-3.0,
                           fabric_axis=np.random.default_rng(11).normal(size=3))
    smalls = []
    for k in (0, 7):                               # grid sides 128 and 256
        n = 128 + 128 * k // 7
        nd, nzd = n - 46, max(n // 2 - 46, 12)
        smalls.append(sweep.disk_build(nd * h, nzd * h, h, 60 + 40 * k, seed=100 + k, kappa=5.0,
                                       fabric_axis=(np.cos(0.5), np.sin(0.5), 0.0)))
    return h, f0, big, smalls


def _oracle_of(shot, h):
    from oracle import oracle as O
    d = shot.desc
    co = np.array([d.coeffs[j] for j in range(d.nc)])
    damp = None if shot.damp is None else shot.damp.cpu().numpy()
    cp = None if shot.cpml is None else (shot.cpml.width, shot.cpml.prof)
    return O.OracleShot(shot.mats.cpu().numpy(), shot.Ct.cpu().numpy(), shot.rho_t.cpu().numpy(),
                        damp, co, h, d.dt,
                        (shot.src_idx.cpu().numpy(), shot.src_w.cpu().numpy().astype(float)),
                        (shot.rec_ptr.cpu().numpy(), shot.rec_idx.cpu().numpy(),
                         shot.rec_w.cpu().numpy().astype(float)),
                        precision="f32", cpml=cp)


@pytest.mark.parametrize("boundary", ["sponge", "cpml"])
def test_batch_at_sweep_shapes_matches_oracle(torch_gpu, disks, boundary):
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import host, sweep
    h, f0, big, smalls = disks
    cw = 20 if boundary == "cpml" else 0
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True)
    prep_big = sweep.GpuAzimuthPrep(big, h, co, f0, sponge=20, cpml_width=cw)
    shots, wavs = [], []
    for az in (0.0, 35.0):
        sh, wav, _, _ = prep_big.shot(az, nt=NSTEP)
        shots.append(sh)
        wavs.append(wav)
    for b in smalls:
        sh, wav, _, _ = sweep.GpuAzimuthPrep(b, h, co, f0, sponge=20, cpml_width=cw).shot(
            0.0, nt=NSTEP)
        shots.append(sh)
        wavs.append(wav)
    shapes = [s.shape for s in shots]
    print("\n[gpu case] batch %s at sweep shapes %s" % (boundary, shapes), flush=True)
    assert shapes[0] == shapes[1] == (169, 397, 397)
    assert shapes[2] == (64, 128, 128) and shapes[3] == (128, 256, 256)
    torch.cuda.synchronize()
    oracles = []
    for k, sh in enumerate(shots):
        osh = _oracle_of(sh, h)
        H.random_state(torch, sh, osh, seed=20 + k)
        oracles.append(osh)
    batch = sweep.ShotBatch(shots, wavs, [NSTEP] * len(shots))
    out = batch.run(graph_steps=2)
    torch.cuda.synchronize()
    for k, (sh, osh, wav) in enumerate(zip(shots, oracles, wavs)):
        otr = osh.run(wav.cpu().numpy(), 0, NSTEP)
        print("[gpu case] shot %d %s compared" % (k, sh.shape), flush=True)
        H.compare_fields(torch, sh, osh, out[k], otr)
