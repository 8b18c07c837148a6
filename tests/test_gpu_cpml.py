"""GPU parity of the CPML boundary (through the C ABI) against the CPU
oracle's band path: fields, traces and memory variables bit for bit in
float32 and float64, for the direct kernels (corrections computed in line)
and the tiled kernels (band pre-pass, corrections added by every tile
variant), with z-chunks short enough that neighbouring workgroups share
band planes and ring columns (which read the corrections the pre-pass wrote
once).
"""
from __future__ import annotations

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _medium(shape=(22, 25, 29), seed=5, L=4):
    from pulse_echo_azimuthal_simulation_amd import cpml
    inp = H.random_medium(shape, n_grain=7, seed=seed, sponge=0, fluid_pad=2, with_mask=True)
    v = cpml.vmax_table(inp["Ct"], inp["rho_t"])
    cp = cpml.Cpml(shape, L, inp["h"], inp["dt"], v, 2e6, R=1e-6, kappa_max=1.3)
    return inp, cp


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("kernel", ["direct", "tiled"])
def test_cpml_matches_oracle(torch_gpu, kernel, precision):
    print("\n[gpu case] cpml %s %s" % (kernel, precision), flush=True)
    from oracle import oracle as O
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    inp, cp = _medium()
    nt = inp["nt"]
    sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"], inp["h"],
                    inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"], precision=precision,
                    kernel=kernel, zchunk=(7, 7), cpml=cp)
    wav = sh.wavelet(inp["wavelet"])
    tr = sh.new_traces(nt)
    sh.run(wav, 0, nt, tr)
    v, s = sh.fields()
    tr = tr.cpu().numpy()
    psi = [p.cpu().numpy() for p in sh.psi]
    osh = O.OracleShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                       inp["h"], inp["dt"], inp["src"], inp["rec"], precision=precision,
                       cpml=(cp.width, cp.prof))
    otr = osh.run(inp["wavelet"], 0, nt)
    assert np.abs(otr).max() > 0 and any(np.abs(p).max() > 0 for p in osh.psi)
    for k in range(3):
        assert np.array_equal(v[k], osh.v[k]), "v%d" % k
    for k in range(6):
        assert np.array_equal(s[k], osh.s[k]), "s%d" % (k + 1)
    for a in range(3):
        assert np.array_equal(psi[a].reshape(osh.psi[a].shape), osh.psi[a]), "psi axis %d" % a
    assert H.rel_l2(tr, otr) <= 1e-6


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_cpml_every_tile_variant_matches_oracle(torch_gpu, precision):
    """Every stress and velocity tile variant of the tiled kernels with a
    CPML (tiles inside, across and clear of every band; 5-plane chunks)
    gives the oracle's fields and memory variables bit for bit."""
    from oracle import oracle as O
    from pulse_echo_azimuthal_simulation_amd import _abi
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    shape, L = (24, 40, 136), 4
    inp, cp = _medium(shape=shape, seed=13, L=L)
    nt = 24
    osh = O.OracleShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                       inp["h"], inp["dt"], inp["src"], inp["rec"], precision=precision,
                       cpml=(cp.width, cp.prof))
    otr = osh.run(inp["wavelet"], 0, nt)
    assert np.abs(otr).max() > 0
    nst = len(_abi.tile_variants(precision, 0))
    nve = len(_abi.tile_variants(precision, 1))
    for tile in [(v, 0) for v in range(nst)] + [(0, v) for v in range(1, nve)]:
        print("\n[gpu case] cpml tile variants %s %s" % (precision, tile), flush=True)
        sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                        inp["h"], inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"],
                        precision=precision, kernel="tiled", zchunk=(5, 5), cpml=cp, tile=tile)
        wav = sh.wavelet(inp["wavelet"])
        tr = sh.new_traces(nt)
        sh.run(wav, 0, nt, tr)
        v, s = sh.fields()
        for k in range(3):
            assert np.array_equal(v[k], osh.v[k]), ("v%d" % k, tile)
        for k in range(6):
            assert np.array_equal(s[k], osh.s[k]), ("s%d" % (k + 1), tile)
        for a in range(3):
            assert np.array_equal(sh.psi[a].cpu().numpy().reshape(osh.psi[a].shape),
                                  osh.psi[a]), ("psi axis %d" % a, tile)
        assert H.rel_l2(tr.cpu().numpy(), otr) <= 1e-6
        del sh


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("nz", [10, 11])
def test_cpml_thin_grid_without_band_free_middle(torch_gpu, nz, precision):
    """nz = 2L+2 and 2L+3 with L = 4: every plane lies in a z band (nz = 10)
    or one plane is clear of them (nz = 11); the band enumeration of the
    pre-pass has no middle planes, or one."""
    print("\n[gpu case] cpml thin grid nz %d %s" % (nz, precision), flush=True)
    from oracle import oracle as O
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    inp, cp = _medium(shape=(nz, 25, 29), seed=11)
    nt = inp["nt"]
    sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"], inp["h"],
                    inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"], precision=precision,
                    kernel="tiled", zchunk=(4, 3), cpml=cp)
    wav = sh.wavelet(inp["wavelet"])
    tr = sh.new_traces(nt)
    sh.run(wav, 0, nt, tr)
    v, s = sh.fields()
    osh = O.OracleShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                       inp["h"], inp["dt"], inp["src"], inp["rec"], precision=precision,
                       cpml=(cp.width, cp.prof))
    otr = osh.run(inp["wavelet"], 0, nt)
    assert np.abs(otr).max() > 0
    for k in range(3):
        assert np.array_equal(v[k], osh.v[k]), "v%d" % k
    for k in range(6):
        assert np.array_equal(s[k], osh.s[k]), "s%d" % (k + 1)
    assert H.rel_l2(tr.cpu().numpy(), otr) <= 1e-6


@pytest.mark.parametrize("precision,tile", [("f32", (0, 0)), ("f64", (0, 0)), ("f32", (8, 2)),
                                            ("f32", (12, 3)), ("f32", (13, 1)), ("f64", (4, 1))])
def test_cpml_batch_equals_single_shots(torch_gpu, precision, tile):
    """Two CPML shots of different sizes, band widths and step counts and a
    sponge shot in one batch (graph replays of 7 steps): every trace, field
    and memory variable bit-identical to that shot's own run.  Tile
    variants: the defaults, the frame-ring stress tile with 32x8 velocity
    tiles, and the one-barrier stress tiles (which a batch with CPML shots
    replaces by the plain kernel of the same tile, as a single shot does:
    32x13 and 30x14 in f32, 32x6 in f64) with 64x4, 32x16 and 64x4 velocity
    tiles."""
    print("\n[gpu case] cpml batch %s %s" % (precision, tile), flush=True)
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import sweep
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    items = []
    for shape, seed, L, nt in (((24, 40, 136), 13, 4, 40), ((22, 25, 29), 5, 3, 33)):
        inp, cp = _medium(shape=shape, seed=seed, L=L)
        sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                        inp["h"], inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"],
                        precision=precision, kernel="tiled", cpml=cp, tile=tile)
        items.append((sh, sh.wavelet(np.resize(inp["wavelet"], nt)), nt))
    inp = H.random_medium((20, 26, 31), n_grain=6, seed=1)
    sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"], inp["h"],
                    inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"], precision=precision,
                    kernel="tiled", tile=tile)
    items.append((sh, sh.wavelet(np.resize(inp["wavelet"], 37)), 37))
    singles = []
    for sh, wav, nt in items:
        tr = sh.new_traces(nt)
        sh.run(wav, 0, nt, tr)
        v, s = sh.fields()
        singles.append((tr.cpu().numpy(), v + s, [p.cpu().numpy() for p in sh.psi]))
        sh.reset()
    batch = sweep.ShotBatch([i[0] for i in items], [i[1] for i in items], [i[2] for i in items])
    out = batch.run(graph_steps=7)
    torch.cuda.synchronize()
    for k, ((sh, _, _), tr, (ref, fields, psi)) in enumerate(zip(items, out, singles)):
        assert np.abs(ref).max() > 0, k
        assert np.array_equal(tr.cpu().numpy(), ref), k
        v, s = sh.fields()
        for a, b in zip(v + s, fields):
            assert np.array_equal(a, b), k
        for a, b in zip(sh.psi, psi):
            assert np.array_equal(a.cpu().numpy(), b), k


def test_cpml_drop_in_boundary_option(torch_gpu):
    """forward_fused_labels(boundary='cpml') runs the same kernels as a
    DeviceShot built with the same profiles."""
    print("\n[gpu case] cpml drop-in", flush=True)
    from pulse_echo_azimuthal_simulation_amd import cpml, solver
    inp, _ = _medium()
    tr = solver.forward_fused_labels(inp["lab"], inp["axes"], inp["h"], inp["dt"], inp["nt"],
                                     inp["src_pts"], inp["wavelet"], inp["groups"],
                                     coeffs=inp["coeffs"], damp_mask=inp["damp"],
                                     boundary="cpml", cpml_width=5, cpml_f0=2e6)
    prep = solver.prepare_shot(inp["lab"], inp["axes"], inp["h"], inp["dt"], inp["src_pts"],
                               inp["groups"], coeffs=inp["coeffs"], damp_mask=inp["damp"])
    cp = cpml.Cpml(inp["shape"], 5, inp["h"], inp["dt"],
                   cpml.vmax_table(prep["Ct"], prep["rho_t"]), 2e6)
    sh = solver.DeviceShot(prep["mats"], prep["Ct"], prep["rho_t"], prep["damp"],
                           prep["coeffs"], inp["h"], inp["dt"], prep["src_idx"], prep["src_w"],
                           prep["rec_ptr"], prep["rec_idx"], prep["rec_w"], cpml=cp)
    wav = sh.wavelet(inp["wavelet"])
    t2 = sh.new_traces(inp["nt"])
    sh.run(wav, 0, inp["nt"], t2)
    assert np.abs(tr).max() > 0
    assert np.array_equal(tr, t2.cpu().numpy())
    with pytest.raises(ValueError):
        solver.forward_fused_labels(inp["lab"], inp["axes"], inp["h"], inp["dt"], 4,
                                    inp["src_pts"], inp["wavelet"], inp["groups"],
                                    coeffs=inp["coeffs"], sponge_width=3, boundary="cpml")
