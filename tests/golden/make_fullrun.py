"""Write the full-run golden traces of the benchmark workloads with the CPU
oracle (float32, the arithmetic of the upstream float32 solver, pinned bit
for bit to the upstream archives in tests/golden/reference/).

Cases (tests/golden/fullrun/<name>.npz):

* c0_cpml_2000   96x96x48 disk configuration (+23-cell couplant pads,
                 200 grains, 1 MHz 3-cycle Hann burst, 6 mm top piston),
                 20-cell CPML + couplant mask, all 2000 steps;
* c0_sponge_2000 the same with the upstream 20-cell sponge;
* c1_sponge_2000 400x400x200 single-shot configuration (3000 grains,
                 girdle fabric, 2.25 MHz Ricker), upstream 20-cell sponge
                 + couplant mask, 2000 steps: past the first back-wall echo
                 (35 mm thick disk, round trip ~18 us = ~1480 steps).

Each file holds the (nt, n_rec) float32 trace, dt, the grid shape and a
CRC-32 of the label volume, so a test can tell a changed input from a
changed solver.  Usage (8 cores: c0 ~3 min each, c1 ~45 min):

    python tests/golden/make_fullrun.py [case ...]
"""
from __future__ import annotations

import os
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fullrun")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CASES = {
    "c0_cpml_2000": ("c0", "cpml", 2000),
    "c0_sponge_2000": ("c0", "sponge", 2000),
    "c1_sponge_2000": ("c1", "sponge", 2000),
}


def config(which, nt):
    from pulse_echo_azimuthal_simulation_amd import specimen
    return specimen.config_c0(nt=nt) if which == "c0" else specimen.config_c1(nt=nt)


def label_crc(lab):
    return zlib.crc32(np.ascontiguousarray(lab, np.int32).tobytes()) & 0xFFFFFFFF


def oracle_shot(cfg, boundary):
    """OracleShot of a workload, host inputs prepared as forward_fused_labels
    prepares them (sponge x couplant mask, or mask only under the CPML)."""
    from oracle import oracle as O
    from pulse_echo_azimuthal_simulation_amd import cpml, solver
    sponge = cfg["sponge_width"] if boundary == "sponge" else 0
    prep = solver.prepare_shot(cfg["lab"], cfg["axes"], cfg["h"], cfg["dt"], cfg["src_pts"],
                               cfg["rec_groups"], coeffs=cfg["coeffs"], sponge_width=sponge,
                               damp_mask=cfg["damp_mask"], mat_tables=cfg["mat_tables"])
    cp = None
    if boundary == "cpml":
        c = cpml.Cpml(prep["mats"].shape, cfg["sponge_width"], cfg["h"], cfg["dt"],
                      cpml.vmax_table(prep["Ct"], prep["rho_t"]),
                      cpml.dominant_frequency(cfg["wavelet"], cfg["dt"]))
        cp = (c.width, c.prof)
    return O.OracleShot(prep["mats"], prep["Ct"], prep["rho_t"], prep["damp"], prep["coeffs"],
                        cfg["h"], cfg["dt"], (prep["src_idx"], prep["src_w"]),
                        (prep["rec_ptr"], prep["rec_idx"], prep["rec_w"]), precision="f32",
                        cpml=cp)


def make(name):
    which, boundary, nt = CASES[name]
    t0 = time.time()
    cfg = config(which, nt)
    osh = oracle_shot(cfg, boundary)
    tr = np.zeros((nt, len(cfg["rec_groups"])), np.float32)
    chunk = 100
    for n0 in range(0, nt, chunk):
        n1 = min(nt, n0 + chunk)
        osh.run(cfg["wavelet"], n0, n1, tr)
        print("[%s] %d / %d steps, %.0f s" % (name, n1, nt, time.time() - t0), flush=True)
    os.makedirs(OUT, exist_ok=True)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), trace=tr, dt=cfg["dt"], h=cfg["h"],
                        shape=np.array(cfg["lab"].shape), lab_crc=label_crc(cfg["lab"]),
                        steps=nt, boundary=boundary)
    print("[%s] written, peak |p| %.3e" % (name, float(np.abs(tr).max())), flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or list(CASES):
        make(name)
