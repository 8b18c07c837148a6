"""Copy the upstream's full-disk reference traces into tests/golden/reference/
as data:

* sim/ref/fw_fabric{00,25,50,100}.npz -- 2 MHz label-path traces of
  ladder.standard_build() with the four fabric sets of
  ladder.standard_fabrics, written by sim/fw_checks/fw_reference_2mhz.py
  (keys trace, dt, f0, E1, level_db, frac, h, order, ppw, seed_specimen,
  seed_fabric, fluid_damp);
* sim/ref/fw_az{17,31,47,62,76,90}.npz -- the same disk rotated by
  rotation_test.rotated_grid, written by sim/pipeline/azimuth_sweep.py:
  104-121 (keys trace, dt, f0, rot_deg, level_db = clean-window envelope
  RMS over E1 in dB, fluid_damp);
* sim/results/tilt_table.npz -- Table 2 of the manuscript
  (sim/runs/tilt_table_run.py): per treatment and ppw the mean absolute
  tilt error (dB), the per-tilt errors and the echo amplitudes.

    python tests/golden/copy_fulldisk_traces.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/sim/ref"


def main():
    names = ["fw_fabric%s" % t for t in ("00", "25", "50", "100")]
    names += ["fw_az%02d" % r for r in (17, 31, 47, 62, 76, 90)]
    srcs = [os.path.join(SRC, n + ".npz") for n in names]
    srcs.append("/root/reference/sim/results/tilt_table.npz")
    for src in srcs:
        d = np.load(src, allow_pickle=False)
        np.savez_compressed(os.path.join(HERE, "reference", os.path.basename(src)),
                            **{k: d[k] for k in d.files})


if __name__ == "__main__":
    main()
