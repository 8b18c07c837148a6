"""Copy the upstream's full-disk reference traces into tests/golden/reference/
as data: sim/ref/fw_fabric{00,25,50,100}.npz (2 MHz label-path traces of
ladder.standard_build() with the four fabric sets of ladder.standard_fabrics,
written by sim/fw_checks/fw_reference_2mhz.py; keys trace, dt, f0, E1,
level_db, frac, h, order, ppw, seed_specimen, seed_fabric, fluid_damp).

    python tests/golden/copy_fulldisk_traces.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/sim/ref"


def main():
    for tag in ("00", "25", "50", "100"):
        d = np.load(os.path.join(SRC, "fw_fabric%s.npz" % tag), allow_pickle=False)
        np.savez_compressed(os.path.join(HERE, "reference", "fw_fabric%s.npz" % tag),
                            **{k: d[k] for k in d.files})


if __name__ == "__main__":
    main()
