"""Write tests/golden/specimen/disk_s11_p8_k-8.npz from the upstream's stored
specimen build sim/analysis/facet_model/specimen_build_s11_p8_k-8.npz (the
recipe of sim/analysis/facet_model/range_domain_test.py:56 with
range_domain_common.py:37-53: DiskSpecimen(0.100 m, 0.035 m, 100 grains,
size_cv 0.35, kappa -8, spatial_corr 0, axis (1, 0, 0), seed 11).build(h)
with h = 3850 / 2 MHz / 8).

The fixture keeps the build's seeds, weights and axes as they are, and of
the 416 x 416 x 146 label volume its CRC-32, the voxel count of every label
and every 8th cell along each axis -- data only.

    python tests/golden/make_specimen_fixture.py
"""
import os
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/sim/analysis/facet_model/specimen_build_s11_p8_k-8.npz"


def main():
    d = np.load(SRC, allow_pickle=False)
    lab = np.ascontiguousarray(d["labels"], np.int16)
    out = os.path.join(HERE, "specimen")
    os.makedirs(out, exist_ok=True)
    np.savez_compressed(os.path.join(out, "disk_s11_p8_k-8.npz"),
                        seeds=d["seeds"], weights=d["weights"], axes=d["axes"], h=d["h"],
                        shape=np.array(lab.shape), label_crc=zlib.crc32(lab.tobytes()) & 0xFFFFFFFF,
                        counts=np.bincount(lab[lab >= 0].ravel().astype(np.int64)),
                        sub=lab[::8, ::8, ::8])


if __name__ == "__main__":
    main()
