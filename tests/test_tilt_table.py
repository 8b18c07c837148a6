"""The Table-2 tilt testbed's echo measurement (sim/validation/
tilt_testbed.py:229-243), pinned on the CPU by the upstream's own numbers:
the naive-treatment tilted-bicrystal traces stored for tilts 0, 30 and 45
deg (sim/results/rsg_twoway_ssg_ppw{6,8,10}.npz, the receiver at the source)
give exactly the echo amplitudes stored in sim/results/tilt_table.npz for
those tilts.  The GPU reproduction of the whole table is
test_gpu_tilt_table.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.golden import recipes as R


def test_qp_speed_closed_form():
    m = R.materials.ICE
    assert R.qp_speed_ice(0.0) == pytest.approx(np.sqrt(m["C33"] / m["rho"]), rel=1e-15)
    assert R.qp_speed_ice(np.pi / 2) == pytest.approx(np.sqrt(m["C11"] / m["rho"]), rel=1e-15)
    # against the Christoffel eigenvalue of the rotated 3-D tensor
    for psi in (0.3, 0.8867, 1.2):
        C = R.grain_C((np.sin(psi), 0.0, np.cos(psi)))
        T = R.materials.voigt6_to_tensor(C)
        G = np.einsum("ijkl,j,l->ik", T, np.array([0, 0, 1.0]), np.array([0, 0, 1.0]))
        v = np.sqrt(np.linalg.eigvalsh(G).max() / m["rho"])
        assert R.qp_speed_ice(psi) == pytest.approx(v, rel=1e-12)


@pytest.mark.parametrize("ppw", [6, 8, 10])
def test_echo_amp_reproduces_stored_table(ppw):
    ref = R.load("rsg_twoway_ssg_ppw%d.npz" % ppw)
    table = R.load("tilt_table.npz")
    amps = table["naive_ppw%d_amp" % ppw]
    v = R.qp_speed_ice(0.0)
    for theta in (0, 30, 45):
        k = "%d" % theta
        a = R.echo_amp(ref["bi_" + k][:, 0], ref["ho_" + k][:, 0], float(ref["dt_" + k]),
                       float(ref["d_" + k]), v)
        assert a == pytest.approx(float(amps[R.TILTS.index(float(theta))]), rel=1e-12)
        assert float(ref["d_" + k]) == R.twoway_geometry(float(theta), float(ppw))["d"]
