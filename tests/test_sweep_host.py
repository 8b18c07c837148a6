"""Host side of the azimuth sweep: the rotated re-grid, its frame convention
and the rim transducer, against what geometry fixes."""
from __future__ import annotations

import numpy as np
import pytest

from pulse_echo_azimuthal_simulation_amd import sweep


@pytest.fixture(scope="module")
def build():
    return sweep.disk_build(6.0e-3, 2.0e-3, 2.0e-4, 25, seed=3)


def test_zero_rotation_is_identity(build):
    lab, nd, m, axes = sweep.rotated_grid(build, build["h"], 4, 0.0)
    core = lab[m:lab.shape[0] - m, m:m + nd, m:m + nd]
    assert np.array_equal(core, build["labels"].transpose(2, 1, 0))
    assert np.allclose(axes, build["axes"], atol=1e-15)
    assert (lab[:m] == -1).all() and (lab[:, :m] == -1).all() and (lab[:, :, :m] == -1).all()


def test_quarter_turn_is_a_permutation(build):
    """At 90 deg the nearest-neighbour re-grid is exact: solver column (i, j)
    reads source column (nd-1-j, i)."""
    lab, nd, m, _ = sweep.rotated_grid(build, build["h"], 4, 90.0)
    core = lab[m:lab.shape[0] - m, m:m + nd, m:m + nd].transpose(2, 1, 0)   # (i, j, z)
    src = build["labels"]
    i, j = np.meshgrid(np.arange(nd), np.arange(nd), indexing="ij")
    assert np.array_equal(core, src[nd - 1 - j, i])


@pytest.mark.parametrize("rot", [17.0, 45.0, 130.0])
def test_rotation_is_rigid(build, rot):
    """Geometry and c-axes rotate together ('rigid2'): every grain's centroid
    moves by the same rotation that takes its axis a to axes_rot = a @ R."""
    lab, nd, m, axes = sweep.rotated_grid(build, build["h"], 4, rot)
    core = lab[m:lab.shape[0] - m, m:m + nd, m:m + nd].transpose(2, 1, 0)
    a = np.radians(rot)
    R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
    assert np.allclose(axes, build["axes"] @ R)
    h = build["h"]
    c = (np.arange(nd) - (nd - 1) / 2) * h
    src = build["labels"]
    for g in range(0, 25, 5):
        p0 = np.argwhere(src == g)
        p1 = np.argwhere(core == g)
        if len(p0) < 30 or len(p1) < 30:
            continue
        x0 = np.array([c[p0[:, 0]].mean(), c[p0[:, 1]].mean()])
        x1 = np.array([c[p1[:, 0]].mean(), c[p1[:, 1]].mean()])
        # solver point x reads specimen point R(+rot) x: x1 = R(-rot) x0
        assert np.allclose(R[:2, :2].T @ x0, x1, atol=1.5 * h)


def test_single_raster_at_zero_rotation_matches_build(build):
    lab, nd, m, _ = sweep.rotated_grid(build, build["h"], 4, 0.0, single_raster=True)
    core = lab[m:lab.shape[0] - m, m:m + nd, m:m + nd]
    ref = build["labels"].transpose(2, 1, 0)
    assert (core != ref).mean() < 1e-3


def test_rim_footprint_brute_force(build):
    lab, nd, m, _ = sweep.rotated_grid(build, build["h"], 4, 33.0)
    pts, w = sweep.rim_footprint(lab, nd, build["h"], element_d_mm=1.2)
    nz, ny = lab.shape[:2]
    cz, cy = nz // 2, ny // 2
    row = lab[cz, cy]
    outer = max(x for x in range(ny // 2, ny) if row[x] >= 0 and x < ny // 2 + nd // 2)
    ixp = outer - 1
    er = max(int(1.2e-3 / 2 / build["h"]), 1)
    ref = sorted((z, y, ixp) for z in range(nz) for y in range(ny)
                 if (z - cz) ** 2 + (y - cy) ** 2 <= er * er and lab[z, y, ixp] >= 0)
    assert sorted(pts) == ref and w == pytest.approx(1.0 / len(ref))


def test_rotated_tables_equal_serial(build):
    from pulse_echo_azimuthal_simulation_amd import materials
    azs = [0.0, 10.0, 95.0]
    par = sweep.rotated_tables(build["axes"], azs, workers=2)
    for az, (Ct, rho) in zip(azs, par):
        _, _, _, axes = sweep.rotated_grid(build, build["h"], 4, az)
        C2, r2 = materials.material_tables(axes)
        assert np.array_equal(Ct, C2) and np.array_equal(rho, r2)
