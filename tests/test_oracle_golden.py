"""The CPU oracle against the upstream float32 solver archives, the NumPy
array-path restatement and closed-form behaviour."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import numpy_forward as NF
from tests import helpers as H
from tests.golden import recipes as R


@pytest.mark.parametrize("med", ["iso", "ani"])
def test_oracle_reproduces_green_archive_bitwise(oracle_lib, med):
    """Full 275-step runs at ppw 4 (N = 121) match the upstream CUDA output
    bit for bit, which pins the stencils, tap order, moved-strain rule,
    deferred damping, sponge, source and pressure record at once."""
    case = R.green_case(med, 4)
    inp = H.fixture_inputs(case)
    tr = oracle_lib.run_forward(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"],
                                inp["coeffs"], inp["h"], inp["dt"], inp["nt"], inp["src"],
                                inp["rec"], inp["wavelet"])
    E = case["expected"].astype(np.float32)
    assert np.abs(E).max() > 1e-3
    assert np.array_equal(tr, E)


def test_planted_mistake_is_caught(oracle_lib):
    """A plausible mistake (one coefficient off by one float32 ulp, or a
    sign flip of c_3) moves the archive comparison off bit-exactness."""
    case = R.green_case("ani", 4)
    inp = H.fixture_inputs(case)
    nt = 120
    co = np.array(inp["coeffs"], float)
    co[3] = -co[3]
    tr = oracle_lib.run_forward(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], co,
                                inp["h"], inp["dt"], nt, inp["src"], inp["rec"], inp["wavelet"])
    E = case["expected"][:nt].astype(np.float32)
    assert H.rel_l2(tr, E) > 1e-4


def _array_path_case(seed, shape=(9, 10, 11)):
    inp = H.random_medium(shape, n_grain=5, seed=seed, sponge=0, fluid_pad=1, with_mask=False)
    inp["damp"] = None
    return inp


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_f64_matches_array_path(oracle_lib, seed):
    """The label-path oracle (f64) equals the upstream array-path scheme
    restated in NumPy (different evaluation order) to rounding."""
    inp = _array_path_case(seed)
    nt = 25
    tr_o = oracle_lib.run_forward(inp["mats"], inp["Ct"], inp["rho_t"], None, inp["coeffs"],
                                  inp["h"], inp["dt"], nt, inp["src"], inp["rec"],
                                  inp["wavelet"], precision="f64")
    ap = NF.ArrayPathShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["coeffs"], inp["h"],
                          inp["dt"], inp["src"], inp["rec"])
    tr_a = ap.run(inp["wavelet"], nt)
    assert np.abs(tr_a).max() > 0
    assert H.rel_l2(tr_o, tr_a) < 1e-12


def test_oracle_f32_close_to_f64(oracle_lib):
    inp = H.random_medium((14, 15, 16), seed=4)
    t32 = oracle_lib.run_forward(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"],
                                 inp["coeffs"], inp["h"], inp["dt"], inp["nt"], inp["src"],
                                 inp["rec"], inp["wavelet"], precision="f32")
    t64 = oracle_lib.run_forward(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"],
                                 inp["coeffs"], inp["h"], inp["dt"], inp["nt"], inp["src"],
                                 inp["rec"], inp["wavelet"], precision="f64")
    assert H.rel_l2(t32, t64) < 1e-5


def _homogeneous_shot(oracle_lib, shape, axis, damp=None):
    C6 = R.grain_C(axis)
    Ct = np.tile(np.asarray(C6, np.float64).reshape(1, 36), (2, 1))
    return oracle_lib.OracleShot(np.ones(shape, np.int32), Ct, np.full(2, 917.0), damp,
                                 R.coeffs_kh2(), 1e-4, 1e-8,
                                 (np.zeros(0, np.int64), np.zeros(0)),
                                 (np.array([0], np.int32), np.zeros(0, np.int64), np.zeros(0)),
                                 precision="f64")


def test_uniform_stress_has_no_divergence(oracle_lib):
    """Closed form: a spatially uniform stress has zero divergence; every
    stencil term is a difference of two in-range taps, so velocities stay
    exactly zero everywhere, edges included."""
    sh = _homogeneous_shot(oracle_lib, (12, 12, 12), [0.3, 0.4, 0.866])
    for k in range(6):
        sh.s[k][...] = 1.0 + k
    sh.velocity()
    for a in sh.v:
        assert np.all(a == 0.0)


def test_linear_stress_gives_closed_form_velocity(oracle_lib):
    """s1 = x (in cells): the staggered x-derivative is sum_j c_j (2j+1) = 1
    wherever all four taps exist, so one velocity half-step gives exactly
    vx = (dt/h) / rho there and vy = vz = 0 everywhere."""
    shape = (6, 7, 20)
    sh = _homogeneous_shot(oracle_lib, shape, [0.0, 0.0, 1.0])
    sh.s[0][...] = np.arange(shape[2], dtype=float)[None, None, :]
    sh.velocity()
    c = R.coeffs_kh2()
    nx = shape[2]
    for ix in range(nx):
        # term j present iff j <= ix < nx - j - 1 (edge truncation)
        d = sum(c[j] * (2 * j + 1) for j in range(4) if j <= ix < nx - j - 1)
        assert np.allclose(sh.v[0][:, :, ix], (1e-8 / 1e-4) * d / 917.0, rtol=1e-13, atol=0)
    assert np.all(sh.v[1] == 0) and np.all(sh.v[2] == 0)


def test_damping_only_decays(oracle_lib):
    """With damp < 1 everywhere and no source, each step scales s by damp
    before adding dt*C:q; zero velocity keeps q = 0, so s decays geometrically."""
    shape = (6, 7, 8)
    sh = _homogeneous_shot(oracle_lib, shape, [0, 0, 1], damp=np.full(shape, 0.5))
    sh.s[0][...] = 3.0
    sh.stress()
    assert np.all(sh.s[0] == 1.5)
