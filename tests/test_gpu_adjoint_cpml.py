"""GPU adjoint of CPML shots (both kernel families, through the C ABI) against the
CPU oracle's reverse step with its CPML transposes (pinned by dot-product
identities and central differences in tests/test_oracle_adjoint.py) and
against the identities it must satisfy on the device itself.

A CPML step's state holds the memory variables; the reverse pass carries
their adjoints (DeviceAdjoint.psib), checkpoints hold the memory variables,
and the stiffness gradient reads the strain history (the strains depend on
the memory variables as well as on v).

Bars: f64 gradients and wavelet gradients equal the f64 oracle's to 1e-9
relative (summation order only); the one-step dot-product identity holds to
1e-11 in f64 with random states, memory variables and profiles; f32
gradients are within 2e-4 of the f64 oracle's; central differences of the
misfit in table entries (f64) match the device gradient to 1e-5.
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


def _case(name):
    print("\n[gpu case] %s" % name, flush=True)


def _medium(shape=(14, 15, 16), seed=4, nt=40, L=3, kappa=1.5):
    from pulse_echo_azimuthal_simulation_amd import cpml
    inp = H.random_medium(shape, n_grain=6, seed=seed, sponge=0, fluid_pad=1)
    inp["nt"] = nt
    inp["wavelet"] = np.resize(inp["wavelet"], nt) if len(inp["wavelet"]) < nt else inp["wavelet"]
    rng = np.random.default_rng(seed + 100)
    Ct_obs = np.asarray(inp["Ct"], float).reshape(-1, 36).copy()
    Ct_obs[1:] *= 1.0 + 0.05 * rng.uniform(-1, 1, Ct_obs[1:].shape)
    inp["Ct_obs"] = Ct_obs
    inp["cpml"] = cpml.Cpml(inp["shape"], L, inp["h"], inp["dt"],
                            cpml.vmax_table(inp["Ct"], inp["rho_t"]), 2e6, kappa_max=kappa)
    return inp


def _oracle_grad(inp, record="pressure", precision="f64"):
    from oracle import oracle as O
    cp = (inp["cpml"].width, inp["cpml"].prof)
    obs = O.OracleShot(inp["mats"], inp["Ct_obs"], inp["rho_t"], inp["damp"], inp["coeffs"],
                       inp["h"], inp["dt"], inp["src"], inp["rec"], record, precision, cpml=cp)
    d = obs.run(inp["wavelet"], 0, inp["nt"])
    sh = O.OracleShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                      inp["h"], inp["dt"], inp["src"], inp["rec"], record, precision, cpml=cp)
    J, gC, gw, tr = O.adjoint_gradient(sh, inp["wavelet"], inp["nt"],
                                       lambda t: (0.5 * float(np.sum((t - d) ** 2)), t - d))
    return J, gC, gw, d


def _device_shot(inp, precision, record="pressure", Ct=None, kernel="direct"):
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    return DeviceShot(inp["mats"], inp["Ct"] if Ct is None else Ct, inp["rho_t"], inp["damp"],
                      inp["coeffs"], inp["h"], inp["dt"], inp["src"][0], inp["src"][1],
                      *inp["rec"], record=record, precision=precision, kernel=kernel,
                      cpml=inp["cpml"])


def _gpu_grad(torch, inp, d, precision="f64", record="pressure", every=7, history_bytes=0,
              kernel="direct"):
    from pulse_echo_azimuthal_simulation_amd import adjoint
    sh = _device_shot(inp, precision, record, kernel=kernel)
    dd = torch.as_tensor(d, device=sh.device).to(sh.real)

    def resid(tr):
        r = tr - dd
        return 0.5 * float((r.double() ** 2).sum()), r

    adj = adjoint.DeviceAdjoint(sh, hist="strain")
    J, g, tr, gw = adjoint.misfit_gradient(sh, inp["wavelet"], inp["nt"], resid,
                                           checkpoint_every=every, history_bytes=history_bytes,
                                           with_wavelet=True, adjoint=adj)
    torch.cuda.synchronize()
    return J, g.cpu().numpy(), gw.cpu().numpy()


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("record", ["pressure", "vx", "vz"])
def test_cpml_f64_gradient_matches_oracle(torch_gpu, record, kernel):
    _case("cpml adjoint f64 %s %s" % (kernel, record))
    inp = _medium()
    J0, g0, gw0, d = _oracle_grad(inp, record)
    J, g, gw = _gpu_grad(torch_gpu, inp, d, "f64", record, kernel=kernel)
    assert abs(J - J0) <= 1e-12 * J0
    assert np.abs(g0).max() > 0
    assert H.rel_l2(g, g0) <= 1e-9, H.rel_l2(g, g0)
    assert H.rel_l2(gw, gw0) <= 1e-9, H.rel_l2(gw, gw0)


def test_cpml_history_kept_and_recomputed_agree(torch_gpu):
    """Every segment recomputed from its checkpoint (memory variables
    restored) against every segment's history kept from the first pass."""
    _case("cpml adjoint history paths")
    inp = _medium(seed=6)
    _, _, _, d = _oracle_grad(inp)
    _, g_a, gw_a = _gpu_grad(torch_gpu, inp, d, "f64", every=9, history_bytes=0)
    _, g_b, gw_b = _gpu_grad(torch_gpu, inp, d, "f64", every=9, history_bytes=2 ** 34)
    assert np.array_equal(g_a, g_b)
    assert np.array_equal(gw_a, gw_b)


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("shape,L,seed", [((14, 15, 16), 3, 0), ((12, 13, 17), 4, 1),
                                          ((30, 36, 70), 5, 2)])
def test_cpml_one_step_dot_product_on_gpu(torch_gpu, shape, L, seed, kernel):
    """<L x, y> = <x, L^T y> for one device step from a random state (fields
    and memory variables) and random adjoints, f64, with random profiles."""
    _case("cpml adjoint dot product %s %s L %d" % (kernel, shape, L))
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint
    inp = _medium(shape, seed, nt=1, L=L)
    rng = np.random.default_rng(seed)
    cp = inp["cpml"]
    for a in range(3):
        cp.prof[a] = rng.uniform(0.2, 1.2, size=np.shape(cp.prof[a]))
    inp["damp"] = rng.uniform(0.6, 1.0, inp["shape"])
    sh = _device_shot(inp, "f64", kernel=kernel)
    x_v = [rng.normal(size=inp["shape"]) for _ in range(3)]
    x_s = [rng.normal(size=inp["shape"]) for _ in range(6)]
    x_p = [rng.normal(size=tuple(p.shape)) for p in sh.psi]
    for k in range(3):
        sh.v[k].copy_(torch.as_tensor(x_v[k]))
    for k in range(6):
        sh.s[k].copy_(torch.as_tensor(x_s[k]))
    for a in range(3):
        sh.psi[a].copy_(torch.as_tensor(x_p[a]))
    sh.run(sh.wavelet(np.zeros(1)), 0, 1)
    torch.cuda.synchronize()
    y_v = [rng.normal(size=inp["shape"]) for _ in range(3)]
    y_s = [rng.normal(size=inp["shape"]) for _ in range(6)]
    y_p = [rng.normal(size=tuple(p.shape)) for p in sh.psi]
    lhs = sum(np.vdot(sh.v[k].cpu().numpy(), y_v[k]) for k in range(3)) + \
        sum(np.vdot(sh.s[k].cpu().numpy(), y_s[k]) for k in range(6)) + \
        sum(np.vdot(sh.psi[a].cpu().numpy(), y_p[a]) for a in range(3))
    adj = adjoint.DeviceAdjoint(sh, gradient=False, hist="strain")
    adj.reset()
    for k in range(3):
        adj.lv[k].copy_(torch.as_tensor(y_v[k]))
    for k in range(6):
        adj.ls[k].copy_(torch.as_tensor(y_s[k]))
    for a in range(3):
        adj.psib[a].copy_(torch.as_tensor(y_p[a]))
    adj.run(1, 0)
    torch.cuda.synchronize()
    rhs = sum(np.vdot(x_v[k], adj.lv[k].cpu().numpy()) for k in range(3)) + \
        sum(np.vdot(x_s[k], adj.ls[k].cpu().numpy()) for k in range(6)) + \
        sum(np.vdot(x_p[a], adj.psib[a].cpu().numpy()) for a in range(3))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
def test_cpml_f32_gradient_close_to_f64_oracle(torch_gpu, kernel):
    _case("cpml adjoint f32 %s" % kernel)
    inp = _medium((16, 18, 20), seed=2)
    _, g0, gw0, d = _oracle_grad(inp)
    _, g, gw = _gpu_grad(torch_gpu, inp, d, "f32", every=11, kernel=kernel)
    assert H.rel_l2(g, g0) <= 2e-4
    assert H.rel_l2(gw, gw0) <= 2e-4


def test_cpml_gradient_matches_central_differences(torch_gpu):
    """dJ/dCt of the device adjoint against central differences of the
    device forward's misfit (f64), entries of grains and of the couplant."""
    _case("cpml adjoint central differences")
    torch = torch_gpu
    inp = _medium(seed=8, nt=60)
    _, _, _, d = _oracle_grad(inp)
    _, g, _ = _gpu_grad(torch, inp, d, "f64")
    Ct0 = np.asarray(inp["Ct"], float).reshape(-1, 36)

    def J_of(Ct):
        sh = _device_shot(inp, "f64", Ct=Ct)
        tr = sh.new_traces(inp["nt"])
        sh.run(sh.wavelet(inp["wavelet"]), 0, inp["nt"], tr)
        return 0.5 * float(np.sum((tr.cpu().numpy() - d) ** 2))

    scale = np.abs(Ct0).max()
    for m, k in [(1, 0), (2, 7), (3, 21), (0, 0), (4, 35)]:
        hstep = 1e-4 * scale
        Cp, Cm = Ct0.copy(), Ct0.copy()
        Cp[m, k] += hstep
        Cm[m, k] -= hstep
        fd = (J_of(Cp) - J_of(Cm)) / (2 * hstep)
        tol = 1e-6 * np.abs(g).max()
        assert abs(fd - g[m, k]) <= max(tol, 1e-5 * abs(fd)), (m, k, fd, g[m, k])


def test_cpml_autograd_matches_misfit_gradient(torch_gpu):
    _case("cpml adjoint autograd (tiled)")
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint
    inp = _medium(seed=5)
    _, g0, _, d = _oracle_grad(inp)
    sh = _device_shot(inp, "f64", kernel="tiled")
    dd = torch.as_tensor(d, dtype=torch.float64)
    t = torch.tensor(np.asarray(inp["Ct"], float).reshape(-1, 36), requires_grad=True)
    tr = adjoint.traces_from_table(t, sh, inp["wavelet"], inp["nt"], checkpoint_every=8)
    J = 0.5 * ((tr.double().cpu() - dd) ** 2).sum()
    J.backward()
    assert H.rel_l2(t.grad.numpy(), g0) <= 1e-9
