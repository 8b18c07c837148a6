"""GPU adjoint (pe_adjoint_run / pe_adjoint_gradient, through the C ABI)
against the CPU oracle's reverse step and against the identities it must
satisfy.

Bars: float64 gradients equal the f64 oracle's (scatter-form, independent
restatement) to 1e-9 relative (different summation order only); float32
gradients equal the f64 oracle's to 2e-4 relative (the f32 oracle shows the
same distance); dot-product identities hold to 1e-11 in f64; repeated runs
are bitwise identical (deterministic reductions).
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


def _medium(shape=(12, 14, 16), seed=4, nt=40, n_grain=6):
    inp = H.random_medium(shape, n_grain=n_grain, seed=seed, sponge=2)
    inp["nt"] = nt
    inp["wavelet"] = np.resize(inp["wavelet"], nt) if len(inp["wavelet"]) < nt else inp["wavelet"]
    rng = np.random.default_rng(seed + 100)
    Ct_obs = np.asarray(inp["Ct"], float).reshape(-1, 36).copy()
    Ct_obs[1:] *= 1.0 + 0.05 * rng.uniform(-1, 1, Ct_obs[1:].shape)
    inp["Ct_obs"] = Ct_obs
    return inp


def _oracle_grad(inp, record="pressure", precision="f64"):
    from oracle import oracle as O
    obs = O.OracleShot(inp["mats"], inp["Ct_obs"], inp["rho_t"], inp["damp"], inp["coeffs"],
                       inp["h"], inp["dt"], inp["src"], inp["rec"], record, precision)
    d = obs.run(inp["wavelet"], 0, inp["nt"])
    sh = O.OracleShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                      inp["h"], inp["dt"], inp["src"], inp["rec"], record, precision)
    J, gC, gw, tr = O.adjoint_gradient(sh, inp["wavelet"], inp["nt"],
                                       lambda t: (0.5 * float(np.sum((t - d) ** 2)), t - d))
    return J, gC, gw, d


def _device_shot(inp, precision, kernel, record="pressure"):
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    return DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"],
                      inp["h"], inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"],
                      record=record, precision=precision, kernel=kernel)


def _gpu_grad(torch, inp, d, precision="f64", kernel="tiled", record="pressure", every=7,
              n_proj=0, proj=None, history_bytes=0, zbreak=0):
    from pulse_echo_azimuthal_simulation_amd import adjoint
    sh = _device_shot(inp, precision, kernel, record)
    dd = torch.as_tensor(d, device=sh.device).to(sh.real)

    def resid(tr):
        r = tr - dd
        return 0.5 * float((r.double() ** 2).sum()), r

    adj = adjoint.DeviceAdjoint(sh, n_proj=n_proj, proj=proj, zbreak=zbreak)
    J, g, tr, gw = adjoint.misfit_gradient(sh, inp["wavelet"], inp["nt"], resid,
                                           checkpoint_every=every, history_bytes=history_bytes,
                                           with_wavelet=True, adjoint=adj)
    torch.cuda.synchronize()
    return J, g.cpu().numpy(), gw.cpu().numpy()


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("record", ["pressure", "vx", "vz"])
def test_f64_gradient_matches_oracle(torch_gpu, kernel, record):
    _case("adjoint f64 %s %s" % (kernel, record))
    inp = _medium()
    J0, g0, gw0, d = _oracle_grad(inp, record)
    J, g, gw = _gpu_grad(torch_gpu, inp, d, "f64", kernel, record)
    assert abs(J - J0) <= 1e-12 * J0
    assert np.abs(g0).max() > 0
    assert H.rel_l2(g, g0) <= 1e-9
    assert H.rel_l2(gw, gw0) <= 1e-9


@pytest.mark.parametrize("shape,seed", [((10, 9, 11), 1), ((16, 12, 20), 2), ((9, 17, 13), 3)])
def test_f32_gradient_close_to_f64_oracle(torch_gpu, shape, seed):
    _case("adjoint f32 %s seed %d" % (shape, seed))
    inp = _medium(shape, seed)
    _, g0, gw0, d = _oracle_grad(inp)
    _, g, gw = _gpu_grad(torch_gpu, inp, d, "f32", "tiled", every=11)
    assert H.rel_l2(g, g0) <= 2e-4
    assert H.rel_l2(gw, gw0) <= 2e-4


def test_low_order_gradient(torch_gpu):
    _case("adjoint f64 order 4")
    inp = H.random_medium((11, 12, 13), n_grain=5, seed=7, order=4, sponge=2)
    rng = np.random.default_rng(7)
    Ct_obs = np.asarray(inp["Ct"], float).reshape(-1, 36).copy()
    Ct_obs[1:] *= 1.0 + 0.05 * rng.uniform(-1, 1, Ct_obs[1:].shape)
    inp["Ct_obs"] = Ct_obs
    _, g0, gw0, d = _oracle_grad(inp)
    _, g, gw = _gpu_grad(torch_gpu, inp, d, "f64", "direct")
    assert H.rel_l2(g, g0) <= 1e-9
    assert H.rel_l2(gw, gw0) <= 1e-9


def test_projected_gradient_equals_projection_of_full(torch_gpu):
    _case("adjoint projections")
    inp = _medium(seed=5)
    _, _, _, d = _oracle_grad(inp)
    rng = np.random.default_rng(9)
    n_mat = len(inp["rho_t"])
    P = rng.normal(size=(n_mat, 3, 36))
    _, g, _ = _gpu_grad(torch_gpu, inp, d, "f64")
    _, gp, _ = _gpu_grad(torch_gpu, inp, d, "f64", n_proj=3, proj=P)
    ref = np.einsum("mpk,mk->mp", P, g)
    assert H.rel_l2(gp, ref) <= 1e-10


def test_deterministic_and_history_paths_agree(torch_gpu):
    """Two runs give bitwise equal gradients; storing the history in the
    first pass (no recompute) gives the same bits as recomputing it; cutting
    runs every 5 planes regroups the sums only."""
    _case("adjoint determinism")
    inp = _medium(seed=6)
    _, _, _, d = _oracle_grad(inp)
    _, g1, w1 = _gpu_grad(torch_gpu, inp, d, "f32", every=9)
    _, g2, w2 = _gpu_grad(torch_gpu, inp, d, "f32", every=9)
    _, g3, w3 = _gpu_grad(torch_gpu, inp, d, "f32", every=9, history_bytes=1 << 40)
    _, g4, _ = _gpu_grad(torch_gpu, inp, d, "f32", every=9, zbreak=5)
    assert np.array_equal(g1, g2) and np.array_equal(w1, w2)
    assert np.array_equal(g1, g3) and np.array_equal(w1, w3)
    assert H.rel_l2(g4, g1) <= 1e-12


def test_wavelet_dot_product_on_gpu(torch_gpu):
    """<F w, r> = <w, F^T r> with the GPU forward and adjoint (f64)."""
    _case("adjoint dot product")
    from pulse_echo_azimuthal_simulation_amd import adjoint
    inp = _medium(seed=8)
    sh = _device_shot(inp, "f64", "tiled")
    nt = inp["nt"]
    wav = sh.wavelet(inp["wavelet"])
    tr = sh.new_traces(nt)
    sh.run(wav, 0, nt, tr)
    r = torch_gpu.as_tensor(np.random.default_rng(1).normal(size=tuple(tr.shape)),
                            device=sh.device)
    lhs = float((tr * r).sum())
    adj = adjoint.DeviceAdjoint(sh, gradient=False)
    adj.reset()
    gw = torch_gpu.zeros(nt, dtype=torch_gpu.float64, device=sh.device)
    adj.run(nt, 0, None, r, gw)
    rhs = float(np.dot(inp["wavelet"][:nt], gw.cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_autograd_function_matches_misfit_gradient(torch_gpu):
    """traces_from_table + a torch misfit + backward() gives the adjoint's
    table gradient; chaining through the torch Bond table gives the angle
    gradient the projection reduction returns."""
    _case("adjoint autograd")
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint, fabric
    inp = _medium(seed=10)
    theta, phi = fabric.angles_from_axes(inp["axes"])
    Ct, rho = fabric.stiffness_table(theta, phi)
    inp = dict(inp, Ct=Ct, rho_t=rho)
    _, _, _, d = _oracle_grad(dict(inp, Ct_obs=inp["Ct_obs"]), precision="f64")
    sh = _device_shot(inp, "f64", "tiled")
    tt = torch.tensor(theta, requires_grad=True)
    pp = torch.tensor(phi, requires_grad=True)
    table = fabric.stiffness_table_torch(tt, pp)
    tr = adjoint.traces_from_table(table, sh, inp["wavelet"], inp["nt"], checkpoint_every=13)
    dd = torch.as_tensor(d, device=tr.device)
    J = 0.5 * ((tr.double() - dd) ** 2).sum()
    J.backward()
    P = fabric.angle_projections(theta, phi)
    _, gp, _ = _gpu_grad(torch, inp, d, "f64", n_proj=2, proj=P, every=13)
    assert H.rel_l2(tt.grad.numpy(), gp[1:, 0]) <= 1e-9
    assert H.rel_l2(pp.grad.numpy(), gp[1:, 1]) <= 1e-9


def test_autograd_two_forwards_then_backward_of_first(torch_gpu):
    """Two forwards through one shot with different tables (a line search),
    then backward of the first: the gradient is the first table's, because
    the backward restores the table its forward saw."""
    _case("adjoint autograd two forwards")
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint
    inp = _medium(seed=12)
    _, _, _, d = _oracle_grad(inp, precision="f64")
    _, g_ref, _ = _gpu_grad(torch, inp, d, "f64", every=9)
    sh = _device_shot(inp, "f64", "tiled")
    dd = torch.as_tensor(d, device=sh.device)
    t1 = torch.tensor(np.asarray(inp["Ct"], float).reshape(-1, 36), requires_grad=True)
    t2 = torch.tensor(np.asarray(inp["Ct_obs"], float).reshape(-1, 36), requires_grad=True)
    tr1 = adjoint.traces_from_table(t1, sh, inp["wavelet"], inp["nt"], checkpoint_every=9)
    tr2 = adjoint.traces_from_table(t2, sh, inp["wavelet"], inp["nt"], checkpoint_every=9)
    assert H.rel_l2(tr2.detach().cpu().numpy(), d) <= 1e-9
    J1 = 0.5 * ((tr1.double() - dd) ** 2).sum()
    J1.backward()
    assert H.rel_l2(t1.grad.numpy(), g_ref) <= 1e-9


def test_cpml_gradient_refuses_velocity_history(torch_gpu):
    """A CPML step's strains depend on the memory variables as well as on v:
    its gradient needs the strain history."""
    _case("adjoint cpml velocity-history refusal")
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint, cpml
    from pulse_echo_azimuthal_simulation_amd.solver import DeviceShot
    inp = H.random_medium((14, 16, 18), n_grain=4, seed=3, sponge=0)
    cp = cpml.Cpml(inp["shape"], 3, inp["h"], inp["dt"],
                   cpml.vmax_table(inp["Ct"], inp["rho_t"]), 2e6)
    sh = DeviceShot(inp["mats"], inp["Ct"], inp["rho_t"], inp["damp"], inp["coeffs"], inp["h"],
                    inp["dt"], inp["src"][0], inp["src"][1], *inp["rec"], precision="f64",
                    kernel="tiled", cpml=cp)
    with pytest.raises(ValueError, match="strain"):
        adjoint.DeviceAdjoint(sh, hist="velocity")


@pytest.mark.parametrize("precision,shape", [("f32", (40, 44, 150)), ("f64", (36, 40, 90)),
                                             ("f32", (23, 29, 71))])
def test_tiled_adjoint_matches_direct(torch_gpu, precision, shape):
    """The LDS-tiled reverse kernels against the one-thread-per-cell ones on
    grids large enough for their interior fast paths (and a ragged one):
    full table gradient, angle projections (tiled gradient kernel) and the
    wavelet gradient."""
    _case("adjoint tiled vs direct %s %s" % (precision, shape))
    from pulse_echo_azimuthal_simulation_amd import adjoint, fabric
    inp = H.random_medium(shape, n_grain=30, seed=11, sponge=4)
    theta, phi = fabric.angles_from_axes(inp["axes"])
    P = fabric.angle_projections(theta, phi)
    nt = 30
    rng = np.random.default_rng(3)
    out = {}
    for fam in ("direct", "tiled"):
        sh = _device_shot(inp, precision, "tiled")
        wav = sh.wavelet(inp["wavelet"])
        res = torch_gpu.as_tensor(rng.normal(size=(nt, sh.n_rec)) if fam == "direct" else
                                  out["direct"][3], device=sh.device).to(sh.real)
        gs = []
        for n_proj, proj in ((0, None), (2, P)):
            adj = adjoint.DeviceAdjoint(sh, n_proj=n_proj, proj=proj, kernel=fam)
            gw = torch_gpu.zeros(nt, dtype=torch_gpu.float64, device=sh.device)
            J, g, _, gw = adjoint.misfit_gradient(sh, wav, nt, lambda tr: (0.0, res),
                                                  checkpoint_every=8, with_wavelet=True,
                                                  adjoint=adj)
            gs.append(g.cpu().numpy())
        torch_gpu.cuda.synchronize()
        out[fam] = (gs[0], gs[1], gw.cpu().numpy(), res.cpu().numpy())
    tol = 1e-10 if precision == "f64" else 2e-4
    gd, pd, wd, _ = out["direct"]
    gt, pt, wt, _ = out["tiled"]
    assert np.abs(gd).max() > 0 and np.abs(pd).max() > 0
    assert H.rel_l2(gt, gd) <= tol
    assert H.rel_l2(pt, pd) <= tol
    assert H.rel_l2(wt, wd) <= tol


def test_angle_gradient_matches_central_differences(torch_gpu):
    """dJ/d(theta, phi) per grain from the checkpointed adjoint (projection
    reduction) equals central differences of the misfit computed with the
    GPU forward, float64, every grain of a small specimen.  h = 1e-4 rad: the
    central-difference truncation error falls as h^2 (measured on the CPU
    oracle for this case: 7.5e-2, 7.5e-4, 7.5e-6 relative at h = 1e-2,
    1e-3, 1e-4 on the worst component), so the bar is 1e-4."""
    _case("adjoint angle FD")
    torch = torch_gpu
    from pulse_echo_azimuthal_simulation_amd import adjoint, fabric, specimen
    inp = H.random_medium((24, 30, 32), n_grain=8, seed=12, sponge=3)
    nt = 150
    wav = np.resize(inp["wavelet"], nt)
    wav[len(inp["wavelet"]):] = 0.0
    theta, phi = fabric.angles_from_axes(inp["axes"])
    tt, pt = fabric.angles_from_axes(specimen.perturb_axes(inp["axes"], 5.0,
                                                            np.random.default_rng(1)))
    Ct, rho = fabric.stiffness_table(theta, phi)
    inp = dict(inp, Ct=Ct, rho_t=rho)
    sh = _device_shot(inp, "f64", "tiled")

    def traces(table):
        sh.Ct.copy_(torch.as_tensor(table, device=sh.device))
        for f in sh.v + sh.s:
            f.zero_()
        tr = sh.new_traces(nt)
        sh.run(sh.wavelet(wav), 0, nt, tr)
        return tr

    d = traces(fabric.stiffness_table(tt, pt)[0]).clone()
    sh.Ct.copy_(torch.as_tensor(Ct, device=sh.device))
    J, g, _ = adjoint.misfit_gradient(
        sh, wav, nt, lambda tr: (0.5 * float(((tr - d) ** 2).sum()), tr - d),
        checkpoint_every=40, n_proj=2, proj=fabric.angle_projections(theta, phi))
    g = g.cpu().numpy()[1:]
    h = 1e-4
    fd = np.zeros_like(g)
    for gi in range(len(theta)):
        for which in (0, 1):
            ap = [theta.copy(), phi.copy()]
            am = [theta.copy(), phi.copy()]
            ap[which][gi] += h
            am[which][gi] -= h
            Jp = 0.5 * float(((traces(fabric.stiffness_table(*ap)[0]) - d) ** 2).sum())
            Jm = 0.5 * float(((traces(fabric.stiffness_table(*am)[0]) - d) ** 2).sum())
            fd[gi, which] = (Jp - Jm) / (2 * h)
    scale = np.abs(g).max()
    assert scale > 0
    rel = np.abs(fd - g) / np.maximum(np.abs(g), 1e-3 * scale)
    assert rel.max() <= 1e-4, (rel.max(), fd, g)


@pytest.mark.parametrize("precision,kernel", [("f32", "tiled"), ("f64", "tiled"), ("f32", "direct")])
@pytest.mark.parametrize("n_proj", [0, 2])
def test_strain_history_equals_velocity_history(torch_gpu, precision, kernel, n_proj):
    """Storing the strains the forward computes (6 fields) instead of v (3)
    and skipping their recompute gives bitwise the same gradient."""
    _case("adjoint strain history %s %s %d" % (precision, kernel, n_proj))
    from pulse_echo_azimuthal_simulation_amd import adjoint, fabric
    inp = H.random_medium((26, 30, 70), n_grain=12, seed=21, sponge=4)
    theta, phi = fabric.angles_from_axes(inp["axes"])
    P = fabric.angle_projections(theta, phi) if n_proj else None
    nt = 36
    rng = np.random.default_rng(2)
    out = {}
    res_np = None
    for hist in ("velocity", "strain"):
        sh = _device_shot(inp, precision, kernel)
        if res_np is None:
            res_np = rng.normal(size=(nt, sh.n_rec))
        res = torch_gpu.as_tensor(res_np, device=sh.device).to(sh.real)
        adj = adjoint.DeviceAdjoint(sh, n_proj=n_proj, proj=P, hist=hist)
        _, g, _, gw = adjoint.misfit_gradient(sh, np.resize(inp["wavelet"], nt), nt,
                                              lambda tr: (0.0, res), checkpoint_every=10,
                                              with_wavelet=True, adjoint=adj)
        torch_gpu.cuda.synchronize()
        out[hist] = (g.cpu().numpy(), gw.cpu().numpy())
    assert np.abs(out["velocity"][0]).max() > 0
    assert np.array_equal(out["velocity"][0], out["strain"][0])
    assert np.array_equal(out["velocity"][1], out["strain"][1])
