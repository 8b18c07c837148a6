"""GPU fabric inversion: the gradient of the sweep misfit in the Watson
parameters (adjoint tables gradient chained through the reparameterisation)
against central differences of the misfit itself, in float64."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def test_fabric_gradient_matches_central_differences(torch_gpu):
    print("\n[gpu case] fabric gradient vs FD", flush=True)
    from pulse_echo_azimuthal_simulation_amd import fabric_fit as F, host, sweep
    h = 1e-4
    build = sweep.disk_build(3.2e-3, 1.6e-3, h, 24, seed=3)
    lat = F.latents(24, 5)
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True)
    inv = F.FabricInversion(build, [0.0, 70.0, 140.0], h, co, 4e6, nt=260, lat=lat, sponge=8,
                            checkpoint_every=50, precision="f64", element_d_mm=1.0)
    data = [t.clone() for t in inv.traces((0.3, 4.0))]

    def J_of(p):
        return sum(0.5 * float(((t - d) ** 2).sum()) for t, d in zip(inv.traces(p), data))

    p = (0.55, 2.5)
    J, g = inv.misfit_gradient(p, data)
    assert J > 0 and abs(J / J_of(p) - 1) < 1e-12
    ha, hk = 1e-4, 1e-4
    fa = (J_of((p[0] + ha, p[1])) - J_of((p[0] - ha, p[1]))) / (2 * ha)
    fk = (J_of((p[0], p[1] + hk)) - J_of((p[0], p[1] - hk))) / (2 * hk)
    print("dJ/dalpha adjoint %.9e FD %.9e; dJ/dkappa adjoint %.9e FD %.9e"
          % (g[0], fa, g[1], fk))
    assert abs(g[0] - fa) <= 1e-5 * abs(fa)
    assert abs(g[1] - fk) <= 1e-5 * abs(fk)
    # at the truth the misfit and its gradient vanish
    J0, g0 = inv.misfit_gradient((0.3, 4.0), data)
    assert J0 <= 1e-20 * J and np.all(np.abs(g0) <= 1e-8 * np.abs(g))
