"""torch.autograd over the C-ABI: the gradient of a variational cost
J = sum((y - H(x))^2 / sigma^2) matches central finite differences of the
GPU forward model and the channel sum of the oracle's K-matrix."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_sensor

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(coeff_dir):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from crtmv3_amd import api, synthetic
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    ps = synthetic.make_profiles(3, [100, 80, 120], seed=31)
    dev = torch.device("cuda:0")
    dps = api.device_profiles(ps, dev)
    yield ctx, ps, dps, dev
    ctx.destroy()


def _cost(ctx, dps, T, A, Ts, E, y, prec):
    from crtmv3_amd.autograd import crtm_bt
    bt = crtm_bt(ctx, dps, T, A, Ts, E, precision=prec)
    return (((y - bt) / 0.2) ** 2).sum()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "split"])
def test_gradient_of_cost_matches_finite_differences(setup, prec):
    import torch
    ctx, ps, dps, dev = setup
    L = ctx.n_channels
    print(f"case autograd prec={prec}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(0)
    y = (torch.from_numpy(np.full((3, L), 250.0)) + torch.randn(3, L, generator=g, dtype=torch.float64)).to(dev)
    T = dps.temperature.clone().requires_grad_(True)
    A = dps.absorber.clone().requires_grad_(True)
    Ts = dps.water_temperature.clone().requires_grad_(True)
    E = torch.full((3, L), 0.97, dtype=torch.float64, device=dev, requires_grad=True)
    J = _cost(ctx, dps, T, A, Ts, E, y, prec)
    J.backward()
    torch.cuda.synchronize()
    rng = np.random.default_rng(4)
    for idx in rng.choice(int(ps.layer_offset[-1]), 5, replace=False):
        for which, h in (("T", 1e-3), ("q", None)):
            with torch.no_grad():
                if which == "T":
                    Tp, Tm = T.detach().clone(), T.detach().clone()
                    Tp[idx] += h
                    Tm[idx] -= h
                    fd = (_cost(ctx, dps, Tp, A.detach(), Ts.detach(), E.detach(), y, 0) -
                          _cost(ctx, dps, Tm, A.detach(), Ts.detach(), E.detach(), y, 0)) / (2 * h)
                    got = T.grad[idx]
                else:
                    hq = 1e-3 * float(A[0, idx])
                    Ap, Am = A.detach().clone(), A.detach().clone()
                    Ap[0, idx] += hq
                    Am[0, idx] -= hq
                    fd = (_cost(ctx, dps, T.detach(), Ap, Ts.detach(), E.detach(), y, 0) -
                          _cost(ctx, dps, T.detach(), Am, Ts.detach(), E.detach(), y, 0)) / (2 * hq)
                    got = A.grad[0, idx]
            # central differences of a quadratic cost: truncation error O(h^2);
            # the absorber predictors (square roots, powers) make it larger for q
            tol = (1e-5 if which == "T" else 1e-3) if prec == 0 else 2e-3
            assert abs(float(got) - float(fd)) <= tol * max(abs(float(fd)), 1e-3), (which, idx, float(got), float(fd))
    # surface temperature gradient
    with torch.no_grad():
        Tsp, Tsm = Ts.detach().clone(), Ts.detach().clone()
        Tsp[1] += 1e-3
        Tsm[1] -= 1e-3
        fd = (_cost(ctx, dps, T.detach(), A.detach(), Tsp, E.detach(), y, 0) -
              _cost(ctx, dps, T.detach(), A.detach(), Tsm, E.detach(), y, 0)) / 2e-3
    assert abs(float(Ts.grad[1]) - float(fd)) <= (1e-5 if prec == 0 else 2e-3) * abs(float(fd))


def test_gradient_equals_oracle_adjoint(setup, coeff_dir, oracle_mod):
    import torch
    from crtmv3_amd.autograd import crtm_bt
    ctx, ps, dps, dev = setup
    L = ctx.n_channels
    print("case autograd vs oracle", flush=True)
    w = torch.linspace(-1.0, 1.0, L, dtype=torch.float64, device=dev).repeat(3, 1)
    T = dps.temperature.clone().requires_grad_(True)
    A = dps.absorber.clone().requires_grad_(True)
    Ts = dps.water_temperature.clone().requires_grad_(True)
    E = torch.full((3, L), 0.97, dtype=torch.float64, device=dev, requires_grad=True)
    (crtm_bt(ctx, dps, T, A, Ts, E, precision=0) * w).sum().backward()
    tau, spc = load_sensor(coeff_dir, "hirs_syn")
    s = oracle_mod.Sensor(tau, spc)
    wn = w.cpu().numpy()
    for m in range(3):
        ref = oracle_mod.run_profile(s, ps, m, np.full(L, 0.97), "k", bt_k=wn[m])
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        np.testing.assert_allclose(T.grad[a:b].cpu().numpy(), ref["t_k"].sum(0), rtol=1e-9,
                                   atol=1e-12 * np.abs(ref["t_k"]).max())
        np.testing.assert_allclose(A.grad[:, a:b].cpu().numpy(), ref["abs_k"].sum(0), rtol=1e-9,
                                   atol=1e-12 * np.abs(ref["abs_k"]).max())
        np.testing.assert_allclose(E.grad[m].cpu().numpy(), ref["emis_k"], rtol=1e-9, atol=1e-14)
