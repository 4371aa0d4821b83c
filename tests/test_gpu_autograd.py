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


def test_gradient_with_computed_surface_optics(coeff_dir, oracle_mod):
    """emissivity=None: the FASTEM6 optics are computed and the cost's
    gradient reaches wind speed, wind direction and salinity; it equals the
    channel sum of the oracle's Surface_K and central differences of the
    GPU forward model."""
    import torch
    from crtmv3_amd import api, coeffio, synthetic
    from crtmv3_amd.autograd import crtm_bt
    print("autograd with computed surface optics", flush=True)
    synthetic.write_mwwater_coeff(coeff_dir)
    synthetic.write_sensor(coeff_dir, "atms_syn")
    ctx = api.crtm_init(["atms_syn"], coeff_dir)
    try:
        dev = torch.device("cuda:0")
        ps = synthetic.ocean_surface(synthetic.make_profiles(3, [100, 90, 100], seed=33), seed=34)
        dps = api.device_profiles(ps, dev)
        L = ctx.n_channels
        y = torch.full((3, L), 230.0, dtype=torch.float64, device=dev)
        W = dps.wind_speed.clone().requires_grad_(True)
        D = dps.wind_direction.clone().requires_grad_(True)
        S = dps.salinity.clone().requires_grad_(True)

        def cost(w, d, s_, prec=api.PREC_FP64):
            bt = crtm_bt(ctx, dps, dps.temperature, dps.absorber, dps.water_temperature, None, precision=prec,
                         wind_speed=w, wind_direction=d, salinity=s_)
            return (((y - bt) / 0.5) ** 2).sum(), bt

        J, bt = cost(W, D, S)
        J.backward()
        torch.cuda.synchronize()
        tau, spc = load_sensor(coeff_dir, "atms_syn")
        Sn = oracle_mod.Sensor(tau, spc)
        mw = oracle_mod.MWwater(coeffio.read_mwwatercoeff(f"{coeff_dir}/{synthetic.MWWATERCOEFF_FILE}"))
        seed = (-2.0 * (y - bt) / 0.25).detach().cpu().numpy()
        for m in range(3):
            so = oracle_mod.SurfaceOptics(spc, mw=mw, salinity=ps.salinity[m], wind_speed=ps.wind_speed[m],
                                          wind_direction=ps.wind_direction[m], sensor_azimuth=ps.sensor_azimuth[m])
            r = oracle_mod.run_profile_sfc(Sn, ps, m, so, "k", bt_k=seed[m])
            want = r["sfc_k"].sum(0)   # salinity, wind speed, wind direction
            got = np.array([S.grad[m].item(), W.grad[m].item(), D.grad[m].item()])
            np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())
        for which, h in (("wind", 1e-4), ("dir", 1e-3)):
            with torch.no_grad():
                wp, wm, dp, dm = W.detach().clone(), W.detach().clone(), D.detach().clone(), D.detach().clone()
                if which == "wind":
                    wp[1] += h
                    wm[1] -= h
                else:
                    dp[1] += h
                    dm[1] -= h
                fd = (cost(wp, dp, S.detach())[0] - cost(wm, dm, S.detach())[0]) / (2 * h)
            got = (W.grad if which == "wind" else D.grad)[1].item()
            assert abs(got - fd.item()) <= 1e-5 * max(abs(fd.item()), 1e-6), (which, got, fd.item())
    finally:
        ctx.destroy()
