"""CRTM_Adjoint through crtm_adjoint (channel sums formed in the kernels, no
per-channel Jacobian) against the oracle: the oracle's K-matrix seeded with
RTSolution_AD, summed over channels (src/CRTM_Adjoint_Module.f90:1378-1396 is
that sum taken before the predictor adjoint).

Clear-sky group-1 OPTRAN (CrIS-like subset) and group-2 (HIRS-like) sensors,
all-sky microwave (ATMS-like, partial cover, every cloud type) and infrared
with clouds and aerosols, FP64 and FP32-split, host and device memory;
Skip_Profile; the accumulate contract of Atmosphere_AD / Surface_AD.
Tolerances: per variable and profile, the largest error relative to the
largest element, FP64 1e-9 (all-sky 1e-8), split 1e-3.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: 1e-9, api.PREC_FP32_SPLIT: 1e-3}


def _rel(got, want):
    scale = np.abs(want).max()
    return np.abs(np.asarray(got) - want).max() / scale if scale > 0 else np.abs(got).max()


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["cris_small_syn", "hirs_syn"])
def test_clear_adjoint_matches_oracle(coeff_dir, oracle_mod, name, precision):
    print(f"adjoint clear {name} precision={precision}", flush=True)
    tau, spc = load_sensor(coeff_dir, name)
    S = oracle_mod.Sensor(tau, spc)
    ctx = api.crtm_init([name], coeff_dir)
    try:
        ps = synthetic.make_profiles(5, [100, 77, 120, 91, 100], seed=501)
        L = ctx.n_channels
        rng = np.random.default_rng(51)
        eps = rng.uniform(0.9, 0.99, (5, L))
        bt_ad = rng.normal(size=(5, L))
        bt, ad = ctx.adjoint(ps, eps, bt_ad, precision=precision)
        worst = {}
        for m in range(5):
            r = oracle_mod.run_profile(S, ps, m, eps[m], "k", bt_k=bt_ad[m])
            a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
            assert np.abs(bt[m] - r["bt"]).max() < (1e-8 if precision == api.PREC_FP64 else 1e-2)
            for key, got, want in (("T", ad["temperature_ad"][a:b], r["t_k"].sum(0)),
                                   ("absorber", ad["absorber_ad"][:, a:b], r["abs_k"].sum(0)),
                                   ("Ts", ad["water_temperature_ad"][m], r["sfc_t_k"][:, 1].sum()),
                                   ("emissivity", ad["surface_emissivity_ad"][m], r["emis_k"])):
                err = _rel(got, np.asarray(want))
                worst[key] = max(worst.get(key, 0.0), err)
                assert err <= TOL[precision], (m, key, err)
        print("  worst: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)
    finally:
        ctx.destroy()


@pytest.fixture(scope="module")
def cloudy(coeff_dir, oracle_mod):
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir))
    out = {}
    for name in ("atms_syn", "hirs_syn"):
        synthetic.write_sensor(coeff_dir, name)
        tau, spc = load_sensor(coeff_dir, name)
        out[name] = (api.crtm_init([name], coeff_dir), oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac))
    yield out
    for c, _, _ in out.values():
        c.destroy()


@pytest.mark.parametrize("memory", ["host", "device"])
@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["atms_syn", "hirs_syn"])
def test_allsky_adjoint_matches_oracle(cloudy, oracle_mod, name, precision, memory):
    import torch
    ctx, S, C = cloudy[name]
    print(f"adjoint all-sky {name} precision={precision} memory={memory}", flush=True)
    M, L = 5, S.n_channels
    ks = [100, 80, 100, 91, 77]
    if name == "atms_syn":
        ps = synthetic.add_clouds(synthetic.make_profiles(M, ks, seed=511), seed=512, cover="random",
                                  reff_scale=8.0, cloudy_fraction=0.8)
    else:
        ps = synthetic.add_clouds(synthetic.make_profiles(M, ks, seed=511), seed=512, median_gm3=0.005,
                                  cover="random", cloudy_fraction=0.8)
        ps = synthetic.add_aerosols(ps, seed=513)
    rng = np.random.default_rng(52)
    eps = rng.uniform(0.5, 0.95, (M, L))
    bt_ad = rng.normal(size=(M, L))
    ns = 4 if name == "hirs_syn" else 8
    if memory == "device":
        dps = api.device_profiles(ps)
        bt, ad = ctx.adjoint(dps, torch.as_tensor(eps, device="cuda"), torch.as_tensor(bt_ad, device="cuda"),
                             precision=precision, n_streams=ns)
        torch.cuda.synchronize()
        bt, ad = bt.cpu().numpy(), {k: v.cpu().numpy() for k, v in ad.items()}
    else:
        bt, ad = ctx.adjoint(ps, eps, bt_ad, precision=precision, n_streams=ns)
    tol = 1e-8 if precision == api.PREC_FP64 else 1e-3
    nc = ps.cloud_type.shape[1]
    worst = {}
    for m in range(M):
        f, k = oracle_mod.k_matrix_allsky(S, C, ps, m, eps[m], ns, 0, bt_ad[m], None, method="adjoint")
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        assert np.abs(bt[m] - f["bt"]).max() < (1e-6 if precision == api.PREC_FP64 else 1e-2)
        blocks = [("T", ad["temperature_ad"][a:b], k["t_k"].sum(0)),
                  ("absorber", ad["absorber_ad"][:, a:b], k["abs_k"].sum(0)),
                  ("water content", ad["cloud_water_content_ad"][:, a:b], k["wc_k"].sum(0)[:nc]),
                  ("radius", ad["cloud_effective_radius_ad"][:, a:b], k["reff_k"].sum(0)[:nc]),
                  ("fraction", ad["cloud_fraction_ad"][a:b], k["cf_k"].sum(0)),
                  ("Ts", ad["water_temperature_ad"][m], k["sfc_t_k"][:, 1].sum()),
                  ("emissivity", ad["surface_emissivity_ad"][m], k["emis_k"])]
        if "aerosol_concentration_ad" in ad:
            na = ps.aerosol_type.shape[1]
            blocks += [("aerosol", ad["aerosol_concentration_ad"][:, a:b], k["aconc_k"].sum(0)[:na]),
                       ("aerosol radius", ad["aerosol_effective_radius_ad"][:, a:b], k["areff_k"].sum(0)[:na])]
        for key, got, want in blocks:
            want = np.asarray(want)
            if np.abs(want).max() == 0:
                assert np.abs(got).max() == 0, (m, key)
                continue
            err = _rel(got, want)
            worst[key] = max(worst.get(key, 0.0), err)
            assert err <= tol, (m, key, err)
    print("  worst: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)
    assert worst.get("water content", 0) >= 0 and "fraction" in worst


def test_adjoint_accumulates_and_skips(cloudy):
    """Atmosphere_AD / Surface_AD are added onto (the reference's
    INTENT(IN OUT) contract), and a skipped profile adds nothing."""
    ctx, S, _ = cloudy["atms_syn"]
    print("adjoint accumulate + Skip_Profile", flush=True)
    M, L = 4, S.n_channels
    ps = synthetic.add_clouds(synthetic.make_profiles(M, [100, 80, 100, 90], seed=521), seed=522, cover="random",
                              reff_scale=8.0)
    rng = np.random.default_rng(53)
    eps = rng.uniform(0.5, 0.9, (M, L))
    bt_ad = rng.normal(size=(M, L))
    _, ad = ctx.adjoint(ps, eps, bt_ad, n_streams=4)
    NT = int(ps.layer_offset[-1])
    pre = {k: rng.normal(size=v.shape) for k, v in ad.items()}
    acc = {k: v.copy() for k, v in pre.items()}
    bt = np.zeros((M, L))
    ctx._adjoint_call(ps, eps, bt_ad, api.PREC_FP64, 4, 0, True, None, bt, acc, 0)
    for k in ad:
        if k == "surface_emissivity_ad":      # per channel, overwritten as RTSolution_K's
            np.testing.assert_array_equal(acc[k], ad[k])
        else:
            np.testing.assert_allclose(acc[k], pre[k] + ad[k], rtol=1e-12, atol=1e-12 * np.abs(ad[k]).max(),
                                       err_msg=k)
    skip = np.array([0, 1, 0, 0], np.int32)
    _, ad_s = ctx.adjoint(ps, eps, bt_ad, n_streams=4, skip_profile=skip)
    a, b = int(ps.layer_offset[1]), int(ps.layer_offset[2])
    for k in ("temperature_ad", "absorber_ad", "cloud_water_content_ad", "cloud_fraction_ad"):
        assert np.all(ad_s[k][..., a:b] == 0), k
        keep = np.r_[0:a, b:NT]
        np.testing.assert_allclose(ad_s[k][..., keep], ad[k][..., keep], rtol=1e-12,
                                   atol=1e-12 * np.abs(ad[k]).max(), err_msg=k)
    assert ad_s["water_temperature_ad"][1] == 0
