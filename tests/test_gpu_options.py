"""Options%Include_Scattering = .FALSE. and Options(m)%Skip_Profile on the
GPU (through the C-ABI).

Include_Scattering = .FALSE. (the reference's test_ScatteringSwitch
regression): clouds and aerosols absorb only, ke (1 - w), and the radiance is
the emission solution (src/AtmScatter/CRTM_CloudScatter.f90:986-997,
1372-1407; src/RTSolution/Common_RTSolution.f90:263).  Forward and K-matrix
against the oracle's hand-written adjoint, which equals its complex-step
Jacobians for this switch (tests/test_oracle_allsky.py), for microwave
columns with all four cloud types and infrared columns with clouds and
aerosols, FP64 and FP32-split.  Tolerances as the other all-sky tests:
FP64 BT 1e-6 K / Jacobians 1e-8 relative, FP32-split 1e-2 K / 1e-3.

Skip_Profile (src/CRTM_Forward_Module.f90:498, 536; src/CRTM_K_Matrix_Module.f90:
594, 631): a skipped profile's RTSolution is zero, its Jacobians are zero,
its inputs are not checked and it does not fail the call; every other
profile's results are bit-identical to a call without the skipped ones.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: dict(bt=1e-6, jac=1e-8), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


@pytest.fixture(scope="module")
def sensors(coeff_dir, oracle_mod):
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir))
    out = {}
    for name in ("atms_syn", "hirs_syn"):
        synthetic.write_sensor(coeff_dir, name)
        tau, spc = load_sensor(coeff_dir, name)
        out[name] = (api.crtm_init([name], coeff_dir), oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac))
    yield out
    for ctx, _, _ in out.values():
        ctx.destroy()


def _cloudy(name, seed, M):
    ks = [100, 80, 100, 91, 100, 77][:M]
    if name == "atms_syn":
        return synthetic.add_clouds(synthetic.make_profiles(M, ks, seed=seed), seed=seed + 1, cover="random",
                                    reff_scale=8.0)
    ps = synthetic.add_clouds(synthetic.make_profiles(M, ks, seed=seed), seed=seed + 1, median_gm3=0.005,
                              cover="random")
    return synthetic.add_aerosols(ps, seed=seed + 2)


def _blocks(ps, kd, k, m):
    a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
    nc = ps.cloud_type.shape[1]
    out = [("T", kd["temperature_k"][a:b].T, k["t_k"]),
           ("absorber", np.asarray(kd["absorber_k"][:, a:b, :]).transpose(2, 0, 1), k["abs_k"]),
           ("water content", np.asarray(kd["cloud_water_content_k"][:, a:b, :]).transpose(2, 0, 1),
            k["wc_k"][:, :nc]),
           ("radius", np.asarray(kd["cloud_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1), k["reff_k"][:, :nc]),
           ("fraction", kd["cloud_fraction_k"][a:b].T, k["cf_k"]),
           ("Ts", kd["water_temperature_k"][m], k["sfc_t_k"][:, 1]),
           ("emissivity", kd["surface_emissivity_k"][m], k["emis_k"])]
    if "aerosol_concentration_k" in kd:
        na = ps.aerosol_type.shape[1]
        out += [("aerosol concentration", np.asarray(kd["aerosol_concentration_k"][:, a:b, :]).transpose(2, 0, 1),
                 k["aconc_k"][:, :na]),
                ("aerosol radius", np.asarray(kd["aerosol_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1),
                 k["areff_k"][:, :na])]
    return out


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["atms_syn", "hirs_syn"])
def test_no_scattering_k_matches_oracle(sensors, oracle_mod, name, precision):
    ctx, S, C = sensors[name]
    print(f"Include_Scattering=.FALSE. K {name} precision={precision}", flush=True)
    ps = _cloudy(name, 401, 5)
    L = S.n_channels
    rng = np.random.default_rng(41)
    eps = rng.uniform(0.5, 0.97, (ps.n_profiles, L))
    btk = rng.uniform(0.5, 1.5, (ps.n_profiles, L))
    rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=4,
                           include_scattering=False)
    fwd = ctx.forward(ps, eps, precision=precision, n_streams=4, include_scattering=False)
    scat = ctx.forward(ps, eps, precision=precision, n_streams=4)
    tol = TOL[precision]
    worst = {}
    for m in range(ps.n_profiles):
        f, k = oracle_mod.k_matrix_allsky(S, C, ps, m, eps[m], 4, 0, btk[m], None, method="adjoint",
                                          include_scattering=False)
        dbt = np.abs(rts["brightness_temperature"][m] - f["bt"]).max()
        assert dbt <= tol["bt"], f"profile {m}: |dBT| {dbt:.3e}"
        assert np.abs(fwd["brightness_temperature"][m] - f["bt"]).max() <= tol["bt"]
        for what, got, want in _blocks(ps, kd, k, m):
            err, _ = assert_jacobian_close(got, want, tol["jac"], f"{name} profile {m} {what}_K", precision)
            worst[what] = max(worst.get(what, 0.0), err)
    print("  worst block-relative errors: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)
    # the switch acts: scattering changes the brightness temperatures
    assert np.abs(scat["brightness_temperature"] - fwd["brightness_temperature"]).max() > 1e-3


def _check_skip(full, part, skip, keys_m, keys_nt, ps):
    """Profiles not skipped equal the full call bit for bit; skipped ones are
    zero (per-profile [M, L] keys and layer-row [.., NT, L] keys)."""
    for key in keys_m:
        a, b = np.asarray(full[key]), np.asarray(part[key])
        for m, s in enumerate(skip):
            if s:
                assert np.all(b[m] == 0), (key, m)
            else:
                np.testing.assert_array_equal(b[m], a[m], err_msg=f"{key} profile {m}")
    for key in keys_nt:
        a, b = np.asarray(full[key]), np.asarray(part[key])
        for m, s in enumerate(skip):
            r0, r1 = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
            if s:
                assert np.all(b[..., r0:r1, :] == 0), (key, m)
            else:
                np.testing.assert_array_equal(b[..., r0:r1, :], a[..., r0:r1, :], err_msg=f"{key} profile {m}")


@pytest.mark.parametrize("memory", ["host", "device"])
@pytest.mark.parametrize("cloudy", [False, True], ids=["clear", "cloudy"])
def test_skip_profile(sensors, memory, cloudy):
    import torch
    ctx, S, _ = sensors["atms_syn"]
    print(f"Skip_Profile memory={memory} cloudy={cloudy}", flush=True)
    M, L = 6, S.n_channels
    ps = synthetic.make_profiles(M, [100, 80, 100, 91, 100, 77], seed=411)
    if cloudy:
        ps = synthetic.add_clouds(ps, seed=412, cover="random", reff_scale=8.0)
    eps = np.random.default_rng(42).uniform(0.5, 0.9, (M, L))
    skip = np.array([0, 1, 0, 0, 1, 0], np.int32)
    # a skipped profile is not checked: give it an invalid temperature
    bad = dataclasses.replace(ps, temperature=ps.temperature.copy())
    r0, r1 = int(ps.layer_offset[1]), int(ps.layer_offset[2])
    bad.temperature[r0:r1] = -5.0
    if memory == "device":
        run_ps, run_eps = api.device_profiles(bad), torch.as_tensor(eps, device="cuda")
        sk = torch.as_tensor(skip, device="cuda")
    else:
        run_ps, run_eps, sk = bad, eps, skip
    full_f = ctx.forward(ps, eps, extras=True, n_streams=4)
    full_k = ctx.k_matrix(ps, eps, n_streams=4, extras=cloudy)
    part_f = ctx.forward(run_ps, run_eps, extras=True, n_streams=4, skip_profile=sk)
    part_k = ctx.k_matrix(run_ps, run_eps, n_streams=4, extras=cloudy, skip_profile=sk)
    if memory == "device":
        torch.cuda.synchronize()
        to_np = lambda d: {k: v.cpu().numpy() for k, v in d.items()}      # noqa: E731
        part_f, part_k = to_np(part_f), (to_np(part_k[0]), to_np(part_k[1]))
    assert np.all(np.asarray(part_f["status"]) == 0) and np.all(np.asarray(part_k[0]["status"]) == 0)
    fkeys = ["radiance", "brightness_temperature", "up_radiance", "down_radiance", "surface_planck_radiance",
             "surface_emissivity"]
    _check_skip(full_f, part_f, skip, fkeys, [], ps)
    _check_skip(full_k[0], part_k[0], skip, ["radiance", "brightness_temperature"], [], ps)
    kkeys_m = ["surface_emissivity_k", "land_temperature_k", "water_temperature_k", "snow_temperature_k",
               "ice_temperature_k"]
    kkeys_nt = ["temperature_k", "absorber_k"]
    if cloudy:
        kkeys_nt += ["cloud_water_content_k", "cloud_effective_radius_k", "cloud_fraction_k"]
        _check_skip(full_k[0], part_k[0], skip, ["r_clear", "tb_clear", "total_cloud_cover"], [], ps)
    _check_skip(full_k[1], part_k[1], skip, kkeys_m, kkeys_nt, ps)
    # without the skip the invalid profile fails the input check
    with pytest.raises(api.CRTMError):
        ctx.forward(bad, eps, n_streams=4)
