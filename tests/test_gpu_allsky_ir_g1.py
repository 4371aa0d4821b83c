"""All-sky infrared on ODPS group-1 sensors with OPTRAN, on the GPU against
the oracle (through the C-ABI).

* The IASI-like 8461-channel sensor in the shape of the fourth baseline
  configuration, scaled down to five 100-layer profiles: 60/40 clear/cloudy
  with liquid and ice clouds, forward and K-matrix at 8 streams (per-lane
  kernels) and 16 streams (wavefront kernels on the matrix cores).  The GPU
  runs every channel; the oracle checks a sample of 160 channels, 40 of them
  OPTRAN channels, so the hand-off of the OPTRAN gas optical depths and of
  the adjoint seeds between the group-1 clear-sky kernel and the scattering
  path is covered.
* The small CrIS-like group-1 sensor with clouds and aerosols at 8 streams.

Tolerances (BT in K; Jacobians relative to the block maximum, and
element-wise as conftest.assert_jacobian_close states: rtol 1e-7 FP64, 1e-3
FP32-split, plus a small absolute allowance):
  FP64 path        BT 1e-6 K,  Jacobians 1e-8
  FP32-split path  BT 1e-2 K,  Jacobians 1e-3
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor, subset_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: dict(bt=1e-6, jac=1e-8), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


@pytest.fixture(scope="module")
def iasi(coeff_dir, oracle_mod):
    tau, spc = load_sensor(coeff_dir, "iasi_syn")
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    opt = np.flatnonzero(tau.osignificance == 1)
    rest = np.setdiff1d(np.arange(tau.n_channels), opt)
    idx = np.sort(np.concatenate([opt[:: max(1, opt.size // 40)][:40],
                                  rest[np.linspace(0, rest.size - 1, 120).astype(int)]]))
    ts, ss = subset_sensor(tau, spc, idx)
    ctx = api.crtm_init(["iasi_syn"], coeff_dir)
    yield ctx, oracle_mod.Sensor(ts, ss), oracle_mod.CloudCoeff(cc), idx
    ctx.destroy()


def _profiles():
    ps = synthetic.make_profiles(5, 100, seed=4101)
    ps = synthetic.add_clouds(ps, seed=4102, types=(synthetic.WATER_CLOUD, synthetic.ICE_CLOUD),
                              cover="random", median_gm3=0.01)
    for m in (0, 2, 4):                       # profiles 1 and 3 keep their clouds
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        ps.cloud_water_content[:, a:b] = 0.0
        ps.cloud_fraction[a:b] = 0.0
    cloudy = [bool(ps.cloud_water_content[:, int(ps.layer_offset[m]):int(ps.layer_offset[m + 1])].max() > 1e-6)
              for m in range(5)]
    assert sum(cloudy) == 2, cloudy          # 60/40 clear / cloudy
    return ps


def _check_k(ps, rts, kd, refs, idx, prec, label):
    tol = TOL[prec]
    nc = ps.cloud_type.shape[1]
    worst, welem = {}, {}
    for m, (f, k) in enumerate(refs):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        dbt = np.abs(np.asarray(rts["brightness_temperature"])[m, idx] - f["bt"]).max()
        assert dbt <= tol["bt"], f"{label} profile {m}: |dBT| {dbt:.3e}"
        blocks = [
            ("T", np.asarray(kd["temperature_k"][a:b])[:, idx].T, k["t_k"]),
            ("absorber", np.asarray(kd["absorber_k"][:, a:b, :])[:, :, idx].transpose(2, 0, 1), k["abs_k"]),
            ("water content", np.asarray(kd["cloud_water_content_k"][:, a:b, :])[:, :, idx].transpose(2, 0, 1),
             k["wc_k"][:, :nc]),
            ("radius", np.asarray(kd["cloud_effective_radius_k"][:, a:b, :])[:, :, idx].transpose(2, 0, 1),
             k["reff_k"][:, :nc]),
            ("fraction", np.asarray(kd["cloud_fraction_k"][a:b])[:, idx].T, k["cf_k"]),
            ("Ts", np.asarray(kd["water_temperature_k"])[m, idx], k["sfc_t_k"][:, 1]),
            ("emissivity", np.asarray(kd["surface_emissivity_k"])[m, idx], k["emis_k"]),
        ]
        for what, got, want in blocks:
            err, elem = assert_jacobian_close(got, want, tol["jac"], f"{label} profile {m} {what}_K", prec)
            worst[what] = max(worst.get(what, 0.0), err)
            welem[what] = max(welem.get(what, 0.0), elem)
    print(f"  {label} worst Jacobian errors (block-relative / element-wise): "
          + ", ".join(f"{k} {v:.1e}/{welem[k]:.1e}" for k, v in worst.items()), flush=True)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("n_streams", [8, 16])
def test_iasi_allsky_k_matches_oracle(iasi, oracle_mod, n_streams, precision):
    ctx, S, C, idx = iasi
    print(f"allsky K iasi_syn 8461 ch M=5 (2 cloudy) n_streams={n_streams} precision={precision}", flush=True)
    ps = _profiles()
    L = ctx.n_channels
    rng = np.random.default_rng(41)
    eps = rng.uniform(0.95, 0.99, (5, L))
    btk = rng.uniform(0.5, 1.5, (5, L))
    rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision),
                           n_streams=n_streams)
    refs = [oracle_mod.k_matrix_allsky(S, C, ps, m, eps[m, idx], n_streams, 0, btk[m, idx], None,
                                       method="adjoint") for m in range(5)]
    _check_k(ps, rts, kd, refs, idx, precision, f"iasi n={n_streams}")
    fwd = ctx.forward(ps, eps, precision=precision, n_streams=n_streams)
    np.testing.assert_allclose(fwd["brightness_temperature"], rts["brightness_temperature"], rtol=0,
                               atol=1e-9 if precision == api.PREC_FP64 else 1e-3)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_cris_small_clouds_and_aerosols(coeff_dir, oracle_mod, precision):
    print(f"allsky K cris_small_syn clouds + aerosols n_streams=8 precision={precision}", flush=True)
    tau, spc = load_sensor(coeff_dir, "cris_small_syn")
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir))
    S, C = oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac)
    ctx = api.crtm_init(["cris_small_syn"], coeff_dir)
    try:
        ps = synthetic.make_profiles(4, [100, 87, 120, 100], seed=4201)
        ps = synthetic.add_clouds(ps, seed=4202, types=(synthetic.WATER_CLOUD, synthetic.ICE_CLOUD),
                                  cover="random", median_gm3=0.01, cloudy_fraction=0.75)
        ps = synthetic.add_aerosols(ps, seed=4203, aerosol_fraction=0.75)
        L = ctx.n_channels
        rng = np.random.default_rng(42)
        eps = rng.uniform(0.95, 0.99, (4, L))
        btk = rng.uniform(0.5, 1.5, (4, L))
        rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=8)
        tol = TOL[precision]
        nc, na = ps.cloud_type.shape[1], ps.aerosol_type.shape[1]
        for m in range(4):
            f, k = oracle_mod.k_matrix_allsky(S, C, ps, m, eps[m], 8, 0, btk[m], None, method="adjoint")
            a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
            assert np.abs(rts["brightness_temperature"][m] - f["bt"]).max() <= tol["bt"]
            for what, got, want in (
                    ("T", kd["temperature_k"][a:b].T, k["t_k"]),
                    ("absorber", np.asarray(kd["absorber_k"][:, a:b, :]).transpose(2, 0, 1), k["abs_k"]),
                    ("water content", np.asarray(kd["cloud_water_content_k"][:, a:b, :]).transpose(2, 0, 1),
                     k["wc_k"][:, :nc]),
                    ("aerosol conc", np.asarray(kd["aerosol_concentration_k"][:, a:b, :]).transpose(2, 0, 1),
                     k["aconc_k"][:, :na]),
                    ("aerosol radius", np.asarray(kd["aerosol_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1),
                     k["areff_k"][:, :na]),
                    ("Ts", kd["water_temperature_k"][m], k["sfc_t_k"][:, 1])):
                assert_jacobian_close(got, want, tol["jac"], f"cris_small profile {m} {what}_K", precision)
    finally:
        ctx.destroy()
