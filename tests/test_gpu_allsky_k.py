"""All-sky K-matrix on the GPU against the oracle (through the C-ABI).

Jacobians of brightness temperature with respect to temperature, absorbers,
cloud water content, effective radius and fraction, surface temperature and
emissivity, for cloudy columns of the ATMS-like microwave and HIRS-like
infrared sensors: overcast and partial cover under each overlap, user and
Mie-parameter stream counts, mixed clear and cloudy profiles, host and device
memory.  The oracle side is its hand-written adjoint, which equals its
complex-step Jacobians to rounding (tests/test_oracle_allsky.py).

Tolerances (BT in K, Jacobians relative to the largest element of the same
Jacobian block of the profile, and element-wise as conftest.assert_jacobian_close
states: rtol 1e-7 FP64, 1e-3 FP32-split, plus a small absolute allowance):
  FP64 path        BT 1e-6 K,  Jacobians 1e-8
  FP32-split path  BT 1e-2 K,  Jacobians 1e-3
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: dict(bt=1e-6, jac=1e-8), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


@pytest.fixture(scope="module")
def ctx_mw(coeff_dir):
    synthetic.write_sensor(coeff_dir, "atms_syn")
    c = api.crtm_init(["atms_syn"], coeff_dir)
    yield c
    c.destroy()


def _oracle_sensor(coeff_dir, oracle_mod, name):
    tau, spc = load_sensor(coeff_dir, name)
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    return oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc)


def _oracle_k(orc, S, C, ps, eps, btk, n_streams, overlap=0):
    return [orc.k_matrix_allsky(S, C, ps, m, eps[m], n_streams, overlap, btk[m], None, method="adjoint")
            for m in range(ps.n_profiles)]


def _rel(got, want):
    got = np.asarray(got, np.float64)
    scale = np.abs(want).max()
    if scale == 0.0:
        return np.abs(got).max()
    return np.abs(got - want).max() / scale


def _check(ps, rts, kd, refs, prec, label=""):
    tol = TOL[prec]
    nc = ps.cloud_type.shape[1]
    worst, welem = {}, {}
    for m, (f, k) in enumerate(refs):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        dbt = np.abs(rts["brightness_temperature"][m] - f["bt"]).max()
        assert dbt <= tol["bt"], f"{label} profile {m}: |dBT| {dbt:.3e}"
        blocks = [
            ("T", kd["temperature_k"][a:b].T, k["t_k"]),
            ("absorber", np.asarray(kd["absorber_k"][:, a:b, :]).transpose(2, 0, 1), k["abs_k"]),
            ("water content", np.asarray(kd["cloud_water_content_k"][:, a:b, :]).transpose(2, 0, 1),
             k["wc_k"][:, :nc]),
            ("radius", np.asarray(kd["cloud_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1), k["reff_k"][:, :nc]),
            ("fraction", kd["cloud_fraction_k"][a:b].T, k["cf_k"]),
            ("Ts", kd["water_temperature_k"][m], k["sfc_t_k"][:, 1]),
            ("emissivity", kd["surface_emissivity_k"][m], k["emis_k"]),
        ]
        for what, got, want in blocks:
            err, elem = assert_jacobian_close(got, want, tol["jac"], f"{label} profile {m} {what}_K", prec)
            worst[what] = max(worst.get(what, 0.0), err)
            welem[what] = max(welem.get(what, 0.0), elem)
    print(f"  {label} worst Jacobian errors (block-relative / element-wise): "
          + ", ".join(f"{k} {v:.1e}/{welem[k]:.1e}" for k, v in worst.items()), flush=True)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("n_streams", [4, 8])
def test_overcast_k_matches_oracle(ctx_mw, coeff_dir, oracle_mod, precision, n_streams):
    print(f"allsky K overcast atms_syn M=6 n_streams={n_streams} precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(6, 100, seed=121), seed=122)
    rng = np.random.default_rng(7)
    eps = rng.uniform(0.45, 0.75, (6, 22))
    btk = rng.uniform(0.5, 1.5, (6, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision),
                              n_streams=n_streams, extras=True)
    refs = _oracle_k(oracle_mod, S, C, ps, eps, btk, n_streams)
    _check(ps, rts, kd, refs, precision, f"overcast n={n_streams}")
    # the K-matrix call's forward half agrees with the forward call
    fwd = ctx_mw.forward(ps, eps, precision=precision, extras=True, n_streams=n_streams)
    np.testing.assert_allclose(rts["brightness_temperature"], fwd["brightness_temperature"], rtol=0,
                               atol=1e-9 if precision == api.PREC_FP64 else 1e-3)
    np.testing.assert_allclose(rts["tb_clear"], fwd["tb_clear"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("overlap", [1, 2, 3, 4, 5])
def test_partial_cover_k_overlaps(ctx_mw, coeff_dir, oracle_mod, overlap):
    print(f"allsky K partial cover overlap={overlap}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(5, 90, seed=131), seed=132, cover="random")
    eps = np.random.default_rng(8).uniform(0.45, 0.75, (5, 22))
    btk = np.ones((5, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=api.PREC_FP64, n_streams=4, overlap=overlap)
    refs = _oracle_k(oracle_mod, S, C, ps, eps, btk, 4, overlap)
    _check(ps, rts, kd, refs, api.PREC_FP64, f"overlap {overlap}")
    if overlap != 5:   # overcast: constant cover, no fraction sensitivity
        assert np.abs(kd["cloud_fraction_k"]).max() > 0


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_mie_streams_thick_layers_k(ctx_mw, coeff_dir, oracle_mod, precision):
    """Automatic stream choice and large, optically thick scattering layers
    (several Matrix_Exp doublings per layer)."""
    print(f"allsky K auto streams thick layers precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(5, 100, seed=141), seed=142, reff_scale=25.0, median_gm3=0.3)
    eps = np.random.default_rng(9).uniform(0.45, 0.75, (5, 22))
    btk = np.ones((5, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=0)
    refs = _oracle_k(oracle_mod, S, C, ps, eps, btk, 0)
    _check(ps, rts, kd, refs, precision, "auto streams")


def test_mixed_clear_cloudy_k(ctx_mw, coeff_dir, oracle_mod):
    print("allsky K mixed clear/cloudy, uneven layers", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ks = [100, 73, 100, 61, 88, 95]
    ps = synthetic.add_clouds(synthetic.make_profiles(6, ks, seed=151), seed=152, cloudy_fraction=0.5)
    eps = np.random.default_rng(10).uniform(0.45, 0.75, (6, 22))
    btk = np.ones((6, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=api.PREC_FP64, n_streams=8)
    refs = _oracle_k(oracle_mod, S, C, ps, eps, btk, 8)
    _check(ps, rts, kd, refs, api.PREC_FP64, "mixed")


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_infrared_clouds_k(coeff_dir, oracle_mod, precision):
    print(f"allsky K infrared hirs_syn precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "hirs_syn")
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    try:
        ps = synthetic.add_clouds(synthetic.make_profiles(4, 100, seed=161), seed=162, median_gm3=0.005,
                                  cover="random")
        eps = np.full((4, 19), 0.97)
        btk = np.ones((4, 19))
        rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=4)
        refs = _oracle_k(oracle_mod, S, C, ps, eps, btk, 4)
        _check(ps, rts, kd, refs, precision, "infrared")
    finally:
        ctx.destroy()


def test_device_memory_k_equals_host(ctx_mw):
    import torch
    print("allsky K device memory vs host", flush=True)
    ps = synthetic.add_clouds(synthetic.make_profiles(8, 100, seed=171), seed=172)
    eps = np.random.default_rng(11).uniform(0.45, 0.75, (8, 22))
    h_rts, h_k = ctx_mw.k_matrix(ps, eps, precision=api.PREC_FP32_SPLIT, n_streams=4)
    d_rts, d_k = ctx_mw.k_matrix(api.device_profiles(ps), torch.as_tensor(eps, device="cuda"),
                                 precision=api.PREC_FP32_SPLIT, n_streams=4)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_rts["brightness_temperature"].cpu().numpy(), h_rts["brightness_temperature"])
    for key in ("temperature_k", "cloud_water_content_k", "cloud_effective_radius_k", "cloud_fraction_k"):
        np.testing.assert_array_equal(d_k[key].cpu().numpy(), h_k[key])
