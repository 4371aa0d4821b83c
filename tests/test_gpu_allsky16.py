"""All-sky forward at 16 streams (9 quadrature angles + the sensor angle) on
the GPU against the oracle, through the C-ABI.

At 16 streams the layer and adding kernels run one wavefront per work unit
with the 16x16 stream matrices on the matrix cores.  Cases: overcast and
partial cover of the ATMS-like microwave sensor, optically thick layers
(several doublings), mixed clear / cloudy profiles with uneven layer counts,
infrared clouds with aerosols, host and device memory.  Tolerances as for
the per-lane kernels: FP64 path 1e-6 K, FP32 split path 1e-2 K.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: 1e-6, api.PREC_FP32_SPLIT: 1e-2}


@pytest.fixture(scope="module")
def ctx_mw(coeff_dir):
    synthetic.write_sensor(coeff_dir, "atms_syn")
    c = api.crtm_init(["atms_syn"], coeff_dir)
    yield c
    c.destroy()


def _oracle_sensor(coeff_dir, oracle_mod, name, aerosol=False):
    tau, spc = load_sensor(coeff_dir, name)
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir)) if aerosol else None
    return oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac)


def _oracle(orc, S, C, ps, eps, n_streams, overlap=0):
    outs = [orc.run_allsky(S, C, ps, m, eps[m], n_streams=n_streams, overlap=overlap)
            for m in range(ps.n_profiles)]
    return {k: np.stack([o[k] for o in outs]) for k in outs[0]}


def _check(out, ref, prec, label):
    err = np.abs(out["brightness_temperature"] - ref["bt"])
    print(f"  {label}: max |dBT| = {err.max():.3e} K; cloud effect up to "
          f"{np.abs(ref['bt'] - ref['tb_clear']).max():.1f} K", flush=True)
    assert err.max() < TOL[prec], label


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_overcast_16_streams(ctx_mw, coeff_dir, oracle_mod, precision):
    print(f"allsky 16 streams overcast atms_syn M=8 precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(8, 100, seed=201), seed=202)
    eps = np.random.default_rng(20).uniform(0.45, 0.75, (8, 22))
    out = ctx_mw.forward(ps, eps, precision=precision, extras=True, n_streams=16)
    ref = _oracle(oracle_mod, S, C, ps, eps, 16)
    _check(out, ref, precision, "overcast")
    np.testing.assert_allclose(out["tb_clear"], ref["tb_clear"], atol=TOL[precision])


@pytest.mark.parametrize("overlap", [2, 3])
def test_partial_cover_16_streams(ctx_mw, coeff_dir, oracle_mod, overlap):
    print(f"allsky 16 streams partial cover overlap={overlap}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(6, 90, seed=211), seed=212, cover="random")
    eps = np.random.default_rng(21).uniform(0.45, 0.75, (6, 22))
    out = ctx_mw.forward(ps, eps, precision=api.PREC_FP64, extras=True, n_streams=16, overlap=overlap)
    ref = _oracle(oracle_mod, S, C, ps, eps, 16, overlap)
    _check(out, ref, api.PREC_FP64, f"overlap {overlap}")
    np.testing.assert_allclose(out["total_cloud_cover"], ref["total_cloud_cover"], atol=1e-12)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_thick_layers_mixed_batch_16_streams(ctx_mw, coeff_dir, oracle_mod, precision):
    """Large particles and heavy water content (several Matrix_Exp doublings
    per layer), clear and cloudy profiles of uneven layer counts in one call."""
    print(f"allsky 16 streams thick layers, mixed batch precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ks = [100, 73, 100, 61, 88, 95]
    ps = synthetic.add_clouds(synthetic.make_profiles(6, ks, seed=221), seed=222, reff_scale=25.0,
                              median_gm3=0.3, cloudy_fraction=0.67)
    eps = np.random.default_rng(22).uniform(0.45, 0.75, (6, 22))
    out = ctx_mw.forward(ps, eps, precision=precision, n_streams=16)
    ref = _oracle(oracle_mod, S, C, ps, eps, 16)
    _check(out, ref, precision, "thick/mixed")


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_infrared_clouds_aerosols_16_streams(coeff_dir, oracle_mod, precision):
    print(f"allsky 16 streams infrared hirs_syn clouds + aerosols precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "hirs_syn", aerosol=True)
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    try:
        ps = synthetic.add_clouds(synthetic.make_profiles(5, 100, seed=231), seed=232, median_gm3=0.005,
                                  cover="random")
        ps = synthetic.add_aerosols(ps, seed=233)
        eps = np.full((5, 19), 0.97)
        out = ctx.forward(ps, eps, precision=precision, n_streams=16)
        ref = _oracle(oracle_mod, S, C, ps, eps, 16)
        _check(out, ref, precision, "infrared")
    finally:
        ctx.destroy()


def test_device_memory_16_streams(ctx_mw):
    import torch
    print("allsky 16 streams device memory vs host", flush=True)
    ps = synthetic.add_clouds(synthetic.make_profiles(8, 100, seed=241), seed=242)
    eps = np.random.default_rng(24).uniform(0.45, 0.75, (8, 22))
    host = ctx_mw.forward(ps, eps, precision=api.PREC_FP32_SPLIT, n_streams=16)
    dev = ctx_mw.forward(api.device_profiles(ps), torch.as_tensor(eps, device="cuda"),
                         precision=api.PREC_FP32_SPLIT, n_streams=16)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev["brightness_temperature"].cpu().numpy(), host["brightness_temperature"])


# ------------------------------------------------------------- K-matrix
KTOL = {api.PREC_FP64: dict(bt=1e-6, jac=1e-8), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


def _rel(got, want):
    got = np.asarray(got, np.float64)
    scale = np.abs(want).max()
    return np.abs(got).max() if scale == 0.0 else np.abs(got - want).max() / scale


def _check_k(ps, rts, kd, refs, prec, label, aerosols=False):
    """Jacobian blocks relative to the largest element of the block, and
    element-wise (conftest.assert_jacobian_close)."""
    tol = KTOL[prec]
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
        if aerosols:
            na = ps.aerosol_type.shape[1]
            blocks += [
                ("aerosol conc", np.asarray(kd["aerosol_concentration_k"][:, a:b, :]).transpose(2, 0, 1),
                 k["aconc_k"][:, :na]),
                ("aerosol radius", np.asarray(kd["aerosol_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1),
                 k["areff_k"][:, :na]),
            ]
        for what, got, want in blocks:
            err, elem = assert_jacobian_close(got, want, tol["jac"], f"{label} profile {m} {what}_K", prec)
            worst[what] = max(worst.get(what, 0.0), err)
            welem[what] = max(welem.get(what, 0.0), elem)
    print(f"  {label} worst Jacobian errors (block-relative / element-wise): "
          + ", ".join(f"{k} {v:.1e}/{welem[k]:.1e}" for k, v in worst.items()), flush=True)


def _oracle_k(orc, S, C, ps, eps, btk, n_streams, overlap=0):
    return [orc.k_matrix_allsky(S, C, ps, m, eps[m], n_streams, overlap, btk[m], None, method="adjoint")
            for m in range(ps.n_profiles)]


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_k_overcast_16_streams(ctx_mw, coeff_dir, oracle_mod, precision):
    print(f"allsky K 16 streams overcast atms_syn M=5 precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(5, 100, seed=251), seed=252)
    rng = np.random.default_rng(25)
    eps = rng.uniform(0.45, 0.75, (5, 22))
    btk = rng.uniform(0.5, 1.5, (5, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision),
                              n_streams=16, extras=True)
    _check_k(ps, rts, kd, _oracle_k(oracle_mod, S, C, ps, eps, btk, 16), precision, "overcast 16")
    fwd = ctx_mw.forward(ps, eps, precision=precision, n_streams=16)
    np.testing.assert_allclose(rts["brightness_temperature"], fwd["brightness_temperature"], rtol=0,
                               atol=1e-9 if precision == api.PREC_FP64 else 1e-3)


def test_k_thick_random_cover_16_streams(ctx_mw, coeff_dir, oracle_mod):
    """Large snow and rain particles under random cover, 100 layers."""
    print("allsky K 16 streams thick random cover", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(5, 100, seed=281), seed=282, cover="random", reff_scale=25.0)
    rng = np.random.default_rng(28)
    eps = rng.uniform(0.45, 0.75, (5, 22))
    btk = rng.uniform(0.5, 1.5, (5, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=api.PREC_FP64, n_streams=16)
    _check_k(ps, rts, kd, _oracle_k(oracle_mod, S, C, ps, eps, btk, 16), api.PREC_FP64, "thick random 16")


@pytest.mark.parametrize("overlap", [1, 3])
def test_k_partial_cover_thick_16_streams(ctx_mw, coeff_dir, oracle_mod, overlap):
    """Partial cover, large particles and heavy water content (several
    doublings per layer), mixed clear and cloudy profiles."""
    print(f"allsky K 16 streams partial cover, thick layers, overlap={overlap}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ks = [90, 73, 90, 61]
    ps = synthetic.add_clouds(synthetic.make_profiles(4, ks, seed=261), seed=262, cover="random",
                              reff_scale=25.0, median_gm3=0.3, cloudy_fraction=0.75)
    eps = np.random.default_rng(26).uniform(0.45, 0.75, (4, 22))
    btk = np.ones((4, 22))
    rts, kd = ctx_mw.k_matrix(ps, eps, bt_k=btk, precision=api.PREC_FP64, n_streams=16, overlap=overlap)
    _check_k(ps, rts, kd, _oracle_k(oracle_mod, S, C, ps, eps, btk, 16, overlap), api.PREC_FP64,
             f"thick overlap {overlap}")


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
def test_k_infrared_aerosols_16_streams(coeff_dir, oracle_mod, precision):
    print(f"allsky K 16 streams infrared hirs_syn clouds + aerosols precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "hirs_syn", aerosol=True)
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    try:
        ps = synthetic.add_clouds(synthetic.make_profiles(4, 100, seed=271), seed=272, median_gm3=0.005,
                                  cover="random")
        ps = synthetic.add_aerosols(ps, seed=273)
        eps = np.full((4, 19), 0.97)
        btk = np.ones((4, 19))
        rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=16)
        _check_k(ps, rts, kd, _oracle_k(oracle_mod, S, C, ps, eps, btk, 16), precision, "infrared 16",
                 aerosols=True)
    finally:
        ctx.destroy()


# ------------------------------------------- wavefront kernels at 8 streams
@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("thick", [False, True], ids=["thin", "thick"])
def test_wavefront_kernels_at_8_streams(ctx_mw, coeff_dir, oracle_mod, precision, thick):
    """crtm_set_wavefront_angles(6): the 8-stream layers with the sensor angle
    appended (6 angles) run on the wavefront kernels, a coinciding sensor angle
    (5 angles) on the per-lane ones; forward and K against the oracle, and
    against the per-lane kernels alone."""
    n_streams = 8
    print(f"wavefront kernels n_streams={n_streams} thick={thick} precision={precision}", flush=True)
    S, C = _oracle_sensor(coeff_dir, oracle_mod, "atms_syn")
    ps = synthetic.add_clouds(synthetic.make_profiles(5, 100, seed=281), seed=282, cover="random",
                              reff_scale=25.0 if thick else 1.0)
    rng = np.random.default_rng(28)
    eps = rng.uniform(0.45, 0.75, (5, 22))
    btk = rng.uniform(0.5, 1.5, (5, 22))
    kw = dict(bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=n_streams, extras=True)
    try:
        ctx_mw.set_wavefront_angles(6)
        rts_w, kd_w = ctx_mw.k_matrix(ps, eps, **kw)
        fwd_w = ctx_mw.forward(ps, eps, precision=precision, n_streams=n_streams)
    finally:
        ctx_mw.set_wavefront_angles(9)
    rts_l, kd_l = ctx_mw.k_matrix(ps, eps, **kw)
    _check_k(ps, rts_w, kd_w, _oracle_k(oracle_mod, S, C, ps, eps, btk, n_streams), precision,
             f"wavefront n={n_streams}")
    tol = 1e-9 if precision == api.PREC_FP64 else 1e-3
    np.testing.assert_allclose(fwd_w["brightness_temperature"], rts_l["brightness_temperature"], rtol=0, atol=tol)
    for key in ("temperature_k", "cloud_water_content_k", "cloud_effective_radius_k"):
        a, b = np.asarray(kd_w[key], np.float64), np.asarray(kd_l[key], np.float64)
        assert np.abs(a - b).max() <= (1e-8 if precision == api.PREC_FP64 else 1e-3) * np.abs(b).max(), key


def test_wavefront_angles_argument(ctx_mw):
    for bad in (5, 7):
        with pytest.raises(api.CRTMError, match=f"wavefront angle threshold {bad}"):
            ctx_mw.set_wavefront_angles(bad)
