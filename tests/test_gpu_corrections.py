"""GPU parity of the antenna and non-LTE corrections (Post_Process_RTSolution,
src/CRTM_Forward_Module.f90:1166-1199; their adjoints in
Pre_Process_RTSolution_K, src/CRTM_K_Matrix_Module.f90:1725-1761) through
the C-ABI, against the FP64 oracle on the same inputs.

* amsua_ac_syn: a SpcCoeff file with an ACCoeff block (30 fields of view);
  Options%Use_Antenna_Correction with Geometry%iFOV per profile (0 = none).
* cris_nlte_syn: a SpcCoeff file with an NLTECoeff block (the 4.3 um
  channels); the correction is on by default and acts where the sun is above
  the horizon (source zenith angles on both sides of 90 degrees here).

Tolerances as the other parity tests: FP64 BT 1e-8 K and Jacobians 1e-9 of
the block maximum; FP32-split 1e-2 K and 1e-3.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor

pytestmark = pytest.mark.gpu

TOL = {0: dict(bt=1e-8, jac=1e-9), 1: dict(bt=1e-2, jac=1e-3)}


def _profiles(name):
    from crtmv3_amd import synthetic
    M = 12
    ps = synthetic.make_profiles(M, np.random.default_rng(7).integers(70, 131, M).tolist(), seed=4242)
    rng = np.random.default_rng(99)
    ps.sensor_zenith = rng.uniform(0.0, 58.0, M)
    ps.source_zenith = rng.uniform(0.0, 130.0, M)
    ps.source_zenith[:3] = (15.0, 95.0, 89.0)
    ps.ifov = rng.integers(0, synthetic.N_FOVS_AC + 1, M).astype(np.int32)
    ps.ifov[:3] = (1, synthetic.N_FOVS_AC, 0)     # 0: no antenna correction
    return ps


@pytest.mark.parametrize("memory", ["host", "device"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["amsua_ac_syn", "cris_nlte_syn"])
def test_corrections_match_oracle(coeff_dir, oracle_mod, name, prec, memory):
    import torch
    from crtmv3_amd import api, synthetic
    print(f"corrections {name} precision={prec} memory={memory}", flush=True)
    synthetic.write_sensor(coeff_dir, name)
    tau, spc = load_sensor(coeff_dir, name)
    opts = dict(use_antenna_correction=True) if name == "amsua_ac_syn" else {}
    ps = _profiles(name)
    M, L = ps.n_profiles, spc.n_channels
    eps = np.full((M, L), 0.7 if name == "amsua_ac_syn" else 0.98)
    btk = np.random.default_rng(3).uniform(0.5, 1.5, (M, L))
    s = oracle_mod.Sensor(tau, spc)
    ref = [oracle_mod.run_profile(s, ps, m, eps[m], "k", bt_k=btk[m], **opts) for m in range(M)]
    ref0 = [oracle_mod.run_profile(s, ps, m, eps[m], "forward") for m in range(M)] if opts else \
        [oracle_mod.run_profile(s, ps, m, eps[m], "forward", apply_nlte_correction=False) for m in range(M)]
    effect = max(np.abs(r["bt"] - r0["bt"]).max() for r, r0 in zip(ref, ref0))
    assert effect > 0.05, f"the correction changes the BT by {effect} K only"
    ctx = api.crtm_init([name], coeff_dir)
    try:
        if memory == "device":
            dev = torch.device("cuda:0")
            prof = api.device_profiles(ps, dev)
            e = torch.as_tensor(eps, device=dev)
            b = torch.as_tensor(btk, device=dev)
        else:
            prof, e, b = ps, eps, btk
        rts, kd = ctx.k_matrix(prof, e, bt_k=b, precision=prec, jacobian_f32=bool(prec), **opts)
        fw = ctx.forward(prof, e, precision=prec, **opts)
        _, ad = ctx.adjoint(prof, e, b, precision=prec, **opts)
        if memory == "device":
            torch.cuda.synchronize()
            rts = {k: v.cpu().numpy() for k, v in rts.items()}
            kd = {k: v.cpu().numpy() for k, v in kd.items()}
            fw = {k: v.cpu().numpy() for k, v in fw.items()}
            ad = {k: v.cpu().numpy() for k, v in ad.items()}
    finally:
        ctx.destroy()
    tol = TOL[prec]
    bt_ref = np.stack([r["bt"] for r in ref])
    for what, got in (("K-matrix BT", rts["brightness_temperature"]), ("forward BT", fw["brightness_temperature"])):
        err = np.abs(got - bt_ref).max()
        print(f"  {what} error {err:.2e} K (correction effect {effect:.2f} K)", flush=True)
        assert err <= tol["bt"], what
    np.testing.assert_allclose(rts["radiance"], np.stack([r["radiance"] for r in ref]), rtol=1e-3 if prec else 1e-10)
    for m, r in enumerate(ref):
        a, c = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        tk = np.asarray(kd["temperature_k"][a:c], np.float64).T
        ak = np.asarray(kd["absorber_k"][:, a:c, :], np.float64).transpose(2, 0, 1)
        for got, want, what in ((tk, r["t_k"], "T_K"), (ak, r["abs_k"], "absorber_K"),
                                (kd["water_temperature_k"][m], r["sfc_t_k"][:, 1], "Ts_K"),
                                (kd["surface_emissivity_k"][m], r["emis_k"], "emis_K")):
            assert_jacobian_close(got, want, tol["jac"], f"profile {m} {what}", prec)
        # CRTM_Adjoint seeded with the same BT_K: the channel sums of the K rows
        want = r["t_k"].sum(axis=0)
        blk = np.abs(ad["temperature_ad"][a:c] - want).max() / np.abs(want).max()
        assert blk <= tol["jac"], f"profile {m} temperature_ad: block-relative error {blk:.3e}"


def test_corrections_refusals(coeff_dir, oracle_mod):
    """A field-of-view index beyond the ACCoeff table fails its profile; a
    cloudy call where a correction would act is refused before any work."""
    from crtmv3_amd import api, synthetic
    synthetic.write_sensor(coeff_dir, "amsua_ac_syn")
    ctx = api.crtm_init(["amsua_ac_syn"], coeff_dir)
    try:
        ps = _profiles("amsua_ac_syn")
        L = ctx.n_channels
        eps = np.full((ps.n_profiles, L), 0.7)
        ps.ifov[4] = synthetic.N_FOVS_AC + 1
        with pytest.raises(api.CRTMError):
            ctx.forward(ps, eps, precision=0, use_antenna_correction=True)
        ps.ifov[4] = 3
        cps = synthetic.add_clouds(ps, seed=5) if hasattr(synthetic, "add_clouds") else None
        if cps is not None:
            with pytest.raises(api.CRTMError):
                ctx.forward(cps, eps, precision=0, use_antenna_correction=True, n_streams=4)
    finally:
        ctx.destroy()
