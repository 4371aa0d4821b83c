"""Properties at the benchmark's shape (2211-channel CrIS-like ODPS group 1
sensor with OPTRAN, 100-layer profiles, FP32-split): the split path stays
within 1e-2 K of the FP64 GPU path and its Jacobians within 1e-3, the
K-matrix is linear in the BT adjoint seed, and a sample of profiles matches
the FP64 oracle."""
from __future__ import annotations

import numpy as np

from conftest import assert_jacobian_close
import pytest

pytestmark = pytest.mark.gpu

M = 512


@pytest.fixture(scope="module")
def big(tmp_path_factory):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from crtmv3_amd import api, synthetic
    d = synthetic.default_coeff_dir()
    synthetic.write_sensor(d, "cris_syn")
    ctx = api.crtm_init(["cris_syn"], d)
    ps = synthetic.make_profiles(M, 100, seed=20261203 + 2)
    dev = torch.device("cuda:0")
    dps = api.device_profiles(ps, dev)
    eps = torch.full((M, ctx.n_channels), 0.98, dtype=torch.float64, device=dev)
    yield ctx, ps, dps, eps, d
    ctx.destroy()


def test_split_vs_fp64_at_full_channel_count(big):
    from crtmv3_amd import api
    ctx, ps, dps, eps, _ = big
    print("case split-vs-fp64 cris_syn M=512", flush=True)
    r1, k1 = ctx.k_matrix(dps, eps, precision=api.PREC_FP32_SPLIT, jacobian_f32=True)
    r0, k0 = ctx.k_matrix(dps, eps, precision=api.PREC_FP64)
    ctx.synchronize()
    bt_err = (r1["brightness_temperature"] - r0["brightness_temperature"]).abs().max().item()
    assert bt_err <= 1e-2
    for key in ("temperature_k", "absorber_k"):
        rel = ((k1[key].double() - k0[key]).abs().max() / k0[key].abs().max()).item()
        assert rel <= 1e-3, (key, rel)
    # the K-matrix forward half equals CRTM_Forward
    f0 = ctx.forward(dps, eps, precision=api.PREC_FP64)
    ctx.synchronize()
    assert (f0["brightness_temperature"] - r0["brightness_temperature"]).abs().max().item() == 0.0


def test_k_matrix_linear_in_seed(big):
    import torch
    from crtmv3_amd import api
    ctx, ps, dps, eps, _ = big
    print("case seed linearity", flush=True)
    seed = torch.ones_like(eps)
    _, ka = ctx.k_matrix(dps, eps, bt_k=seed, precision=api.PREC_FP64)
    _, kb = ctx.k_matrix(dps, eps, bt_k=-3.0 * seed, precision=api.PREC_FP64)
    ctx.synchronize()
    # exact up to FP64 rounding of the layer sums (relative to the block maximum)
    for key in ("temperature_k", "absorber_k"):
        a, b = ka[key], kb[key]
        assert ((b + 3.0 * a).abs().max() / a.abs().max()).item() <= 1e-13, key


def test_sample_matches_oracle(big, oracle_mod):
    from crtmv3_amd import api, coeffio
    import os
    ctx, ps, dps, eps, d = big
    print("case oracle sample cris_syn", flush=True)
    r, k = ctx.k_matrix(dps, eps, precision=api.PREC_FP32_SPLIT, jacobian_f32=True)
    ctx.synchronize()
    tau = coeffio.read_odps(os.path.join(d, "cris_syn.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(d, "cris_syn.SpcCoeff.bin"))
    s = oracle_mod.Sensor(tau, spc)
    e = eps[0].cpu().numpy()
    bt = r["brightness_temperature"].cpu().numpy()
    tk = k["temperature_k"].cpu().numpy()
    for m in (0, 137, 511):
        ref = oracle_mod.run_profile(s, ps, m, e, "k")
        assert np.abs(bt[m] - ref["bt"]).max() <= 1e-2
        got = tk[100 * m:100 * (m + 1)].T.astype(np.float64)
        assert_jacobian_close(got, ref["t_k"], 1e-3, f"cris_syn profile {m} T_K", api.PREC_FP32_SPLIT)
