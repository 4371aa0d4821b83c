"""The reference's clamps, forced on purpose, on the GPU against the oracle
(through the C-ABI):

* the layer optical-depth clamp to [0, 20] (ODPS_AtmAbsorption.f90:213-222),
  whose adjoint is zero where it acts (:545-555): one channel's
  coefficients negated (OD < 0 in every layer), another's scaled up (OD > 20
  in the lower layers);
* the OPTRAN ln(chi) limits +-20 (Add_OPTRAN_wloOD, :711-718): an OPTRAN
  channel's b0 shifted by +30 and another's by -30;
* the absorber Min/Max clamps of Map_Input (ODPS_CoordinateMapping.f90:
  271-283), which the adjoint ignores (:517): water vapour far above the
  table's maximum in some layers and below its minimum in others, ozone
  above its maximum.

Each case asserts that the clamp really acts (from the oracle's own
intermediate quantities) and compares forward and K-matrix with the same
tolerances as the other parity tests, element-wise included.
"""
from __future__ import annotations

import copy
import ctypes
import os

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: dict(bt=1e-8, jac=1e-9), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


def _write(tmp_path, name, tau, spc):
    d = str(tmp_path)
    coeffio.write_odps(os.path.join(d, f"{name}.TauCoeff.bin"), tau)
    coeffio.write_spccoeff(os.path.join(d, f"{name}.SpcCoeff.bin"), spc)
    return d


def _compare(ctx, S, oracle_mod, ps, label, channels=None):
    M, L = ps.n_profiles, ctx.n_channels
    rng = np.random.default_rng(61)
    eps = rng.uniform(0.6, 0.98, (M, L))
    btk = rng.uniform(0.5, 1.5, (M, L))
    refs = [oracle_mod.run_profile(S, ps, m, eps[m], "k", bt_k=btk[m]) for m in range(M)]
    for prec in (api.PREC_FP64, api.PREC_FP32_SPLIT):
        tol = TOL[prec]
        rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=prec, jacobian_f32=bool(prec))
        for m, r in enumerate(refs):
            a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
            ch = slice(None) if channels is None else channels
            dbt = np.abs(rts["brightness_temperature"][m][ch] - r["bt"][ch]).max()
            assert dbt <= tol["bt"], f"{label} prec {prec} profile {m}: |dBT| {dbt:.3e}"
            tk = np.asarray(kd["temperature_k"][a:b], np.float64).T
            ak = np.asarray(kd["absorber_k"][:, a:b, :], np.float64).transpose(2, 0, 1)
            assert_jacobian_close(tk, r["t_k"], tol["jac"], f"{label} prec {prec} profile {m} T_K", prec)
            assert_jacobian_close(ak, r["abs_k"], tol["jac"], f"{label} prec {prec} profile {m} absorber_K", prec)
    return refs


@pytest.mark.parametrize("name", ["hirs_syn", "cris_small_syn"])
def test_optical_depth_clamps(coeff_dir, oracle_mod, tmp_path, name):
    print(f"OD clamps {name}", flush=True)
    tau, spc = load_sensor(coeff_dir, name)
    t = copy.deepcopy(tau)
    t.c = tau.c.copy()
    Ko = tau.n_layers
    neg, big = 0, 1
    for l, scale in ((neg, -1.0), (big, 3.0e3)):
        if tau.osignificance is not None:
            t.osignificance = t.osignificance.copy()
            t.osignificance[l] = 0            # the plain regression for these two channels
        for j in range(tau.n_components):
            n = int(tau.n_predictors[j, l])
            if n:
                p0 = int(tau.pos_index[j, l]) - 1
                t.c[p0:p0 + n * Ko] = (tau.c[p0:p0 + n * Ko] * scale).astype(np.float32)
    d = _write(tmp_path, name, t, spc)
    S = oracle_mod.Sensor(t, spc)
    ctx = api.crtm_init([name], d)
    try:
        ps = synthetic.make_profiles(4, [100, 90, 120, 100], seed=601)
        refs = _compare(ctx, S, oracle_mod, ps, f"OD clamps {name}")
        # the clamps act: no absorption at all in the negated channel, layers
        # at the 20 cap in the scaled one
        for r in refs:
            assert (r["tau"][neg] == 0.0).all()
            assert np.abs(r["t_k"][neg]).max() == 0.0           # the clamp kills the adjoint
            assert (r["tau"][big] > 0).all()
        assert max((r["tau"][big] * 1.25 >= 19.0).sum() for r in refs) > 0
    finally:
        ctx.destroy()


def test_optran_ln_chi_limits(coeff_dir, oracle_mod, tmp_path):
    name = "cris_small_syn"
    print(f"OPTRAN +-20 limits {name}", flush=True)
    tau, spc = load_sensor(coeff_dir, name)
    opt = np.flatnonzero(tau.osignificance == 1)
    assert opt.size >= 2
    t = copy.deepcopy(tau)
    t.oc = tau.oc.copy()
    hi, lo = int(opt[0]), int(opt[1])
    for l, shift in ((hi, 30.0), (lo, -30.0)):
        t.oc[int(tau.opos_index[l]) - 1] += shift          # b0's constant term
    d = _write(tmp_path, name, t, spc)
    S = oracle_mod.Sensor(t, spc)
    ctx = api.crtm_init([name], d)
    try:
        ps = synthetic.make_profiles(3, [100, 77, 100], seed=611)
        refs = _compare(ctx, S, oracle_mod, ps, "OPTRAN limits")
        base = [oracle_mod.run_profile(oracle_mod.Sensor(tau, spc), ps, m, np.full(tau.n_channels, 0.9))
                for m in range(3)]
        for r, b0 in zip(refs, base):
            assert (r["tau"][lo] < b0["tau"][lo]).any()      # chi = 0: the WLO part vanishes
            assert (r["tau"][hi] > b0["tau"][hi]).all()      # chi = 4.8e8 (then the OD cap)
    finally:
        ctx.destroy()


def _predictors(orc, s, ps, m):
    L = orc.lib()
    L.oracle_predictors.argtypes = [ctypes.POINTER(orc.OrSensor), ctypes.POINTER(orc.OrProfile)] + \
        [ctypes.c_void_p] * 4
    keep = []
    p = orc.make_profile(ps, m, keep)
    Ko, J = s.s.n_layers, s.s.n_components
    x = np.zeros((Ko, J, 18))
    sec, t = np.zeros(Ko), np.zeros(Ko)
    a = np.zeros((s.s.n_absorbers, Ko))
    L.oracle_predictors(ctypes.byref(s.s), ctypes.byref(p), x.ctypes.data, sec.ctypes.data, t.ctypes.data,
                        a.ctypes.data)
    return a


@pytest.mark.parametrize("name", ["amsua_syn", "hirs_syn", "cris_small_syn"])
def test_absorber_min_max_clamps(coeff_dir, oracle_mod, name):
    print(f"absorber Min/Max clamps {name}", flush=True)
    tau, spc = load_sensor(coeff_dir, name)
    S = oracle_mod.Sensor(tau, spc)
    ps = synthetic.make_profiles(3, 100, seed=621)
    absorber = ps.absorber.copy()
    for m in range(3):
        a = int(ps.layer_offset[m])
        absorber[0, a + 55:a + 80] *= 5000.0      # water vapour above the table maximum (1000 x reference)
        absorber[0, a + 10:a + 35] *= 1e-6        # and below its minimum (1e-3 x reference)
        absorber[1, a + 38:a + 52] *= 5000.0      # ozone above its maximum
    ps.absorber[:] = absorber
    ctx = api.crtm_init([name], coeff_dir)
    try:
        _compare(ctx, S, oracle_mod, ps, f"absorber clamps {name}")
        # the mapped absorbers sit on the table bounds in some layers
        a_odps = _predictors(oracle_mod, S, ps, 0)
        ids = list(tau.absorber_id)
        jw = ids.index(1)
        assert np.isclose(a_odps[jw], tau.max_absorber[:, jw], rtol=1e-12).sum() >= 3
        assert np.isclose(a_odps[jw], tau.min_absorber[:, jw], rtol=1e-12).sum() >= 3
        if 3 in ids:                              # the microwave tables carry no ozone
            jo = ids.index(3)
            assert np.isclose(a_odps[jo], tau.max_absorber[:, jo], rtol=1e-12).sum() >= 3
    finally:
        ctx.destroy()
