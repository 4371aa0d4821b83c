"""HIP graph capture of a device-memory K-matrix call (hipGraph via
torch.cuda.CUDAGraph): the clear-sky path issues kernels and device copies
only, so a warmed-up call of fixed shape can be captured once and replayed
with new inputs written in place.  Every replay is checked against an eager
call on the same inputs, with eager work and synchronisations between the
replays.  A cloudy call reads its scattering-layer count back unless the
context has a layer capacity (crtm_set_cloudy_layer_capacity): without one it
fails cleanly under capture, with one it is captured and replayed like the
clear-sky call, and profiles whose layers exceed the capacity fail."""
from __future__ import annotations

import numpy as np
import pytest

from crtmv3_amd import api, synthetic

pytestmark = pytest.mark.gpu


def test_captured_k_matrix_replays_match_eager(coeff_dir):
    import torch
    print("graph capture of a clear-sky K-matrix call, 5 replays", flush=True)
    synthetic.write_sensor(coeff_dir, "hirs_syn")
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    try:
        dev = torch.device("cuda:0")
        ps = synthetic.make_profiles(64, [90, 100, 110, 95] * 16, seed=401)
        dps = api.device_profiles(ps, dev)
        L = ctx.n_channels
        emis = torch.full((64, L), 0.97, dtype=torch.float64, device=dev)
        btk = torch.ones((64, L), dtype=torch.float64, device=dev)
        t0 = dps.temperature.clone()
        s = torch.cuda.Stream()
        kw = dict(bt_k=btk, precision=api.PREC_FP32_SPLIT, jacobian_f32=True, check_input=False)
        with torch.cuda.stream(s):
            out = ctx.k_matrix(dps, emis, **kw)          # sizes the work buffers, binds the stream
            out = ctx.k_matrix(dps, emis, out=out, **kw)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            captured = ctx.k_matrix(dps, emis, out=out, **kw)
        rng = np.random.default_rng(5)
        for rep in range(5):
            shift = torch.as_tensor(rng.normal(0.0, 2.0, t0.shape[0]), device=dev)
            dps.temperature.copy_(t0 + shift)
            g.replay()
            torch.cuda.synchronize()
            got_bt = captured[0]["brightness_temperature"].clone()
            got_tk = captured[1]["temperature_k"].clone()
            eager = ctx.k_matrix(dps, emis, **kw)       # eager work between replays
            torch.cuda.synchronize()
            assert torch.equal(got_bt, eager[0]["brightness_temperature"]), rep
            assert torch.equal(got_tk, eager[1]["temperature_k"]), rep
        dps.temperature.copy_(t0)
    finally:
        ctx.destroy()


def test_cloudy_call_refuses_capture(coeff_dir):
    import torch
    synthetic.write_sensor(coeff_dir, "atms_syn")
    ctx = api.crtm_init(["atms_syn"], coeff_dir)
    try:
        dev = torch.device("cuda:0")
        ps = synthetic.add_clouds(synthetic.make_profiles(4, 80, seed=402), seed=403)
        dps = api.device_profiles(ps, dev)
        emis = torch.full((4, 22), 0.6, dtype=torch.float64, device=dev)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            ctx.forward(dps, emis, n_streams=4)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with pytest.raises(api.CRTMError, match="graph capture"):
            with torch.cuda.graph(g, stream=s):
                ctx.forward(dps, emis, n_streams=4)
    finally:
        ctx.destroy()


def _cloudy_batch(coeff_dir, dev):
    synthetic.write_sensor(coeff_dir, "atms_syn")
    synthetic.write_sensor(coeff_dir, "hirs_syn")
    rng = np.random.default_rng(411)
    ps = synthetic.make_profiles(24, rng.integers(70, 111, 24).tolist(), seed=412)
    ps = synthetic.add_clouds(ps, seed=413, types=(synthetic.WATER_CLOUD, synthetic.ICE_CLOUD), cover="random",
                              cloudy_fraction=0.6)
    ps = synthetic.add_aerosols(ps, seed=414, aerosol_fraction=0.5)
    return ps, api.device_profiles(ps, dev)


def _k_keys(kd):
    return [k for k in ("temperature_k", "absorber_k", "cloud_water_content_k", "cloud_effective_radius_k",
                        "cloud_fraction_k", "aerosol_concentration_k") if k in kd]


def test_capacity_mode_matches_counted_launches(coeff_dir):
    """Launches sized for a layer capacity give the same bits as launches sized
    for the counted layers."""
    import torch
    print("cloudy K-matrix with a layer capacity vs the counted layers", flush=True)
    dev = torch.device("cuda:0")
    ps, dps = _cloudy_batch(coeff_dir, dev)
    ref = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    cap = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    try:
        cap.set_cloudy_layer_capacity(int(ps.layer_offset[-1]))
        L = ref.n_channels
        emis = torch.full((24, L), 0.9, dtype=torch.float64, device=dev)
        btk = torch.ones((24, L), dtype=torch.float64, device=dev)
        kw = dict(bt_k=btk, precision=api.PREC_FP32_SPLIT, jacobian_f32=True, n_streams=4)
        r0, k0 = ref.k_matrix(dps, emis, **kw)
        r1, k1 = cap.k_matrix(dps, emis, **kw)
        torch.cuda.synchronize()
        assert int((r0["status"] != 0).sum()) == 0
        assert torch.equal(r0["brightness_temperature"], r1["brightness_temperature"])
        assert torch.equal(r0["status"], r1["status"])
        for key in _k_keys(k0):
            assert torch.equal(k0[key], k1[key]), key
    finally:
        ref.destroy()
        cap.destroy()


def test_captured_cloudy_k_matrix_replays_match_eager(coeff_dir):
    """A cloudy K-matrix call (clouds and aerosols, two sensors) captured with a
    layer capacity; each replay on new cloud water contents and temperatures
    equals an eager call of a context without a capacity."""
    import torch
    print("graph capture of a cloudy K-matrix call, 4 replays", flush=True)
    dev = torch.device("cuda:0")
    ps, dps = _cloudy_batch(coeff_dir, dev)
    ref = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    ctx = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    try:
        ctx.set_cloudy_layer_capacity(int(ps.layer_offset[-1]))
        L = ctx.n_channels
        emis = torch.full((24, L), 0.9, dtype=torch.float64, device=dev)
        btk = torch.ones((24, L), dtype=torch.float64, device=dev)
        kw = dict(bt_k=btk, precision=api.PREC_FP32_SPLIT, jacobian_f32=True, n_streams=4, check_input=False)
        t0, w0 = dps.temperature.clone(), dps.cloud_water_content.clone()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            out = ctx.k_matrix(dps, emis, **kw)          # sizes the work space, binds the stream
            out = ctx.k_matrix(dps, emis, out=out, **kw)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            captured = ctx.k_matrix(dps, emis, out=out, **kw)
        rng = np.random.default_rng(415)
        for rep in range(4):
            dps.temperature.copy_(t0 + torch.as_tensor(rng.normal(0.0, 2.0, t0.shape[0]), device=dev))
            dps.cloud_water_content.copy_(w0 * torch.as_tensor(rng.uniform(0.5, 2.0, tuple(w0.shape)), device=dev))
            g.replay()
            torch.cuda.synchronize()
            got_r = {k: v.clone() for k, v in captured[0].items() if k in ("brightness_temperature", "status")}
            got_k = {k: captured[1][k].clone() for k in _k_keys(captured[1])}
            r_e, k_e = ref.k_matrix(dps, emis, **kw)     # eager work between replays
            torch.cuda.synchronize()
            assert int((r_e["status"] != 0).sum()) == 0, rep
            assert torch.equal(got_r["status"], r_e["status"]), rep
            assert torch.equal(got_r["brightness_temperature"], r_e["brightness_temperature"]), rep
            for key in got_k:
                assert torch.equal(got_k[key], k_e[key]), (rep, key)
        dps.temperature.copy_(t0)
        dps.cloud_water_content.copy_(w0)
    finally:
        ref.destroy()
        ctx.destroy()


def test_captured_cloudy_call_with_per_profile_streams(coeff_dir):
    """Per-profile stream counts (Options(m)%n_Streams) in device memory
    inside a captured cloudy K-matrix call: replays equal eager calls with
    the same counts, and a count the kernels do not take fails its profile
    (status 8) without failing the others."""
    import torch
    print("graph capture of a cloudy K-matrix call with per-profile stream counts", flush=True)
    dev = torch.device("cuda:0")
    ps, dps = _cloudy_batch(coeff_dir, dev)
    ref = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    ctx = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    try:
        ctx.set_cloudy_layer_capacity(int(ps.layer_offset[-1]))
        L = ctx.n_channels
        M = ps.n_profiles
        emis = torch.full((M, L), 0.9, dtype=torch.float64, device=dev)
        btk = torch.ones((M, L), dtype=torch.float64, device=dev)
        counts = torch.as_tensor(np.resize([4, 8, 16, 0, 6], M), dtype=torch.int32, device=dev)
        kw = dict(bt_k=btk, precision=api.PREC_FP32_SPLIT, jacobian_f32=True, check_input=False)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            out = ctx.k_matrix(dps, emis, n_streams=counts, **kw)
            out = ctx.k_matrix(dps, emis, n_streams=counts, out=out, **kw)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            captured = ctx.k_matrix(dps, emis, n_streams=counts, out=out, **kw)
        base = counts.clone()
        for rep, bad in enumerate((None, 3)):
            counts.copy_(base)
            if bad is not None:
                counts[bad] = 5
            g.replay()
            torch.cuda.synchronize()
            got_st = captured[0]["status"].clone()
            got_bt = captured[0]["brightness_temperature"].clone()
            got_k = {k: captured[1][k].clone() for k in _k_keys(captured[1])}
            if bad is None:
                r_e, k_e = ref.k_matrix(dps, emis, n_streams=counts, **kw)
                torch.cuda.synchronize()
                assert int((r_e["status"] != 0).sum()) == 0
                assert torch.equal(got_st, r_e["status"])
                assert torch.equal(got_bt, r_e["brightness_temperature"])
                for key in got_k:
                    assert torch.equal(got_k[key], k_e[key]), key
            else:
                st = got_st.cpu().numpy()
                assert st[bad] == 8 and np.all(np.delete(st, bad) == 0), st
                assert torch.isnan(got_bt[bad]).all()
                keep = [m for m in range(M) if m != bad]
                assert torch.equal(got_bt[keep], r_e["brightness_temperature"][keep])
        counts.copy_(base)
    finally:
        ref.destroy()
        ctx.destroy()


def _scattering_layer_counts(ps):
    """Scattering layers per profile as the cloud preparation counts them:
    a typed cloud with water content above 1e-6 kg/m^2 in a layer whose cloud
    fraction is at least 1e-12, or a typed aerosol above 1e-9 kg/m^2."""
    M = ps.n_profiles
    counts = np.zeros(M, np.int64)
    for m in range(M):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        scat = np.zeros(b - a, bool)
        for n in range(ps.cloud_type.shape[1]):
            if 1 <= ps.cloud_type[m, n] <= 6:
                scat |= (ps.cloud_water_content[n, a:b] > 1e-6) & (ps.cloud_fraction[a:b] >= 1e-12)
        for n in range(ps.aerosol_type.shape[1]):
            if ps.aerosol_type[m, n] != 0:
                scat |= ps.aerosol_concentration[n, a:b] > 1e-9
        counts[m] = int(scat.sum())
    return counts


def test_layer_capacity_overflow_fails_profiles(coeff_dir):
    """With room for the scattering layers of the first half of the batch,
    slots are numbered in profile order: exactly the profiles whose layers end
    beyond the capacity fail (status 6, NaN brightness temperatures), the same
    ones on every call; every other profile equals the counted-layer result."""
    import torch
    print("layer capacity overflow", flush=True)
    dev = torch.device("cuda:0")
    ps, dps = _cloudy_batch(coeff_dir, dev)
    counts = _scattering_layer_counts(ps)
    base = np.concatenate([[0], np.cumsum(counts)[:-1]])
    capacity = int(np.cumsum(counts)[11])          # the first 12 profiles fit
    expect_bad = (counts > 0) & (base + counts > capacity)
    assert expect_bad.any() and (~expect_bad).any()
    ref = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    cap = api.crtm_init(["atms_syn", "hirs_syn"], coeff_dir)
    try:
        cap.set_cloudy_layer_capacity(capacity)
        L = ref.n_channels
        emis = torch.full((24, L), 0.9, dtype=torch.float64, device=dev)
        r0 = ref.forward(dps, emis, precision=api.PREC_FP64, n_streams=4)
        r1 = cap.forward(dps, emis, precision=api.PREC_FP64, n_streams=4)
        r2 = cap.forward(dps, emis, precision=api.PREC_FP64, n_streams=4)
        torch.cuda.synchronize()
        st1, st2 = r1["status"].cpu().numpy(), r2["status"].cpu().numpy()
        np.testing.assert_array_equal(st1 != 0, expect_bad)
        np.testing.assert_array_equal(st1, st2)
        assert (st1[expect_bad] == 6).all()
        bad = expect_bad
        bt0, bt1 = r0["brightness_temperature"].cpu().numpy(), r1["brightness_temperature"].cpu().numpy()
        assert np.isnan(bt1[bad]).all()
        np.testing.assert_array_equal(bt1[~bad], bt0[~bad])
        with pytest.raises(api.CRTMError, match="capacity"):
            cap.set_cloudy_layer_capacity(-1)
    finally:
        ref.destroy()
        cap.destroy()
