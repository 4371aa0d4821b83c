"""HIP graph capture of a device-memory K-matrix call (hipGraph via
torch.cuda.CUDAGraph): the clear-sky path issues kernels and device copies
only, so a warmed-up call of fixed shape can be captured once and replayed
with new inputs written in place.  Every replay is checked against an eager
call on the same inputs, with eager work and synchronisations between the
replays.  Cloudy and host-memory calls, which synchronise with the host, fail
cleanly under capture."""
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
