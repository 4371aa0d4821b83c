"""crtm_k_matrix_noalloc (the K-matrix entry point that allocates nothing)
against crtm_k_matrix, both through the C-ABI on device-memory batches.

The two entry points run the same kernels; only the work space differs (the
library's fixed scratch array instead of its pool), so every output must be
bit-identical.  Two no-allocation calls in a row, the second on a smaller
batch, check that the array's leftover contents from the first call do not
leak into the second.  Batches beyond the array, cloudy batches and host
batches are refused before any work is issued.
"""
from __future__ import annotations

import numpy as np
import pytest

from crtmv3_amd import api, synthetic

pytestmark = pytest.mark.gpu

_KEYS_R = ("radiance", "brightness_temperature", "status")


def _same(a, b, what):
    import torch
    for k in a:
        if not torch.is_tensor(a[k]):
            continue
        assert torch.equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("precision", [api.PREC_FP32_SPLIT, api.PREC_FP64], ids=["split", "fp64"])
def test_noalloc_equals_k_matrix(coeff_dir, precision):
    import torch
    print(f"no-allocation K-matrix vs crtm_k_matrix, precision={precision}", flush=True)
    synthetic.write_sensor(coeff_dir, "cris_small_syn")
    synthetic.write_sensor(coeff_dir, "amsua_syn")
    ctx = api.crtm_init(["cris_small_syn", "amsua_syn"], coeff_dir)
    dev = torch.device("cuda:0")
    try:
        L = ctx.n_channels
        kw = dict(precision=precision, jacobian_f32=precision == api.PREC_FP32_SPLIT, check_input=False)
        for M, seed in ((48, 901), (13, 902)):     # the second call is smaller
            ps = synthetic.make_profiles(M, np.random.default_rng(seed).integers(60, 121, M).tolist(), seed=seed)
            dps = api.device_profiles(ps, dev)
            emis = torch.full((M, L), 0.97, dtype=torch.float64, device=dev)
            btk = torch.rand((M, L), dtype=torch.float64, device=dev,
                             generator=torch.Generator(device=dev).manual_seed(seed)) + 0.5
            r0, k0 = ctx.k_matrix(dps, emis, bt_k=btk, **kw)
            r1, k1 = ctx.k_matrix(dps, emis, bt_k=btk, no_alloc=True, **kw)
            torch.cuda.synchronize()
            print(f"  M={M}: max layers {int(ps.n_layers.max())}", flush=True)
            assert int((r1["status"] != 0).sum()) == 0
            _same(r0, r1, f"M={M} RTSolution")
            _same(k0, k1, f"M={M} K")
            assert torch.isfinite(k1["temperature_k"]).all()
    finally:
        ctx.destroy()


def test_noalloc_refusals(coeff_dir):
    import torch
    print("no-allocation K-matrix refusals", flush=True)
    synthetic.write_sensor(coeff_dir, "hirs_syn")
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    dev = torch.device("cuda:0")
    try:
        L = ctx.n_channels
        # beyond the 128 MiB scratch array: refused with the size in the message
        big = synthetic.make_profiles(2000, 137, seed=911)
        dbig = api.device_profiles(big, dev)
        ebig = torch.full((2000, L), 0.97, dtype=torch.float64, device=dev)
        with pytest.raises(api.CRTMError, match="scratch array"):
            ctx.k_matrix(dbig, ebig, precision=api.PREC_FP64, no_alloc=True)
        del dbig, ebig
        # cloudy batches and host batches are refused
        ps = synthetic.add_clouds(synthetic.make_profiles(4, 80, seed=912), seed=913)
        with pytest.raises(api.CRTMError, match="clear-sky"):
            ctx.k_matrix(api.device_profiles(ps, dev), torch.full((4, L), 0.97, dtype=torch.float64, device=dev),
                         no_alloc=True)
        clear = synthetic.make_profiles(4, 80, seed=914)
        with pytest.raises(api.CRTMError, match="device-memory"):
            ctx.k_matrix(clear, np.full((4, L), 0.97), no_alloc=True)
        # the context still works, in both entry points
        dcl = api.device_profiles(clear, dev)
        e = torch.full((4, L), 0.97, dtype=torch.float64, device=dev)
        r0, k0 = ctx.k_matrix(dcl, e, precision=api.PREC_FP64)
        r1, k1 = ctx.k_matrix(dcl, e, precision=api.PREC_FP64, no_alloc=True)
        torch.cuda.synchronize()
        _same(r0, r1, "after refusals RTSolution")
        _same(k0, k1, "after refusals K")
    finally:
        ctx.destroy()
