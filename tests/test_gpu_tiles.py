"""Tile shapes of the fused clear-sky kernel and the autotuner that picks them
(crtm_set_tile / crtm_get_tile / crtm_autotune_k_matrix through the C-ABI).

A tile changes how many profiles share a workgroup (and with it the workgroup
order); every profile-channel still runs the same arithmetic in one lane, so
the forward and K-matrix outputs of every tile must be bit-identical to the
default's.  The autotuner must time every tile, keep the fastest per sensor,
and leave outputs equal to a plain call's.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from crtmv3_amd import api, synthetic

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    import torch
    for k in a:
        if torch.is_tensor(a[k]):
            assert torch.equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("precision", [api.PREC_FP32_SPLIT, api.PREC_FP64], ids=["split", "fp64"])
def test_tiles_bit_identical(coeff_dir, precision):
    import torch
    print(f"tiles bit-identical, precision={precision}", flush=True)
    names = ["cris_small_syn", "atms_syn", "hirs_syn"]   # ODPS groups 1, 3, 2
    for n in names:
        synthetic.write_sensor(coeff_dir, n)
    ctx = api.crtm_init(names, coeff_dir)
    dev = torch.device("cuda:0")
    try:
        L = ctx.n_channels
        M = 37                                           # not a multiple of any tile
        ps = synthetic.make_profiles(M, np.random.default_rng(5).integers(60, 121, M).tolist(), seed=77)
        dps = api.device_profiles(ps, dev)
        emis = torch.full((M, L), 0.9, dtype=torch.float64, device=dev)
        kw = dict(precision=precision, jacobian_f32=precision == api.PREC_FP32_SPLIT, check_input=False)
        for s in range(len(names)):
            assert ctx.get_tile(s, precision) == 1      # the default
        r0, k0 = ctx.k_matrix(dps, emis, **kw)
        f0 = ctx.forward(dps, emis, precision=precision, check_input=False)
        torch.cuda.synchronize()
        assert int((r0["status"] != 0).sum()) == 0
        for t in range(api.N_TILES):
            for s in range(len(names)):
                ctx.set_tile(s, precision, t)
                assert ctx.get_tile(s, precision) == t
            print(f"  tile {t}: {api.tile_profiles(precision, t)} profiles per workgroup", flush=True)
            r1, k1 = ctx.k_matrix(dps, emis, **kw)
            f1 = ctx.forward(dps, emis, precision=precision, check_input=False)
            torch.cuda.synchronize()
            _same(r0, r1, f"tile {t} RTSolution")
            _same(k0, k1, f"tile {t} K")
            _same(f0, f1, f"tile {t} forward")
        other = api.PREC_FP64 if precision == api.PREC_FP32_SPLIT else api.PREC_FP32_SPLIT
        assert ctx.get_tile(0, other) == 1               # tiles are kept per precision
        with pytest.raises(api.CRTMError):
            ctx.set_tile(0, precision, api.N_TILES)
        with pytest.raises(api.CRTMError):
            ctx.set_tile(len(names), precision, 0)
    finally:
        ctx.destroy()


def test_autotune_picks_fastest_and_caches(coeff_dir, tmp_path):
    import torch
    print("autotune: cris_small_syn + atms_syn, split", flush=True)
    names = ["cris_small_syn", "atms_syn"]
    for n in names:
        synthetic.write_sensor(coeff_dir, n)
    ctx = api.crtm_init(names, coeff_dir)
    dev = torch.device("cuda:0")
    try:
        L = ctx.n_channels
        M = 96
        ps = synthetic.make_profiles(M, 100, seed=123)
        dps = api.device_profiles(ps, dev)
        emis = torch.full((M, L), 0.95, dtype=torch.float64, device=dev)
        prec = api.PREC_FP32_SPLIT
        r0, k0 = ctx.k_matrix(dps, emis, precision=prec, jacobian_f32=True)
        cache = str(tmp_path / "tiles.json")
        rep = ctx.autotune(dps, emis, precision=prec, repeats=2, cache=cache)
        print(f"  {rep}", flush=True)
        ms = np.asarray(rep["ms"])
        assert ms.shape == (2, api.N_TILES) and (ms > 0).all()
        for s in range(2):
            assert rep["best"][s] == int(np.argmin(ms[s]))
            assert ctx.get_tile(s, prec) == rep["best"][s]
            assert rep["profiles_per_workgroup"][s] == api.tile_profiles(prec, rep["best"][s])
        assert ctx.get_tile(0, api.PREC_FP64) == 1       # the other precision is untouched
        r1, k1 = ctx.k_matrix(dps, emis, precision=prec, jacobian_f32=True)
        torch.cuda.synchronize()
        _same(r0, r1, "autotuned RTSolution")
        _same(k0, k1, "autotuned K")
        d = json.load(open(cache))
        assert len(d) == 2 and sorted(d.values()) == sorted(rep["best"])
        for s in range(2):
            ctx.set_tile(s, prec, 1)
        assert ctx.apply_tile_cache(cache, prec, M)
        assert [ctx.get_tile(s, prec) for s in range(2)] == rep["best"]
        assert not ctx.apply_tile_cache(cache, prec, M + 1)
    finally:
        ctx.destroy()
