"""The C-ABI library without a GPU: it loads, exports every function
include/crtm_mi355x.h declares, and the ctypes mirror in crtmv3_amd/api.py
has the header's struct layouts (field names and order)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "crtm_mi355x.h")


def _declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(crtm_\w+)\s*\(", txt, re.M)))


def _header_structs():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(\w+);", txt, re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if decl:
                fields.append(re.findall(r"(\w+)\s*$", decl)[0])
        out[name] = fields
    return out


def test_header_declares_api():
    fns = _declared_functions()
    for f in ("crtm_init", "crtm_destroy", "crtm_forward", "crtm_k_matrix", "crtm_channel_info",
              "crtm_n_channels_total", "crtm_synchronize", "crtm_last_error"):
        assert f in fns


def test_library_exports_every_declared_symbol():
    from crtmv3_amd import api
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libcrtm_mi355x.so not built")
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [f for f in _declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    # and they are C symbols (no C++ mangling)
    nm = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    for f in _declared_functions():
        assert re.search(rf"\bT {f}$", nm, re.M), f


def test_ctypes_structs_match_header():
    from crtmv3_amd import api
    hs = _header_structs()
    pairs = {"crtm_atmosphere": api.CAtmosphere, "crtm_surface": api.CSurface, "crtm_geometry": api.CGeometry,
             "crtm_options": api.COptions, "crtm_rtsolution": api.CRTSolution, "crtm_k_output": api.CKOutput,
             "crtm_ad_output": api.CADOutput}
    for cname, py in pairs.items():
        assert [f[0] for f in py._fields_] == hs[cname], cname


def test_null_context_is_an_error_not_a_crash():
    """Entry points validate their arguments before touching the GPU."""
    from crtmv3_amd import api
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libcrtm_mi355x.so not built")
    L = api.lib()
    assert L.crtm_forward(None, None, None, None, None, None) == api.FAILURE
    assert L.crtm_k_matrix(None, None, None, None, None, None, None) == api.FAILURE
    assert L.crtm_k_matrix_noalloc(None, None, None, None, None, None, None) == api.FAILURE
    assert L.crtm_adjoint(None, None, None, None, None, None, None) == api.FAILURE
    assert L.crtm_destroy(None) == api.FAILURE
    assert L.crtm_last_error(None) is not None


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from crtmv3_amd import api
    monkeypatch.setattr(api, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(api, "_lib", None)
    with pytest.raises(api.CRTMError):
        api.lib()


def test_profile_slice_rebases_offsets():
    """api.profile_slice: a sub-batch of profiles with its own layer offsets."""
    import numpy as np
    from crtmv3_amd import api, synthetic
    ps = synthetic.add_aerosols(synthetic.add_clouds(synthetic.make_profiles(5, [60, 80, 100, 70, 90], seed=1),
                                                     seed=2), seed=3)
    sub = api.profile_slice(ps, 1, 4)
    a, b = int(ps.layer_offset[1]), int(ps.layer_offset[4])
    np.testing.assert_array_equal(sub.n_layers, [80, 100, 70])
    np.testing.assert_array_equal(sub.layer_offset, [0, 80, 180, 250])
    np.testing.assert_array_equal(sub.temperature, ps.temperature[a:b])
    np.testing.assert_array_equal(sub.absorber, ps.absorber[:, a:b])
    np.testing.assert_array_equal(sub.cloud_water_content, ps.cloud_water_content[:, a:b])
    np.testing.assert_array_equal(sub.aerosol_type, ps.aerosol_type[1:4])
    np.testing.assert_array_equal(sub.level_pressure, ps.level_pressure[a + 1:b + 4])
    np.testing.assert_array_equal(sub.water_temperature, ps.water_temperature[1:4])
