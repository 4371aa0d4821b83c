"""CRTM binary files of Atmosphere, Cloud, Aerosol, Geometry and RTSolution
structures (crtmv3_amd/fileio.py).

Pins: Geometry and Cloud files are written by the reference's own
CRTM_Geometry_WriteFile / CRTM_Cloud_WriteFile (built unmodified in
oracle/_ref) and must be byte-identical to this package's files of the same
values, and each side reads the other's file to the same values.  The
Atmosphere file embeds its cloud block exactly as a Cloud file after the
magic number (the reference reads it with CRTM_Cloud_ReadFile on the open
unit, src/Atmosphere/CRTM_Atmosphere_Define.f90:2555-2562), so its cloud part
is checked against the reference-written Cloud file byte for byte.  The
Atmosphere and RTSolution modules themselves reach the netCDF library
(CRTM_AerosolCoeff, `USE netcdf`) and cannot be built here: their own
records are checked by round trips and by record sizes fixed by the record
lists (src/RTSolution/CRTM_RTSolution_Define.f90:2272-2356), so those
records are "parity unpinned" against a reference build.
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np
import pytest

from crtmv3_amd import fileio, synthetic


def _ref():
    from oracle import oracle as orc
    if not os.path.exists(orc.REF_PATH):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    L = ctypes.CDLL(orc.REF_PATH)
    c, i, v = ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p
    L.ref_geometry_write.argtypes = [c, i, i, v, v]
    L.ref_geometry_read.argtypes = [c, i, i, v, v, v]
    L.ref_cloud_write.argtypes = [c, i, i, i, v, v, v, v]
    L.ref_cloud_read.argtypes = [c, i, i, i, v, v, v, v, v, v]
    return L


def _records(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        out.append(raw[pos + 4:pos + 4 + n])
        assert struct.unpack_from("<i", raw, pos + 4 + n)[0] == n
        pos += n + 8
    return out


def _geo(M, seed=3):
    rng = np.random.default_rng(seed)
    g = {k: rng.uniform(-80, 80, M) for k in fileio.GEO_REALS}
    g["ifov"] = rng.integers(0, 31, M).astype(np.int32)
    g["year"] = rng.integers(1990, 2030, M).astype(np.int32)
    g["month"] = rng.integers(1, 13, M).astype(np.int32)
    g["day"] = rng.integers(1, 29, M).astype(np.int32)
    return g


def test_geometry_file_matches_reference(tmp_path):
    L = _ref()
    M = 7
    g = _geo(M)
    ours, theirs = str(tmp_path / "ours.bin"), str(tmp_path / "ref.bin")
    fileio.write_geometry(ours, g)
    ints = np.ascontiguousarray(np.stack([g[k] for k in fileio.GEO_INTS]).T, np.int32)     # [M, 4]
    reals = np.ascontiguousarray(np.stack([g[k] for k in fileio.GEO_REALS]).T, np.float64)  # [M, 9]
    p = theirs.encode()
    assert L.ref_geometry_write(p, len(p), M, ints.ctypes.data, reals.ctypes.data) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    back = fileio.read_geometry(theirs)
    for k in fileio.GEO_INTS + fileio.GEO_REALS:
        np.testing.assert_array_equal(back[k], g[k], err_msg=k)
    mout = ctypes.c_int()
    ri = np.zeros((M, 4), np.int32)
    rr = np.zeros((M, 9))
    q = ours.encode()
    assert L.ref_geometry_read(q, len(q), M, ctypes.byref(mout), ri.ctypes.data, rr.ctypes.data) == 0
    assert mout.value == M
    np.testing.assert_array_equal(ri, ints)
    np.testing.assert_array_equal(rr, reals)


def _clouds(nc, K, seed=5):
    rng = np.random.default_rng(seed)
    return (rng.integers(1, 7, nc).astype(np.int32), rng.uniform(5, 500, (nc, K)), rng.uniform(0, 1, (nc, K)),
            rng.lognormal(-3, 1, (nc, K)))


def test_cloud_file_matches_reference(tmp_path):
    L = _ref()
    nc, K = 3, 11
    ty, re_, rv, wc = _clouds(nc, K)
    ours, theirs = str(tmp_path / "ours.bin"), str(tmp_path / "ref.bin")
    fileio.write_cloud_file(ours, ty, re_, rv, wc)
    f = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731  [nc, K] C order = (K, nc) Fortran
    p = theirs.encode()
    assert L.ref_cloud_write(p, len(p), nc, K, ty.ctypes.data, f(re_).ctypes.data, f(rv).ctypes.data,
                             f(wc).ctypes.data) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    t2, r2, v2, w2 = fileio.read_cloud_file(theirs)
    assert t2 == ty.tolist()
    np.testing.assert_array_equal(np.stack(r2), re_)
    np.testing.assert_array_equal(np.stack(v2), rv)
    np.testing.assert_array_equal(np.stack(w2), wc)
    nco, ko = ctypes.c_int(), ctypes.c_int()
    rt = np.zeros(nc, np.int32)
    rr, rvv, rw = np.zeros((nc, K)), np.zeros((nc, K)), np.zeros((nc, K))
    q = ours.encode()
    assert L.ref_cloud_read(q, len(q), nc, K, ctypes.byref(nco), ctypes.byref(ko), rt.ctypes.data, rr.ctypes.data,
                            rvv.ctypes.data, rw.ctypes.data) == 0
    assert (nco.value, ko.value) == (nc, K)
    np.testing.assert_array_equal(rt, ty)
    np.testing.assert_array_equal(rr, re_)
    np.testing.assert_array_equal(rw, wc)


def _cloudy_batch():
    ps = synthetic.make_profiles(4, [40, 55, 31, 62], seed=8)
    ps = synthetic.add_clouds(ps, seed=9)
    ps = synthetic.add_aerosols(ps, seed=10)
    ps.climatology = np.array([1, 6, 3, 2], np.int32)
    return ps


def test_atmosphere_round_trip(tmp_path):
    ps = _cloudy_batch()
    rh = np.random.default_rng(1).uniform(0, 1, int(ps.layer_offset[-1]))
    path = str(tmp_path / "atm.bin")
    fileio.write_atmosphere(path, ps, relative_humidity=rh)
    back, rh2 = fileio.read_atmosphere(path)
    np.testing.assert_array_equal(rh2, rh)
    for k in ("n_layers", "layer_offset", "absorber_id", "absorber_units", "level_pressure", "pressure",
              "temperature", "absorber", "climatology", "cloud_fraction", "aerosol_type"):
        np.testing.assert_array_equal(getattr(back, k), getattr(ps, k), err_msg=k)
    # cloud slots that are empty in a profile are dropped on write and the
    # remaining ones packed to the front on read
    for m in range(ps.n_profiles):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        keep = [n for n in range(ps.cloud_type.shape[1]) if ps.cloud_type[m, n] != 0]
        got = [n for n in range(back.cloud_type.shape[1]) if back.cloud_type[m, n] != 0]
        assert len(got) == len(keep)
        for i, n in zip(got, keep):
            assert back.cloud_type[m, i] == ps.cloud_type[m, n]
            np.testing.assert_array_equal(back.cloud_water_content[i, a:b], ps.cloud_water_content[n, a:b])
            np.testing.assert_array_equal(back.cloud_effective_radius[i, a:b], ps.cloud_effective_radius[n, a:b])
        np.testing.assert_array_equal(back.aerosol_concentration[:, a:b], ps.aerosol_concentration[:, a:b])


def test_atmosphere_cloud_block_is_a_reference_cloud_file(tmp_path):
    """One profile with clouds and no aerosols: the records after the three
    profile records are the reference's Cloud file without its magic number."""
    L = _ref()
    ps = synthetic.add_clouds(synthetic.make_profiles(1, 23, seed=4), seed=6)
    path = str(tmp_path / "atm.bin")
    fileio.write_atmosphere(path, ps)
    recs = _records(path)
    K = int(ps.n_layers[0])
    keep = [n for n in range(ps.cloud_type.shape[1]) if ps.cloud_type[0, n] != 0]
    assert keep, "the sample profile has no cloud"
    assert struct.unpack("<4i", recs[2]) == (K, ps.absorber.shape[0], len(keep), 0)
    ty = np.ascontiguousarray(ps.cloud_type[0, keep], np.int32)
    f = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    theirs = str(tmp_path / "ref.bin")
    p = theirs.encode()
    assert L.ref_cloud_write(p, len(p), len(keep), K, ty.ctypes.data, f(ps.cloud_effective_radius[keep]).ctypes.data,
                             f(np.zeros((len(keep), K))).ctypes.data, f(ps.cloud_water_content[keep]).ctypes.data) == 0
    assert recs[5:] == _records(theirs)[1:]


def test_rtsolution_round_trip_and_record_sizes(tmp_path):
    rng = np.random.default_rng(2)
    M, L, K = 3, 5, 7
    rts = {k: rng.normal(size=(M, L)) for k in fileio.RTS_SCALARS + fileio.RTS_RESULTS}
    for k in fileio.RTS_LAYERED:
        rts[k] = rng.normal(size=(M, L, K))
    rts["stokes"] = rng.normal(size=(M, L, 4))
    ci = [("cris399_npp", 224, 620, 30 + l) for l in range(L)]
    path = str(tmp_path / "rts.bin")
    fileio.write_rtsolution(path, rts, ci)
    back = fileio.read_rtsolution(path)
    for k in fileio.RTS_SCALARS + fileio.RTS_RESULTS + fileio.RTS_LAYERED + ("stokes",):
        np.testing.assert_array_equal(back[k], rts[k], err_msg=k)
    assert back["channel_info"] == ci and back["algorithm"] == "ADA"
    recs = _records(path)
    # magic, [n_Channels, n_Profiles], then 6 records per element:
    # n_Layers; Sensor_Id(20) + 3 ints; algorithm name(20); 11 reals;
    # 5 layer arrays; 4 reals + Stokes(4)
    assert struct.unpack("<2i", recs[1]) == (L, M)
    assert len(recs) == 2 + 6 * M * L
    assert [len(r) for r in recs[2:8]] == [4, 20 + 12, 20, 11 * 8, 5 * K * 8, 8 * 8]
    # element (l=2, m=1) sits at position m*L + l: its radiance is the first real of its last record
    e = 2 + 6 * (1 * L + 2)
    assert struct.unpack_from("<d", recs[e + 5])[0] == rts["radiance"][1, 2]
    # a forward-call dict (no layers) writes n_Layers = 0 and no layer record
    fileio.write_rtsolution(path, {"brightness_temperature": rts["brightness_temperature"],
                                   "radiance": rts["radiance"]})
    recs = _records(path)
    assert len(recs) == 2 + 5 * M * L
    back = fileio.read_rtsolution(path)
    np.testing.assert_array_equal(back["brightness_temperature"], rts["brightness_temperature"])
    assert "upwelling_radiance" not in back or back["upwelling_radiance"].size == 0
