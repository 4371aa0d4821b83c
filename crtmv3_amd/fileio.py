"""CRTM binary files of Atmosphere, Cloud, Aerosol, Geometry and RTSolution
structures (test fixtures and batch inputs/outputs of the drop-in).

Every file is Fortran sequential-unformatted (4-byte record markers),
little-endian, opened by the magic-number record of
src/Utility/Binary_File_Utility.f90 (`Open_Binary_File`).  Record lists:

* Atmosphere (src/Atmosphere/CRTM_Atmosphere_Define.f90:1674-1790 file,
  :2474-2712 record): [n_Channels = 0, n_Profiles]; per profile
  [n_Layers, n_Absorbers, n_Clouds, n_Aerosols],
  [Climatology, Absorber_ID(n_Absorbers), Absorber_Units(n_Absorbers)],
  [Level_Pressure(0:K), Pressure(K), Temperature(K), Relative_Humidity(K),
  Absorber(K, n_Absorbers), Cloud_Fraction(K)], then a Cloud block and an
  Aerosol block in the same unit when the profile has any.
* Cloud block (src/Atmosphere/Cloud/CRTM_Cloud_Define.f90:1087-1214 file,
  :1640-1750 record): [n_Clouds]; per cloud [n_Layers],
  [Type, Effective_Radius(K), Effective_Variance(K), Water_Content(K)].
* Aerosol block (src/Atmosphere/Aerosol/CRTM_Aerosol_Define.f90, the
  'CRTM' aerosol scheme): [n_Aerosols]; per aerosol [n_Layers],
  [Type, Effective_Radius(K), Concentration(K)].
* Geometry (src/GeometryInfo/CRTM_Geometry_Define.f90:1061-1194 file,
  :1407-1541 record): [n_Profiles]; per profile [iFOV, Longitude, Latitude,
  Surface_Altitude, Sensor_Scan_Angle, Sensor_Zenith_Angle,
  Sensor_Azimuth_Angle, Source_Zenith_Angle, Source_Azimuth_Angle,
  Flux_Zenith_Angle, Year, Month, Day].
* RTSolution (src/RTSolution/CRTM_RTSolution_Define.f90:1416-1521 file,
  2272-2356 file write, Write_Record): [n_Channels, n_Profiles]; per
  profile and channel (channel fastest) [n_Layers], [Sensor_Id (20 chars),
  WMO_Satellite_Id, WMO_Sensor_Id, Sensor_Channel], [RT_Algorithm_Name
  (20 chars)], [SOD, Surface_Emissivity, Surface_Reflectivity, Up_Radiance,
  Down_Radiance, Down_Solar_Radiance, Surface_Planck_Radiance,
  Total_Cloud_Cover, R_clear, Tb_clear, Reflectance_clear], with n_Layers
  > 0 [Upwelling_Overcast_Radiance(K), Upwelling_Radiance(K),
  Layer_Optical_Depth(K), Reflectivity(K), Reflectivity_Attenuated(K)],
  then [Radiance, Brightness_Temperature, Solar_Irradiance, Reflectance,
  Stokes(4)].

Batches are the package's flat layout (synthetic.ProfileSet: per-layer
arrays concatenated over profiles, layer_offset[m] the first layer of
profile m); an RTSolution batch is a dict of [M, L] arrays.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Optional

import numpy as np

from .coeffio import RecordReader, RecordWriter

RTS_SCALARS = ("sod", "surface_emissivity", "surface_reflectivity", "up_radiance", "down_radiance",
               "down_solar_radiance", "surface_planck_radiance", "total_cloud_cover", "r_clear", "tb_clear",
               "reflectance_clear")
RTS_LAYERED = ("upwelling_overcast_radiance", "upwelling_radiance", "layer_optical_depth", "reflectivity",
               "reflectivity_attenuated")
RTS_RESULTS = ("radiance", "brightness_temperature", "solar_irradiance", "reflectance")
GEO_REALS = ("longitude", "latitude", "surface_altitude", "sensor_scan", "sensor_zenith", "sensor_azimuth",
             "source_zenith", "source_azimuth", "flux_zenith")
GEO_INTS = ("ifov", "year", "month", "day")
# src/GeometryInfo/CRTM_Geometry_Define.f90:99-121 defaults
GEO_DEFAULTS = dict(ifov=0, longitude=0.0, latitude=0.0, surface_altitude=0.0, sensor_scan=0.0,
                    sensor_zenith=0.0, sensor_azimuth=999.9, source_zenith=100.0, source_azimuth=0.0,
                    flux_zenith=53.130102, year=2001, month=1, day=1)


def _i(x):
    return np.int32(x)


def _f(x):
    return np.asarray(x, np.float64)


# --------------------------------------------------------------------------
# Cloud and aerosol blocks
# --------------------------------------------------------------------------
def _write_cloud_block(w: RecordWriter, types, reff, rvar, wc) -> None:
    """types [nc]; reff, rvar, wc [nc, K]."""
    w.record(_i(len(types)))
    for n in range(len(types)):
        K = reff.shape[1]
        w.record(_i(K))
        w.record(_i(types[n]), _f(reff[n]), _f(rvar[n]), _f(wc[n]))


def _read_cloud_block(r: RecordReader):
    nc = int(r.read_ints(1)[0])
    types, reff, rvar, wc = [], [], [], []
    for _ in range(nc):
        K = int(r.read_ints(1)[0])
        t, a, b, c = r.read([("i4", ()), ("f8", (K,)), ("f8", (K,)), ("f8", (K,))])
        types.append(int(t))
        reff.append(a)
        rvar.append(b)
        wc.append(c)
    return types, reff, rvar, wc


def _write_aerosol_block(w: RecordWriter, types, reff, conc) -> None:
    w.record(_i(len(types)))
    for n in range(len(types)):
        w.record(_i(reff.shape[1]))
        w.record(_i(types[n]), _f(reff[n]), _f(conc[n]))


def _read_aerosol_block(r: RecordReader):
    na = int(r.read_ints(1)[0])
    types, reff, conc = [], [], []
    for _ in range(na):
        K = int(r.read_ints(1)[0])
        t, a, c = r.read([("i4", ()), ("f8", (K,)), ("f8", (K,))])
        types.append(int(t))
        reff.append(a)
        conc.append(c)
    return types, reff, conc


def write_cloud_file(path: str, types, reff, rvar, wc) -> None:
    """A Cloud file (CRTM_Cloud_WriteFile): the cloud block behind the magic number."""
    with RecordWriter(path) as w:
        _write_cloud_block(w, types, np.atleast_2d(reff), np.atleast_2d(rvar), np.atleast_2d(wc))


def read_cloud_file(path: str):
    """(types, effective radius, effective variance, water content), one entry per cloud."""
    return _read_cloud_block(RecordReader(path))


# --------------------------------------------------------------------------
# Atmosphere
# --------------------------------------------------------------------------
def write_atmosphere(path: str, ps, relative_humidity: Optional[np.ndarray] = None) -> None:
    """Atmosphere file of a ProfileSet (CRTM_Atmosphere_WriteFile, rank-1).

    Cloud slots of type 0 (no cloud) are dropped, as a profile's Cloud array
    holds only its n_Clouds clouds; effective variances are written as zero
    (the ADA path does not read them).  relative_humidity [sum K] defaults
    to zero, the reference's initial value."""
    M = ps.n_profiles
    NT = int(ps.layer_offset[-1])
    rh = np.zeros(NT) if relative_humidity is None else np.asarray(relative_humidity, np.float64)
    clim = np.full(M, 6, np.int32) if ps.climatology is None else np.asarray(ps.climatology)
    with RecordWriter(path) as w:
        w.record(_i(0), _i(M))
        for m in range(M):
            a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
            K = b - a
            clouds = [] if ps.cloud_type is None else [n for n in range(ps.cloud_type.shape[1])
                                                        if ps.cloud_type[m, n] != 0]
            aers = [] if ps.aerosol_type is None else [n for n in range(ps.aerosol_type.shape[1])
                                                        if ps.aerosol_type[m, n] != 0]
            na = ps.absorber.shape[0]
            w.record(_i(K), _i(na), _i(len(clouds)), _i(len(aers)))
            w.record(_i(clim[m]), np.asarray(ps.absorber_id, np.int32), np.asarray(ps.absorber_units, np.int32))
            cf = np.zeros(K) if ps.cloud_fraction is None else ps.cloud_fraction[a:b]
            w.record(_f(ps.level_pressure[a + m:b + m + 1]), _f(ps.pressure[a:b]), _f(ps.temperature[a:b]),
                     _f(rh[a:b]), _f(ps.absorber[:, a:b].T), _f(cf))
            if clouds:
                _write_cloud_block(w, [ps.cloud_type[m, n] for n in clouds],
                                   ps.cloud_effective_radius[clouds, a:b], np.zeros((len(clouds), K)),
                                   ps.cloud_water_content[clouds, a:b])
            if aers:
                _write_aerosol_block(w, [ps.aerosol_type[m, n] for n in aers],
                                     ps.aerosol_effective_radius[aers, a:b], ps.aerosol_concentration[aers, a:b])


def read_atmosphere(path: str):
    """ProfileSet (atmosphere part; surface and geometry at their defaults)
    and the relative humidity [sum K] of an Atmosphere file.  Every profile
    must carry the same absorbers (the batch layout holds one absorber list);
    cloud and aerosol slots are padded to the largest count of any profile."""
    from .synthetic import ProfileSet
    r = RecordReader(path)
    n_ch, M = (int(v) for v in r.read_ints(2))
    if n_ch != 0:
        raise ValueError(f"{path}: n_Channels = {n_ch}; a rank-1 Atmosphere file has 0")
    profs = []
    for m in range(M):
        K, na, nc, nae = (int(v) for v in r.read_ints(4))
        clim, ids, units = r.read([("i4", ()), ("i4", (na,)), ("i4", (na,))])
        lp, p, t, rh, ab, cf = r.read([("f8", (K + 1,)), ("f8", (K,)), ("f8", (K,)), ("f8", (K,)),
                                       ("f8", (K, na)), ("f8", (K,))])
        cl = _read_cloud_block(r) if nc > 0 else ([], [], [], [])
        ae = _read_aerosol_block(r) if nae > 0 else ([], [], [])
        profs.append(dict(K=K, clim=int(clim), ids=np.array(ids), units=np.array(units), lp=lp, p=p, t=t, rh=rh,
                          ab=ab, cf=cf, cl=cl, ae=ae))
    if M == 0:
        raise ValueError(f"{path}: no profiles")
    ids, units = profs[0]["ids"], profs[0]["units"]
    for q in profs[1:]:
        if not (np.array_equal(q["ids"], ids) and np.array_equal(q["units"], units)):
            raise ValueError(f"{path}: profiles with different absorber lists")
    nl = np.array([q["K"] for q in profs], np.int32)
    off = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
    NT = int(off[-1])
    cat = lambda k: np.concatenate([q[k] for q in profs])  # noqa: E731
    z = np.zeros(M)
    ps = ProfileSet(n_layers=nl, layer_offset=off, absorber_id=ids.astype(np.int32),
                    absorber_units=units.astype(np.int32), level_pressure=cat("lp"), pressure=cat("p"),
                    temperature=cat("t"), absorber=np.concatenate([q["ab"].T for q in profs], axis=1),
                    land_coverage=z.copy(), water_coverage=np.ones(M), snow_coverage=z.copy(),
                    ice_coverage=z.copy(), land_temperature=np.full(M, 283.0), water_temperature=np.full(M, 283.0),
                    snow_temperature=np.full(M, 263.0), ice_temperature=np.full(M, 263.0), sensor_zenith=z.copy(),
                    sensor_scan=z.copy(), source_zenith=np.full(M, 100.0),
                    climatology=np.array([q["clim"] for q in profs], np.int32))
    nc = max(len(q["cl"][0]) for q in profs)
    if nc:
        ps.cloud_type = np.zeros((M, nc), np.int32)
        ps.cloud_water_content = np.zeros((nc, NT))
        ps.cloud_effective_radius = np.zeros((nc, NT))
        for m, q in enumerate(profs):
            a, b = int(off[m]), int(off[m + 1])
            for n, (ty, re_, _, wc) in enumerate(zip(*q["cl"])):
                ps.cloud_type[m, n] = ty
                ps.cloud_effective_radius[n, a:b] = re_
                ps.cloud_water_content[n, a:b] = wc
    if nc or any(q["cf"].any() for q in profs):
        ps.cloud_fraction = cat("cf")
    nae = max(len(q["ae"][0]) for q in profs)
    if nae:
        ps.aerosol_type = np.zeros((M, nae), np.int32)
        ps.aerosol_effective_radius = np.zeros((nae, NT))
        ps.aerosol_concentration = np.zeros((nae, NT))
        for m, q in enumerate(profs):
            a, b = int(off[m]), int(off[m + 1])
            for n, (ty, re_, cc) in enumerate(zip(*q["ae"])):
                ps.aerosol_type[m, n] = ty
                ps.aerosol_effective_radius[n, a:b] = re_
                ps.aerosol_concentration[n, a:b] = cc
    return ps, cat("rh")


# --------------------------------------------------------------------------
# Geometry
# --------------------------------------------------------------------------
def write_geometry(path: str, geo: Dict[str, np.ndarray]) -> None:
    """Geometry file (CRTM_Geometry_WriteFile) of per-profile arrays keyed by
    GEO_INTS + GEO_REALS; missing keys take the reference defaults."""
    M = max(len(np.atleast_1d(v)) for v in geo.values())
    col = {k: np.broadcast_to(np.asarray(geo.get(k, GEO_DEFAULTS[k])), (M,)) for k in GEO_INTS + GEO_REALS}
    with RecordWriter(path) as w:
        w.record(_i(M))
        for m in range(M):
            w.record(_i(col["ifov"][m]), *[_f(col[k][m]) for k in GEO_REALS],
                     _i(col["year"][m]), _i(col["month"][m]), _i(col["day"][m]))


def read_geometry(path: str) -> Dict[str, np.ndarray]:
    r = RecordReader(path)
    M = int(r.read_ints(1)[0])
    rows = [r.read([("i4", ())] + [("f8", ())] * len(GEO_REALS) + [("i4", ())] * 3) for _ in range(M)]
    out = {"ifov": np.array([q[0] for q in rows], np.int32)}
    for i, k in enumerate(GEO_REALS):
        out[k] = np.array([q[1 + i] for q in rows], np.float64)
    for i, k in enumerate(("year", "month", "day")):
        out[k] = np.array([q[1 + len(GEO_REALS) + i] for q in rows], np.int32)
    return out


def geometry_of(ps) -> Dict[str, np.ndarray]:
    """The Geometry fields a ProfileSet carries."""
    g = {"sensor_zenith": ps.sensor_zenith, "sensor_scan": ps.sensor_scan, "source_zenith": ps.source_zenith}
    if ps.sensor_azimuth is not None:
        g["sensor_azimuth"] = ps.sensor_azimuth
    if ps.ifov is not None:
        g["ifov"] = ps.ifov
    return g


def apply_geometry(ps, geo: Dict[str, np.ndarray]):
    """ProfileSet with the geometry fields of a Geometry file."""
    return dataclasses.replace(ps, sensor_zenith=geo["sensor_zenith"], sensor_scan=geo["sensor_scan"],
                               source_zenith=geo["source_zenith"], sensor_azimuth=geo["sensor_azimuth"],
                               ifov=geo["ifov"])


# --------------------------------------------------------------------------
# RTSolution
# --------------------------------------------------------------------------
def write_rtsolution(path: str, rts: Dict[str, np.ndarray], channel_info=None,
                     algorithm: str = "ADA") -> None:
    """RTSolution file (CRTM_RTSolution_WriteFile, rank-2 [L, M]) of an
    RTSolution dict of [M, L] arrays (as Context.forward returns).  Fields the
    dict lacks are written as zero and n_Layers as 0 unless the dict holds
    the per-layer fields as [M, L, K] arrays.  channel_info: per channel
    (sensor_id, wmo_satellite_id, wmo_sensor_id, sensor_channel)."""
    bt = np.asarray(rts["brightness_temperature"])
    M, L = bt.shape
    get = lambda k: np.asarray(rts[k], np.float64) if k in rts else np.zeros((M, L))  # noqa: E731
    sc = {k: get(k) for k in RTS_SCALARS + RTS_RESULTS}
    lay = {k: np.asarray(rts[k], np.float64) for k in RTS_LAYERED if k in rts}
    K = next(iter(lay.values())).shape[2] if lay else 0
    stokes = np.asarray(rts["stokes"], np.float64) if "stokes" in rts else np.zeros((M, L, 4))
    ci = channel_info or [("", -1, -1, l + 1) for l in range(L)]
    with RecordWriter(path) as w:
        w.record(_i(L), _i(M))
        for m in range(M):
            for l in range(L):
                sid, wsat, wsen, ch = ci[l]
                w.record(_i(K))
                w.record(str(sid), _i(wsat), _i(wsen), _i(ch))
                w.record(algorithm)
                w.record(*[_f(sc[k][m, l]) for k in RTS_SCALARS])
                if K > 0:
                    w.record(*[_f(lay[k][m, l]) if k in lay else np.zeros(K) for k in RTS_LAYERED])
                w.record(*[_f(sc[k][m, l]) for k in RTS_RESULTS], _f(stokes[m, l]))


def read_rtsolution(path: str) -> Dict[str, np.ndarray]:
    """RTSolution dict of [M, L] arrays (+ [M, L, K] per-layer fields when
    the file has layers, 'stokes' [M, L, 4], 'channel_info' per channel and
    'algorithm')."""
    r = RecordReader(path)
    L, M = (int(v) for v in r.read_ints(2))
    out = {k: np.zeros((M, L)) for k in RTS_SCALARS + RTS_RESULTS}
    out["stokes"] = np.zeros((M, L, 4))
    ci, alg, K0 = [None] * L, "", None
    for m in range(M):
        for l in range(L):
            K = int(r.read_ints(1)[0])
            if K0 is None:
                K0 = K
                for k in RTS_LAYERED:
                    if K:
                        out[k] = np.zeros((M, L, K))
            elif K != K0:
                raise ValueError(f"{path}: channels with different layer counts ({K} and {K0})")
            ci[l] = tuple(v if isinstance(v, str) else int(v)
                          for v in r.read([("s", ()), ("i4", ()), ("i4", ()), ("i4", ())]))
            alg = r.read([("s", ())])[0]
            for k, v in zip(RTS_SCALARS, r.read([("f8", ())] * len(RTS_SCALARS))):
                out[k][m, l] = v
            if K:
                for k, v in zip(RTS_LAYERED, r.read([("f8", (K,))] * len(RTS_LAYERED))):
                    out[k][m, l] = v
            vals = r.read([("f8", ())] * len(RTS_RESULTS) + [("f8", (4,))])
            for k, v in zip(RTS_RESULTS, vals[:-1]):
                out[k][m, l] = v
            out["stokes"][m, l] = vals[-1]
    out["channel_info"] = ci
    out["algorithm"] = alg
    return out
