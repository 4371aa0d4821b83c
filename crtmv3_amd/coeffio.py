"""CRTM binary coefficient files: Fortran sequential-unformatted records.

Writers and readers for the on-disk formats the native loader
(crtmv3_amd/csrc/coeff_load.cpp) consumes, so synthetic coefficient sets can
be produced offline in exactly the byte layout the reference reads.

Record framing follows gfortran/flang sequential unformatted I/O: every
record is ``int32 nbytes | payload | int32 nbytes`` little-endian.  The first
record of every file is the magic number 123456789
(reference: src/Utility/Binary_File_Utility.f90:56, 241).

Formats:
  * SpcCoeff, release 8 / version 3
    (src/Coefficients/SpcCoeff/SpcCoeff_Binary_IO.f90:374-470,
     type at src/Coefficients/SpcCoeff/SpcCoeff_Define.f90:163-189)
  * ODPS TauCoeff, release 2
    (src/Coefficients/TauCoeff/ODPS/ODPS_Binary_IO.f90:617-800,
     type at src/Coefficients/TauCoeff/ODPS/ODPS_Define.f90:109-205)
  * CloudCoeff, release 3
    (src/Coefficients/CloudCoeff/CloudCoeff_Binary_IO.f90:340-460,
     type at src/Coefficients/CloudCoeff/CloudCoeff_Define.f90:95-154)
"""
from __future__ import annotations

import dataclasses
import struct
from typing import List, Optional

import numpy as np

MAGIC_NUMBER = 123456789
SENSOR_ID_LEN = 20           # CHARACTER(SL), SL = 20
SPCCOEFF_RELEASE = 8
SPCCOEFF_VERSION = 3
ODPS_RELEASE = 2
ODPS_VERSION = 1
ODPS_ALGORITHM = 2           # src/CRTM_Parameters.f90:128
N_PREDICTOR_USED_OPTRAN = 6  # ODPS_Define.f90:103

CLOUDCOEFF_RELEASE = 3
CLOUDCOEFF_VERSION = 1

MICROWAVE_SENSOR = 1
INFRARED_SENSOR = 2
VISIBLE_SENSOR = 3
ULTRAVIOLET_SENSOR = 4


class RecordWriter:
    def __init__(self, path: str):
        self._f = open(path, "wb")
        self.record(np.int32(MAGIC_NUMBER))

    def record(self, *items) -> None:
        payload = b"".join(_to_bytes(x) for x in items)
        n = struct.pack("<i", len(payload))
        self._f.write(n + payload + n)

    def close(self) -> None:
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _to_bytes(x) -> bytes:
    if isinstance(x, bytes):
        return x
    if isinstance(x, str):
        return x.encode("ascii").ljust(SENSOR_ID_LEN)[:SENSOR_ID_LEN]
    a = np.asarray(x)
    if a.dtype.kind == "i":
        a = a.astype("<i4")
    elif a.dtype == np.float32:
        a = a.astype("<f4")
    elif a.dtype.kind == "f":
        a = a.astype("<f8")
    # Fortran arrays are column-major
    return np.asfortranarray(a).tobytes(order="F")


class RecordReader:
    def __init__(self, path: str):
        with open(path, "rb") as f:
            self._buf = f.read()
        self._pos = 0
        magic = self.read_ints(1)[0]
        if magic != MAGIC_NUMBER:
            raise ValueError(f"{path}: bad magic number {magic}")

    def _next(self) -> bytes:
        (n,) = struct.unpack_from("<i", self._buf, self._pos)
        p = self._pos + 4
        payload = self._buf[p:p + n]
        (n2,) = struct.unpack_from("<i", self._buf, p + n)
        if n2 != n:
            raise ValueError("corrupt record framing")
        self._pos = p + n + 4
        return payload

    def read(self, fmt: List[tuple]):
        """fmt: list of (kind, shape) with kind in 'i4','f4','f8','s'."""
        payload = self._next()
        out, off = [], 0
        for kind, shape in fmt:
            if kind == "s":
                out.append(payload[off:off + SENSOR_ID_LEN].decode("ascii").rstrip())
                off += SENSOR_ID_LEN
                continue
            dt = np.dtype("<" + kind)
            cnt = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(payload, dt, cnt, off)
            off += cnt * dt.itemsize
            out.append(a.reshape(shape, order="F") if shape else a[0])
        if off != len(payload):
            raise ValueError(f"record length mismatch {off} != {len(payload)}")
        return out

    def remaining(self) -> bool:
        return self._pos < len(self._buf)

    def read_ints(self, n):
        payload = self._next()
        return np.frombuffer(payload, "<i4", n)


# --------------------------------------------------------------------------
# SpcCoeff
# --------------------------------------------------------------------------
@dataclasses.dataclass
class SpcCoeff:
    sensor_id: str
    sensor_type: int
    wmo_satellite_id: int
    wmo_sensor_id: int
    sensor_channel: np.ndarray          # int32 [L]
    polarization: np.ndarray            # int32 [L]
    pol_angle: np.ndarray               # f8 [L]
    channel_flag: np.ndarray            # int32 [L]
    frequency: np.ndarray               # GHz [L]
    wavenumber: np.ndarray              # cm^-1 [L]
    planck_c1: np.ndarray
    planck_c2: np.ndarray
    band_c1: np.ndarray
    band_c2: np.ndarray
    cosmic_background_radiance: np.ndarray
    solar_irradiance: np.ndarray
    ac: Optional["ACCoeff"] = None      # antenna correction block (microwave sensors)
    nc: Optional["NLTECoeff"] = None    # non-LTE correction block (infrared sensors)

    @property
    def n_channels(self) -> int:
        return int(self.sensor_channel.size)


ACCOEFF_RELEASE = 1              # src/Coefficients/ACCoeff/ACCoeff_Define.f90:71-72
ACCOEFF_VERSION = 1
NLTECOEFF_RELEASE = 2            # src/Coefficients/NLTECoeff/NLTECoeff_Define.f90
NLTECOEFF_VERSION = 1
N_NLTE_LAYERS = 2                # src/NLTE/NLTE_Parameters.f90


@dataclasses.dataclass
class ACCoeff:
    """Antenna correction (ACCoeff_type, src/Coefficients/ACCoeff/ACCoeff_Define.f90:78-96):
    antenna temperature = (A_earth + A_platform) * BT + A_space * TSPACE per field of view."""
    sensor_id: str
    wmo_satellite_id: int
    wmo_sensor_id: int
    sensor_channel: np.ndarray          # int32 [L]
    a_earth: np.ndarray                 # f8 [n_fovs, L]
    a_space: np.ndarray
    a_platform: np.ndarray

    @property
    def n_fovs(self) -> int:
        return int(self.a_earth.shape[0])


@dataclasses.dataclass
class NLTECoeff:
    """Non-LTE correction (NLTECoeff_type, src/Coefficients/NLTECoeff/NLTECoeff_Define.f90:87-121)."""
    sensor_id: str
    wmo_satellite_id: int
    wmo_sensor_id: int
    sensor_channel: np.ndarray          # int32 [L]
    upper_plevel: np.ndarray            # f8 [2]
    lower_plevel: np.ndarray            # f8 [2]
    min_tm: np.ndarray                  # f8 [2]
    max_tm: np.ndarray
    mean_tm: np.ndarray
    nlte_channel: np.ndarray            # int32 [n_nlte]  sensor channel numbers
    secant_sensor_zenith: np.ndarray    # f8 [n_sen]
    secant_solar_zenith: np.ndarray     # f8 [n_sol]
    c_index: np.ndarray                 # int32 [L]: 1-based index into C's last axis, 0 = not NLTE
    c: np.ndarray                       # f8 [n_pred, n_sen, n_sol, n_nlte]


def _write_ac_block(w, ac: ACCoeff) -> None:
    """Record order of ACCoeff_Binary_ReadFile (src/Coefficients/ACCoeff/ACCoeff_Binary_IO.f90)."""
    n, L = ac.a_earth.shape
    w.record(np.int32(ACCOEFF_RELEASE), np.int32(ACCOEFF_VERSION))
    w.record(np.int32(n), np.int32(L))
    w.record(ac.sensor_id, np.int32(ac.wmo_satellite_id), np.int32(ac.wmo_sensor_id))
    w.record(np.asarray(ac.sensor_channel, np.int32))
    w.record(np.asarray(ac.a_earth, np.float64), np.asarray(ac.a_space, np.float64),
             np.asarray(ac.a_platform, np.float64))


def _read_ac_block(r) -> ACCoeff:
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != ACCOEFF_RELEASE:
        raise ValueError(f"ACCoeff release {rel} unsupported")
    n, L = r.read([("i4", ()), ("i4", ())])
    sid, wsat, wsen = r.read([("s", None), ("i4", ()), ("i4", ())])
    (ch,) = r.read([("i4", (L,))])
    ae, asp, ap = r.read([("f8", (n, L))] * 3)
    return ACCoeff(sid, int(wsat), int(wsen), ch.copy(), ae.copy(), asp.copy(), ap.copy())


def _write_nc_block(w, nc: NLTECoeff) -> None:
    """Record order of NLTECoeff_Binary_ReadFile (src/Coefficients/NLTECoeff/NLTECoeff_Binary_IO.f90)."""
    n1, n2, n3, n4 = nc.c.shape
    L = int(nc.sensor_channel.size)
    w.record(np.int32(NLTECOEFF_RELEASE), np.int32(NLTECOEFF_VERSION))
    w.record(np.int32(n1), np.int32(n2), np.int32(n3), np.int32(n4), np.int32(L))
    w.record(nc.sensor_id, np.int32(nc.wmo_satellite_id), np.int32(nc.wmo_sensor_id),
             np.asarray(nc.sensor_channel, np.int32))
    w.record(np.asarray(nc.upper_plevel, np.float64), np.asarray(nc.lower_plevel, np.float64))
    w.record(np.asarray(nc.min_tm, np.float64), np.asarray(nc.max_tm, np.float64),
             np.asarray(nc.mean_tm, np.float64))
    w.record(np.asarray(nc.nlte_channel, np.int32))
    w.record(np.asarray(nc.secant_sensor_zenith, np.float64), np.asarray(nc.secant_solar_zenith, np.float64))
    w.record(np.asarray(nc.c_index, np.int32), np.asarray(nc.c, np.float64))


def _read_nc_block(r) -> NLTECoeff:
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != NLTECOEFF_RELEASE:
        raise ValueError(f"NLTECoeff release {rel} unsupported")
    n1, n2, n3, n4, L = r.read([("i4", ())] * 5)
    sid, wsat, wsen, ch = r.read([("s", None), ("i4", ()), ("i4", ()), ("i4", (L,))])
    up, lo = r.read([("f8", (N_NLTE_LAYERS,))] * 2)
    mn, mx, me = r.read([("f8", (N_NLTE_LAYERS,))] * 3)
    (nch,) = r.read([("i4", (n4,))])
    ss, so = r.read([("f8", (n2,)), ("f8", (n3,))])
    ci, c = r.read([("i4", (L,)), ("f8", (n1, n2, n3, n4))])
    return NLTECoeff(sid, int(wsat), int(wsen), ch.copy(), up.copy(), lo.copy(), mn.copy(), mx.copy(),
                     me.copy(), nch.copy(), ss.copy(), so.copy(), ci.copy(), c.copy())


def write_accoeff(path: str, ac: ACCoeff) -> None:
    """A standalone ACCoeff file (magic number, then the block)."""
    with RecordWriter(path) as w:
        _write_ac_block(w, ac)


def write_nltecoeff(path: str, nc: NLTECoeff) -> None:
    """A standalone NLTECoeff file (magic number, then the block)."""
    with RecordWriter(path) as w:
        _write_nc_block(w, nc)


def write_spccoeff(path: str, sc: SpcCoeff) -> None:
    """Record order of SpcCoeff_Binary_IO.f90:374-470 (Version > 2)."""
    with RecordWriter(path) as w:
        w.record(np.int32(SPCCOEFF_RELEASE), np.int32(SPCCOEFF_VERSION))
        w.record(np.int32(sc.n_channels))
        w.record(sc.sensor_id, np.int32(sc.sensor_type),
                 np.int32(sc.wmo_satellite_id), np.int32(sc.wmo_sensor_id))
        w.record(np.asarray(sc.sensor_channel, np.int32),
                 np.asarray(sc.polarization, np.int32),
                 np.asarray(sc.pol_angle, np.float64),
                 np.asarray(sc.channel_flag, np.int32),
                 np.asarray(sc.frequency, np.float64),
                 np.asarray(sc.wavenumber, np.float64),
                 np.asarray(sc.planck_c1, np.float64),
                 np.asarray(sc.planck_c2, np.float64),
                 np.asarray(sc.band_c1, np.float64),
                 np.asarray(sc.band_c2, np.float64),
                 np.asarray(sc.cosmic_background_radiance, np.float64),
                 np.asarray(sc.solar_irradiance, np.float64))
        # antenna- and NLTE-correction blocks, each behind a DATA_PRESENT flag
        # (SpcCoeff_Binary_IO.f90:466-527)
        w.record(np.int32(1 if sc.ac is not None else 0))
        if sc.ac is not None:
            _write_ac_block(w, sc.ac)
        w.record(np.int32(1 if sc.nc is not None else 0))
        if sc.nc is not None:
            _write_nc_block(w, sc.nc)


def read_spccoeff(path: str) -> SpcCoeff:
    r = RecordReader(path)
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != SPCCOEFF_RELEASE:
        raise ValueError(f"SpcCoeff release {rel} unsupported")
    (n,) = r.read([("i4", ())])
    sid, stype, wsat, wsen = r.read([("s", None), ("i4", ()), ("i4", ()), ("i4", ())])
    if ver > 2:
        a = r.read([("i4", (n,)), ("i4", (n,)), ("f8", (n,)), ("i4", (n,))] + [("f8", (n,))] * 8)
        ch, pol, pang, flag = a[:4]
        rest = a[4:]
    else:
        a = r.read([("i4", (n,)), ("i4", (n,)), ("i4", (n,))] + [("f8", (n,))] * 8)
        ch, pol, flag = a[:3]
        pang = np.zeros(n)
        rest = a[3:]
    ac = nc = None
    if r.remaining():
        (acp,) = r.read([("i4", ())])
        if acp == 1:
            ac = _read_ac_block(r)
        (ncp,) = r.read([("i4", ())])
        if ncp == 1:
            nc = _read_nc_block(r)
    for blk, what in ((ac, "antenna correction"), (nc, "NLTE correction")):
        if blk is not None and (blk.sensor_id != sid or blk.wmo_satellite_id != wsat
                                or blk.wmo_sensor_id != wsen or not np.array_equal(blk.sensor_channel, ch)):
            raise ValueError(f"{path}: {what} sensor information is inconsistent with SpcCoeff")
    return SpcCoeff(sid, int(stype), int(wsat), int(wsen), ch.copy(), pol.copy(), pang.copy(),
                    flag.copy(), *[x.copy() for x in rest], ac=ac, nc=nc)


# --------------------------------------------------------------------------
# ODPS TauCoeff
# --------------------------------------------------------------------------
@dataclasses.dataclass
class ODPS:
    group_index: int
    sensor_id: str
    wmo_satellite_id: int
    wmo_sensor_id: int
    sensor_type: int
    sensor_channel: np.ndarray      # int32 [L]
    component_id: np.ndarray        # int32 [J]
    absorber_id: np.ndarray         # int32 [Jm]
    ref_level_pressure: np.ndarray  # f8 [Ko+1]
    ref_pressure: np.ndarray        # f8 [Ko]
    ref_temperature: np.ndarray     # f8 [Ko]
    ref_absorber: np.ndarray        # f8 [Ko, Jm]
    min_absorber: np.ndarray        # f8 [Ko, Jm]
    max_absorber: np.ndarray        # f8 [Ko, Jm]
    n_predictors: np.ndarray        # int32 [J, L]
    pos_index: np.ndarray           # int32 [J, L]  (1-based start into C)
    c: np.ndarray                   # f4 [n_coeffs]
    # OPTRAN water-vapour-line block (optional)
    osignificance: Optional[np.ndarray] = None   # int32 [L]
    order: Optional[np.ndarray] = None           # int32 [L]
    op_index: Optional[np.ndarray] = None        # int32 [7, L]  (row 0 = count)
    opos_index: Optional[np.ndarray] = None      # int32 [L]     (1-based start into OC)
    oc: Optional[np.ndarray] = None              # f8 [n_ocoeffs]
    alpha: float = 0.0
    alpha_c1: float = 0.0
    alpha_c2: float = 0.0
    ocomponent_index: int = -1

    @property
    def n_layers(self) -> int:
        return int(self.ref_temperature.size)

    @property
    def n_components(self) -> int:
        return int(self.component_id.size)

    @property
    def n_absorbers(self) -> int:
        return int(self.absorber_id.size)

    @property
    def n_channels(self) -> int:
        return int(self.sensor_channel.size)

    @property
    def n_ocoeffs(self) -> int:
        return 0 if self.oc is None else int(self.oc.size)


def write_odps(path: str, t: ODPS) -> None:
    """Record order of Read_ODPS_Data, ODPS_Binary_IO.f90:617-800."""
    L, J, Jm, Ko = t.n_channels, t.n_components, t.n_absorbers, t.n_layers
    with RecordWriter(path) as w:
        w.record(np.int32(ODPS_RELEASE), np.int32(ODPS_VERSION))
        w.record(np.int32(ODPS_ALGORITHM))
        w.record(*[np.int32(x) for x in (Ko, J, Jm, L, t.c.size, N_PREDICTOR_USED_OPTRAN, t.n_ocoeffs)])
        w.record(np.int32(t.group_index))
        w.record(t.sensor_id, np.int32(t.wmo_satellite_id), np.int32(t.wmo_sensor_id),
                 np.int32(t.sensor_type))
        w.record(np.asarray(t.sensor_channel, np.int32))
        w.record(np.asarray(t.component_id, np.int32))
        w.record(np.asarray(t.absorber_id, np.int32))
        w.record(np.asarray(t.ref_level_pressure, np.float64),
                 np.asarray(t.ref_pressure, np.float64),
                 np.asarray(t.ref_temperature, np.float64),
                 np.asarray(t.ref_absorber, np.float64).reshape(Ko, Jm),
                 np.asarray(t.min_absorber, np.float64).reshape(Ko, Jm),
                 np.asarray(t.max_absorber, np.float64).reshape(Ko, Jm))
        w.record(np.asarray(t.n_predictors, np.int32).reshape(J, L),
                 np.asarray(t.pos_index, np.int32).reshape(J, L))
        if t.c.size > 0:
            w.record(np.asarray(t.c, np.float32))
        if t.n_ocoeffs > 0:
            w.record(np.asarray(t.osignificance, np.int32),
                     np.asarray(t.order, np.int32),
                     np.asarray(t.op_index, np.int32).reshape(N_PREDICTOR_USED_OPTRAN + 1, L),
                     np.asarray(t.opos_index, np.int32),
                     np.asarray(t.oc, np.float64),
                     np.float64(t.alpha), np.float64(t.alpha_c1), np.float64(t.alpha_c2),
                     np.int32(t.ocomponent_index))


def read_odps(path: str) -> ODPS:
    r = RecordReader(path)
    rel, _ver = r.read([("i4", ()), ("i4", ())])
    if rel != ODPS_RELEASE:
        raise ValueError(f"ODPS release {rel} unsupported")
    (alg,) = r.read([("i4", ())])
    if alg != ODPS_ALGORITHM:
        raise ValueError(f"not an ODPS file (algorithm {alg})")
    Ko, J, Jm, L, nc, noi, noc = [int(x) for x in r.read([("i4", ())] * 7)]
    (grp,) = r.read([("i4", ())])
    sid, wsat, wsen, stype = r.read([("s", None), ("i4", ()), ("i4", ()), ("i4", ())])
    (ch,) = r.read([("i4", (L,))])
    (cid,) = r.read([("i4", (J,))])
    (aid,) = r.read([("i4", (Jm,))])
    rlp, rp, rt, ra, mn, mx = r.read([("f8", (Ko + 1,)), ("f8", (Ko,)), ("f8", (Ko,)),
                                      ("f8", (Ko, Jm)), ("f8", (Ko, Jm)), ("f8", (Ko, Jm))])
    npred, pos = r.read([("i4", (J, L)), ("i4", (J, L))])
    c = r.read([("f4", (nc,))])[0] if nc > 0 else np.zeros(0, np.float32)
    t = ODPS(int(grp), sid, int(wsat), int(wsen), int(stype), ch.copy(), cid.copy(), aid.copy(),
             rlp.copy(), rp.copy(), rt.copy(), ra.copy(), mn.copy(), mx.copy(),
             npred.copy(), pos.copy(), c.copy())
    if noc > 0:
        o = r.read([("i4", (L,)), ("i4", (L,)), ("i4", (noi + 1, L)), ("i4", (L,)),
                    ("f8", (noc,)), ("f8", ()), ("f8", ()), ("f8", ()), ("i4", ())])
        t.osignificance, t.order, t.op_index, t.opos_index, t.oc = [x.copy() for x in o[:5]]
        t.alpha, t.alpha_c1, t.alpha_c2 = float(o[5]), float(o[6]), float(o[7])
        t.ocomponent_index = int(o[8])
    return t


# --------------------------------------------------------------------------
# CloudCoeff
# --------------------------------------------------------------------------
@dataclasses.dataclass
class CloudCoeff:
    """Mie/TAMU cloud optical property tables.

    Shapes (Fortran order kept, numpy arrays indexed the same way):
      ke/w/g_l_mw (nF_mw, nR_mw, nT), pcoeff_l_mw (nF_mw, nR_mw, nT, nL+1, nP)
      ke/w/g_s_mw (nF_mw, nR_mw, nD), pcoeff_s_mw (nF_mw, nR_mw, nD, nL+1, nP)
      ke/w/g_ir  (nF_ir, nR_ir, nD+1), pcoeff_ir  (nF_ir, nR_ir, nD+1, nL, nP)
    The backscatter (kb) tables are not part of the file.
    """
    frequency_mw: np.ndarray
    frequency_ir: np.ndarray
    reff_mw: np.ndarray
    reff_ir: np.ndarray
    temperature: np.ndarray
    density_mw: np.ndarray
    ke_l_mw: np.ndarray
    w_l_mw: np.ndarray
    g_l_mw: np.ndarray
    pcoeff_l_mw: np.ndarray
    ke_s_mw: np.ndarray
    w_s_mw: np.ndarray
    g_s_mw: np.ndarray
    pcoeff_s_mw: np.ndarray
    ke_ir: np.ndarray
    w_ir: np.ndarray
    g_ir: np.ndarray
    pcoeff_ir: np.ndarray
    release: int = CLOUDCOEFF_RELEASE
    version: int = CLOUDCOEFF_VERSION

    @property
    def n_legendre_terms(self) -> int:
        return int(self.pcoeff_l_mw.shape[3] - 1)

    @property
    def n_phase_elements(self) -> int:
        return int(self.pcoeff_l_mw.shape[4])

    @property
    def n_ir_densities(self) -> int:
        return int(self.density_mw.size + 1)


def write_cloudcoeff(path: str, cc: CloudCoeff) -> None:
    """Record order of CloudCoeff_Binary_ReadFile (CloudCoeff_Binary_IO.f90:367-446)."""
    f8 = lambda a: np.asarray(a, np.float64)
    with RecordWriter(path) as w:
        w.record(np.int32(cc.release), np.int32(cc.version))
        w.record(*[np.int32(x) for x in (cc.frequency_mw.size, cc.reff_mw.size, cc.frequency_ir.size,
                                         cc.reff_ir.size, cc.temperature.size, cc.density_mw.size,
                                         cc.n_legendre_terms, cc.n_phase_elements)])
        w.record(f8(cc.frequency_mw), f8(cc.frequency_ir), f8(cc.reff_mw), f8(cc.reff_ir),
                 f8(cc.temperature), f8(cc.density_mw))
        w.record(f8(cc.ke_l_mw), f8(cc.w_l_mw), f8(cc.g_l_mw), f8(cc.pcoeff_l_mw))
        w.record(f8(cc.ke_s_mw), f8(cc.w_s_mw), f8(cc.g_s_mw), f8(cc.pcoeff_s_mw))
        w.record(f8(cc.ke_ir), f8(cc.w_ir), f8(cc.g_ir), f8(cc.pcoeff_ir))


def read_cloudcoeff(path: str) -> CloudCoeff:
    r = RecordReader(path)
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel < CLOUDCOEFF_RELEASE:
        raise ValueError(f"CloudCoeff release {rel} unsupported (need {CLOUDCOEFF_RELEASE})")
    nfm, nrm, nfi, nri, nt, ndm, nl, npe = [int(x) for x in r.read([("i4", ())] * 8)]
    ndi = ndm + 1
    fm, fi, rm, ri, t, dm = r.read([("f8", (nfm,)), ("f8", (nfi,)), ("f8", (nrm,)), ("f8", (nri,)),
                                    ("f8", (nt,)), ("f8", (ndm,))])
    lm = r.read([("f8", (nfm, nrm, nt))] * 3 + [("f8", (nfm, nrm, nt, nl + 1, npe))])
    sm = r.read([("f8", (nfm, nrm, ndm))] * 3 + [("f8", (nfm, nrm, ndm, nl + 1, npe))])
    ir = r.read([("f8", (nfi, nri, ndi))] * 3 + [("f8", (nfi, nri, ndi, nl, npe))])
    return CloudCoeff(fm.copy(), fi.copy(), rm.copy(), ri.copy(), t.copy(), dm.copy(),
                      *[a.copy() for a in lm], *[a.copy() for a in sm], *[a.copy() for a in ir],
                      release=int(rel), version=int(ver))


# --------------------------------------------------------------------------
# AerosolCoeff (binary, release 4, 'CRTM' aerosol scheme)
# --------------------------------------------------------------------------
AEROSOLCOEFF_RELEASE = 4
AEROSOLCOEFF_VERSION = 301   # > 300: the file carries the backscatter table kb
AEROSOL_STRLEN = 80          # CHARACTER(SL), SL = 80 (AerosolCoeff_Define.f90)


@dataclasses.dataclass
class AerosolCoeff:
    """Aerosol optical property tables (AerosolCoeff_Define.f90:100-128).

    Shapes (Fortran order): reff (nR, nT), rsig (nS, nT), rh (n_RH_Radii,);
    ke/w/g/kb (nW, nRH, nR, nS, nT); pcoeff (nW, nRH, nR, nS, nT, nL+1, nP).
    Wavelengths in microns, stored in descending order so that the
    frequencies (10^4 / wavelength, cm^-1) the lookup interpolates in ascend.
    """
    type_id: np.ndarray
    type_name: List[str]
    wavelength: np.ndarray
    reff: np.ndarray
    rsig: np.ndarray
    rh: np.ndarray
    ke: np.ndarray
    w: np.ndarray
    g: np.ndarray
    kb: np.ndarray
    pcoeff: np.ndarray
    data_source: str = "synthetic"
    release: int = AEROSOLCOEFF_RELEASE
    version: int = AEROSOLCOEFF_VERSION

    @property
    def frequency(self) -> np.ndarray:
        """AerosolCoeff_Frequency: micron_to_inverse_cm of the wavelengths."""
        return 10000.0 / self.wavelength

    @property
    def n_legendre_terms(self) -> int:
        return int(self.pcoeff.shape[5] - 1)


def _fixed(s: str, n: int = AEROSOL_STRLEN) -> bytes:
    return s.encode("ascii").ljust(n)[:n]


def write_aerosolcoeff(path: str, ac: AerosolCoeff) -> None:
    """Record order of AerosolCoeff_Binary_ReadFile (AerosolCoeff_Binary_IO.f90:396-520)."""
    f8 = lambda a: np.asarray(a, np.float64)
    nW, nRH, nR, nS, nT = ac.ke.shape
    with RecordWriter(path) as w:
        w.record(np.int32(ac.release), np.int32(ac.version))
        w.record(*[np.int32(x) for x in (nW, nR, nS, nT, nRH, ac.n_legendre_terms, ac.pcoeff.shape[6])])
        w.record(np.int32(ac.rh.size))                 # n_RH_Radii ('CRTM' and 'CMAQ' schemes)
        w.record(np.int32(AEROSOL_STRLEN))
        w.record(_fixed(ac.data_source))
        w.record(np.asarray(ac.type_id, np.int32))
        w.record(np.int32(AEROSOL_STRLEN))
        w.record(b"".join(_fixed(n) for n in ac.type_name))
        w.record(f8(ac.wavelength), f8(ac.reff), f8(ac.rsig), f8(ac.rh))
        if ac.version <= 300:
            w.record(f8(ac.ke), f8(ac.w), f8(ac.g), f8(ac.pcoeff))
        else:
            w.record(f8(ac.ke), f8(ac.w), f8(ac.g), f8(ac.kb), f8(ac.pcoeff))


def read_aerosolcoeff(path: str) -> AerosolCoeff:
    r = RecordReader(path)
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != AEROSOLCOEFF_RELEASE:
        raise ValueError(f"AerosolCoeff release {rel} unsupported (need {AEROSOLCOEFF_RELEASE})")
    nW, nR, nS, nT, nRH, nL, nP = [int(x) for x in r.read([("i4", ())] * 7)]
    (nRHR,) = r.read([("i4", ())])
    (sl,) = r.read([("i4", ())])
    if sl != AEROSOL_STRLEN:
        raise ValueError("AerosolCoeff data source string length mismatch")
    src = r._next().decode("ascii").rstrip()
    (tid,) = r.read([("i4", (nT,))])
    (sl,) = r.read([("i4", ())])
    if sl != AEROSOL_STRLEN:
        raise ValueError("AerosolCoeff type name string length mismatch")
    raw = r._next()
    names = [raw[i * sl:(i + 1) * sl].decode("ascii").rstrip() for i in range(nT)]
    wl, reff, rsig, rh = r.read([("f8", (nW,)), ("f8", (nR, nT)), ("f8", (nS, nT)), ("f8", (int(nRHR),))])
    shp = (nW, nRH, nR, nS, nT)
    if ver <= 300:
        ke, w, g, pc = r.read([("f8", shp)] * 3 + [("f8", shp + (nL + 1, nP))])
        kb = np.zeros(shp)
    else:
        ke, w, g, kb, pc = r.read([("f8", shp)] * 4 + [("f8", shp + (nL + 1, nP))])
    return AerosolCoeff(tid.copy(), names, wl.copy(), reff.copy(), rsig.copy(), rh.copy(), ke.copy(), w.copy(),
                        g.copy(), kb.copy(), pc.copy(), data_source=src, release=int(rel), version=int(ver))


# --------------------------------------------------------------------------
# Surface emissivity model coefficients (EmisCoeff)
# --------------------------------------------------------------------------
GATT_COUNT = 5               # write_module, created_on, title, history, comment
FITCOEFF_RELEASE = 1
MWWATERCOEFF_RELEASE = 1
FASTEM6_VERSION = 6          # Compute_FastemX selects the FASTEM6 azimuth model on it
IRWATERCOEFF_RELEASE = 4
IRWATERCOEFF_VERSION = 2
IRWATER_NAME_LEN = 80        # CHARACTER(SL), SL = 80


GATT_NAMES = ("write_module", "created_on", "title", "history", "comment")


def _write_gatts(w: RecordWriter, title: str = "synthetic coefficients") -> None:
    """Global attributes of the binary files: per attribute a length record
    and, when the length is not zero, a 'name: value' text record
    (WriteGAtts_Binary_File / ReadGAtts_Binary_File, src/Utility/Binary_File_Utility.f90:333-357)."""
    values = {"write_module": "crtmv3_amd.coeffio", "title": title}
    for name in GATT_NAMES:
        v = values.get(name, "")
        text = f"{name}: {v}".encode("ascii") if v else b""
        w.record(np.int32(len(text)))
        if text:
            w.record(text)


def _skip_gatts(r: RecordReader) -> None:
    for _ in range(GATT_COUNT):
        (n,) = r.read([("i4", ())])
        if n > 0 and len(r._next()) != int(n):
            raise ValueError("global attribute length mismatch")


def _write_fitcoeff(w: RecordWriter, c: np.ndarray) -> None:
    """One FitCoeff_nD block (FitCoeff_ReadFile.inc: release/version, rank,
    dimensions, attributes, coefficients)."""
    c = np.asarray(c, np.float64)
    w.record(np.int32(FITCOEFF_RELEASE), np.int32(1))
    w.record(np.int32(c.ndim))
    w.record(np.asarray(c.shape, np.int32))
    _write_gatts(w)
    w.record(c)


def _read_fitcoeff(r: RecordReader, rank: int) -> np.ndarray:
    rel, _ = r.read([("i4", ()), ("i4", ())])
    if rel != FITCOEFF_RELEASE:
        raise ValueError(f"FitCoeff release {rel} unsupported")
    (nd,) = r.read([("i4", ())])
    if nd != rank:
        raise ValueError(f"FitCoeff rank {nd}, expected {rank}")
    dims = tuple(int(x) for x in r.read_ints(nd))
    _skip_gatts(r)
    (c,) = r.read([("f8", dims)])
    return c.copy()


@dataclasses.dataclass
class MWwaterCoeff:
    """FASTEM coefficient set (src/Coefficients/EmisCoeff/MW_Water/MWwaterCoeff_Define.f90:78-92).

    fc  foam coverage, C(1) * U**C(2)                           [2]
    fr  foam reflectivity (zenith polynomial, frequency factor) [6]
    rc  reflection correction, (term, predictor, polarisation)  [3, 7, 2]
    az  azimuth emissivity, (coefficient, frequency, pol)       [6, 6, 2]
    ssc small-scale correction                                  [8]
    lsc large-scale correction, (frequency power, term, pol)    [3, 6, 2]
    """
    fc: np.ndarray
    fr: np.ndarray
    rc: np.ndarray
    az: np.ndarray
    ssc: np.ndarray
    lsc: np.ndarray
    release: int = MWWATERCOEFF_RELEASE
    version: int = FASTEM6_VERSION


def write_mwwatercoeff(path: str, c: MWwaterCoeff) -> None:
    """Record order of MWwaterCoeff_ReadFile (MWwaterCoeff_Define.f90:699-922):
    release/version, attributes, FC, FR, RC, AZ, SSC, LSC FitCoeff blocks and
    the look-up-table flag (0: no table; 1 = DATA_PRESENT)."""
    with RecordWriter(path) as w:
        w.record(np.int32(c.release), np.int32(c.version))
        _write_gatts(w)
        for a in (c.fc, c.fr, c.rc, c.az, c.ssc, c.lsc):
            _write_fitcoeff(w, a)
        w.record(np.int32(0))


def read_mwwatercoeff(path: str) -> MWwaterCoeff:
    r = RecordReader(path)
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != MWWATERCOEFF_RELEASE:
        raise ValueError(f"MWwaterCoeff release {rel} unsupported")
    _skip_gatts(r)
    fc, fr, rc, az, ssc, lsc = [_read_fitcoeff(r, k) for k in (1, 1, 3, 3, 1, 3)]
    (lut,) = r.read([("i4", ())])
    if lut != 0:
        raise ValueError("MWwaterCoeff look-up tables are not supported")
    if fc.size < 2 or fr.size < 6 or rc.shape != (3, 7, 2) or az.shape[0] < 6 or az.shape[1] != 6 \
            or ssc.size < 8 or lsc.shape[1:] != (6, 2) or lsc.shape[0] < 3:
        raise ValueError("MWwaterCoeff: unexpected coefficient dimensions")
    return MWwaterCoeff(fc, fr, rc, az, ssc, lsc, release=int(rel), version=int(ver))


@dataclasses.dataclass
class IRwaterCoeff:
    """IR sea-surface emissivity table (src/Coefficients/EmisCoeff/IR_Water/IRwaterCoeff_Define.f90:84-106):
    emissivity [angle, frequency, wind speed, temperature] with angles in
    degrees and frequencies in cm^-1.  A 'Nalli2' classification
    interpolates in temperature too; otherwise temperature has one entry."""
    classification: str
    angle: np.ndarray
    frequency: np.ndarray
    wind_speed: np.ndarray
    temperature: np.ndarray
    emissivity: np.ndarray
    release: int = IRWATERCOEFF_RELEASE
    version: int = IRWATERCOEFF_VERSION


def write_irwatercoeff(path: str, c: IRwaterCoeff) -> None:
    """Record order of IRwaterCoeff_ReadFile (IRwaterCoeff_Define.f90:805-984)."""
    f8 = lambda a: np.asarray(a, np.float64)
    with RecordWriter(path) as w:
        w.record(np.int32(c.release), np.int32(c.version))
        w.record(*[np.int32(x) for x in (c.angle.size, c.frequency.size, c.wind_speed.size,
                                         c.temperature.size)])
        _write_gatts(w)
        w.record(_fixed(c.classification, IRWATER_NAME_LEN))
        w.record(f8(c.angle), f8(c.frequency), f8(c.wind_speed), f8(c.temperature))
        w.record(f8(c.emissivity))


def read_irwatercoeff(path: str) -> IRwaterCoeff:
    r = RecordReader(path)
    rel, ver = r.read([("i4", ()), ("i4", ())])
    if rel != IRWATERCOEFF_RELEASE:
        raise ValueError(f"IRwaterCoeff release {rel} unsupported")
    na, nf, nv, nt = [int(x) for x in r.read([("i4", ())] * 4)]
    _skip_gatts(r)
    name = r._next().decode("ascii").rstrip()
    a, f, v, t = r.read([("f8", (na,)), ("f8", (nf,)), ("f8", (nv,)), ("f8", (nt,))])
    (e,) = r.read([("f8", (na, nf, nv, nt))])
    return IRwaterCoeff(name, a.copy(), f.copy(), v.copy(), t.copy(), e.copy(), release=int(rel),
                        version=int(ver))
