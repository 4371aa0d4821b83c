"""Python mirror of the CRTM API over the C-ABI (include/crtm_mi355x.h).

Names follow the reference interface (src/CRTM_Module.F90:30-38):
  crtm_init(sensor_ids, file_path)            -> Context     (CRTM_Init)
  Context.forward(profiles, emissivity, ...)  -> RTSolution  (CRTM_Forward)
  Context.k_matrix(profiles, emissivity, ...) -> (RTSolution, KMatrix)  (CRTM_K_Matrix)
  Context.destroy()                                          (CRTM_Destroy)

Errors surface as CRTMError carrying the reference status code
(SUCCESS=0, FAILURE=3, src/Utility/Message_Handler.f90:27-31).

Host arrays (numpy) go through CRTM_MEM_HOST; torch tensors already on the
GPU go through CRTM_MEM_DEVICE with no host round trip.  The library must be
built (python -m crtmv3_amd.build); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcrtm_mi355x.so")

SUCCESS, FAILURE = 0, 3
PREC_FP64, PREC_FP32_SPLIT = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class CRTMError(RuntimeError):
    def __init__(self, msg, status=FAILURE):
        super().__init__(msg)
        self.status = status


class CAtmosphere(ctypes.Structure):
    _fields_ = [("n_profiles", ctypes.c_int), ("max_layers", ctypes.c_int), ("total_layers", ctypes.c_int64),
                ("n_layers", _vp), ("layer_offset", _vp),
                ("n_absorbers", ctypes.c_int), ("absorber_id", _vp), ("absorber_units", _vp),
                ("level_pressure", _vp), ("pressure", _vp), ("temperature", _vp), ("absorber", _vp),
                ("climatology", _vp), ("n_clouds", ctypes.c_int), ("cloud_type", _vp),
                ("cloud_water_content", _vp), ("cloud_effective_radius", _vp), ("cloud_fraction", _vp),
                ("n_aerosols", ctypes.c_int), ("aerosol_type", _vp), ("aerosol_concentration", _vp),
                ("aerosol_effective_radius", _vp)]


class CSurface(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("land_coverage", "water_coverage", "snow_coverage", "ice_coverage",
                                   "land_temperature", "water_temperature", "snow_temperature",
                                   "ice_temperature", "wind_speed", "wind_direction", "salinity")]


class CGeometry(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("sensor_zenith_angle", "sensor_scan_angle", "source_zenith_angle",
                                   "surface_altitude", "sensor_azimuth_angle", "ifov")]


class COptions(ctypes.Structure):
    _fields_ = [("use_emissivity", ctypes.c_int), ("emissivity", _vp), ("precision", ctypes.c_int),
                ("memory", ctypes.c_int), ("jacobian_f32", ctypes.c_int), ("check_input", ctypes.c_int),
                ("n_streams", ctypes.c_int), ("overlap_id", ctypes.c_int), ("n_streams_profile", _vp),
                ("no_scattering", ctypes.c_int), ("skip_profile", _vp),
                ("use_antenna_correction", ctypes.c_int), ("no_nlte_correction", ctypes.c_int)]


class CRTSolution(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("radiance", "brightness_temperature", "up_radiance", "down_radiance",
                                   "surface_planck_radiance", "surface_emissivity", "status", "r_clear",
                                   "tb_clear", "total_cloud_cover", "surface_reflectivity")]


class CKOutput(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("brightness_temperature_k", "radiance_k", "surface_emissivity_k",
                                   "temperature_k", "absorber_k", "land_temperature_k",
                                   "water_temperature_k", "snow_temperature_k", "ice_temperature_k",
                                   "cloud_water_content_k", "cloud_effective_radius_k", "cloud_fraction_k",
                                   "aerosol_concentration_k", "aerosol_effective_radius_k", "salinity_k",
                                   "wind_speed_k", "wind_direction_k")]


# adjoint outputs laid out per layer ([NT] or [n, NT]); the others are per profile
_LAYER_AD = ("temperature_ad", "absorber_ad", "cloud_water_content_ad", "cloud_effective_radius_ad",
             "cloud_fraction_ad", "aerosol_concentration_ad", "aerosol_effective_radius_ad")


class CADOutput(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("brightness_temperature_ad", "radiance_ad", "surface_emissivity_ad",
                                   "temperature_ad", "absorber_ad", "land_temperature_ad",
                                   "water_temperature_ad", "snow_temperature_ad", "ice_temperature_ad",
                                   "cloud_water_content_ad", "cloud_effective_radius_ad", "cloud_fraction_ad",
                                   "aerosol_concentration_ad", "aerosol_effective_radius_ad", "salinity_ad",
                                   "wind_speed_ad", "wind_direction_ad")]


_lib = None

N_TILES = 3   # tile shapes of the fused clear-sky kernel (CRTM_N_TILES)


def tile_profiles(precision: int, tile: int) -> int:
    """Profiles per workgroup of a tile index (crtm_tile_profiles)."""
    return int(lib().crtm_tile_profiles(int(precision), int(tile)))


def _tile_key(sensor_id: str, precision: int, n_profiles: int) -> str:
    try:
        import torch
        dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    except Exception:
        dev = "unknown"
    return f"{dev}|{sensor_id}|prec{int(precision)}|M{int(n_profiles)}"


def lib():
    """Load the native library; fails loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CRTMError(f"{LIB_PATH} missing: run `python -m crtmv3_amd.build`")
        _torch_device_first()
        L = ctypes.CDLL(LIB_PATH)
        L.crtm_init.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p,
                                ctypes.POINTER(_vp)]
        L.crtm_destroy.argtypes = [_vp]
        L.crtm_n_sensors.argtypes = [_vp]
        L.crtm_n_channels_total.argtypes = [_vp]
        L.crtm_channel_info.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip, ctypes.c_char_p]
        L.crtm_forward.argtypes = [_vp, ctypes.POINTER(CAtmosphere), ctypes.POINTER(CSurface),
                                   ctypes.POINTER(CGeometry), ctypes.POINTER(COptions),
                                   ctypes.POINTER(CRTSolution)]
        L.crtm_k_matrix.argtypes = L.crtm_forward.argtypes + [ctypes.POINTER(CKOutput)]
        L.crtm_k_matrix_noalloc.argtypes = L.crtm_k_matrix.argtypes
        L.crtm_adjoint.argtypes = L.crtm_forward.argtypes + [ctypes.POINTER(CADOutput)]
        L.crtm_synchronize.argtypes = [_vp]
        L.crtm_set_stream.argtypes = [_vp, _vp, ctypes.c_int]
        L.crtm_set_wavefront_angles.argtypes = [_vp, ctypes.c_int]
        L.crtm_set_cloudy_layer_capacity.argtypes = [_vp, ctypes.c_int64]
        L.crtm_tile_profiles.argtypes = [ctypes.c_int, ctypes.c_int]
        L.crtm_set_tile.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.crtm_get_tile.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _ip]
        L.crtm_autotune_k_matrix.argtypes = L.crtm_k_matrix.argtypes + [ctypes.c_int, _dp, _ip]
        L.crtm_last_error.argtypes = [_vp]
        L.crtm_last_error.restype = ctypes.c_char_p
        L.crtm_profile_enable.argtypes = [_vp, ctypes.c_int]
        L.crtm_profile_query.argtypes = [_vp, _dp, _dp, _ip]
        L.crtm_profile_query_kinds.argtypes = [_vp, _dp, _ip, ctypes.c_int]
        _lib = L
    return _lib


def _torch_device_first():
    """PyTorch-ROCm ships its own HIP runtime next to the system one the
    library links.  Both can drive the GPU of one process when torch's runtime
    opens the device first; the other order leaves torch without a device.
    So, when torch is installed and sees a GPU, it is initialised before the
    library is loaded."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        return x.data_ptr()
    return x.ctypes.data


@dataclasses.dataclass
class ChannelInfo:
    """CRTM_ChannelInfo_type (src/ChannelInfo/CRTM_ChannelInfo_Define.f90:70-85)."""
    sensor_id: str
    sensor_type: int
    n_channels: int
    sensor_channel: np.ndarray
    channel_offset: int


class Context:
    def __init__(self, handle, channel_info):
        self._h = handle
        self.channel_info = channel_info
        self.n_channels = sum(c.n_channels for c in channel_info)

    def destroy(self):
        if self._h:
            lib().crtm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _use_torch_stream(self):
        """Device-memory calls run on torch's current stream, so they are
        ordered after the kernels that produced their inputs and before the
        ones that consume their outputs."""
        import torch
        h = torch.cuda.current_stream().cuda_stream
        if h != getattr(self, "_stream", None):
            if lib().crtm_set_stream(self._h, h, 0) != SUCCESS:
                self._err("crtm_set_stream")
            self._stream = h

    def set_wavefront_angles(self, min_angles: int):
        """Angle counts >= min_angles (6 or 9) run on the wavefront
        (matrix-core) scattering kernels, smaller ones on the per-lane kernels."""
        if lib().crtm_set_wavefront_angles(self._h, int(min_angles)) != SUCCESS:
            self._err("crtm_set_wavefront_angles")

    def set_tile(self, sensor: int, precision: int, tile: int):
        """Tile index (0..N_TILES-1) of the fused clear-sky kernel for one
        sensor and precision: 64 channels x tile_profiles(precision, tile)
        profiles per workgroup."""
        if lib().crtm_set_tile(self._h, int(sensor), int(precision), int(tile)) != SUCCESS:
            self._err("crtm_set_tile")

    def get_tile(self, sensor: int, precision: int) -> int:
        t = ctypes.c_int()
        if lib().crtm_get_tile(self._h, int(sensor), int(precision), ctypes.byref(t)) != SUCCESS:
            self._err("crtm_get_tile")
        return t.value

    def autotune(self, prof, emissivity, precision=PREC_FP32_SPLIT, repeats=3, jacobian_f32=True, cache=None,
                 **kw):
        """Tile-shape autotuner (crtm_autotune_k_matrix): times the fused
        clear-sky kernel of every sensor at each tile on this batch and keeps
        the fastest tile per sensor at `precision`.  Returns {"ms": [n_sensors]
        [N_TILES] mean kernel times, "best": [n_sensors] tile indices,
        "profiles_per_workgroup": [n_sensors]}.  cache: a JSON file that
        records the choice under (device, sensor id, precision, profiles,
        channels); apply_tile_cache() applies it in a later process."""
        ns = len(self.channel_info)
        ms = np.zeros((ns, N_TILES), np.float64)
        best = np.zeros(ns, np.int32)
        self.k_matrix(prof, emissivity, precision=precision, jacobian_f32=jacobian_f32,
                      _tune=(int(repeats), ms, best), **kw)
        rep = {"ms": ms.tolist(), "best": best.tolist(),
               "profiles_per_workgroup": [tile_profiles(precision, int(b)) for b in best]}
        if cache is not None:
            import json
            try:
                d = json.load(open(cache))
            except (OSError, ValueError):
                d = {}
            for i, ci in enumerate(self.channel_info):
                d[_tile_key(ci.sensor_id, precision, prof.n_profiles)] = int(best[i])
            with open(cache, "w") as f:
                json.dump(d, f, indent=1, sort_keys=True)
        return rep

    def apply_tile_cache(self, cache, precision, n_profiles) -> bool:
        """Set the tiles an earlier autotune() recorded in `cache` for these
        sensors at this precision and batch size; False if any is missing."""
        import json
        try:
            d = json.load(open(cache))
        except (OSError, ValueError):
            return False
        keys = [_tile_key(ci.sensor_id, precision, n_profiles) for ci in self.channel_info]
        if not all(k in d for k in keys):
            return False
        for i, k in enumerate(keys):
            self.set_tile(i, precision, int(d[k]))
        return True

    def set_cloudy_layer_capacity(self, max_layers: int):
        """Size the scattering launches for max_layers scattering layers of a
        batch instead of reading the count back (0: read back, the default).
        With a capacity, device-memory calls with clouds or aerosols can be
        captured in a HIP graph; profiles whose layers do not fit fail."""
        if lib().crtm_set_cloudy_layer_capacity(self._h, int(max_layers)) != SUCCESS:
            self._err("crtm_set_cloudy_layer_capacity")

    def _err(self, what):
        msg = lib().crtm_last_error(self._h).decode()
        raise CRTMError(f"{what}: {msg}")

    # -- argument marshalling ------------------------------------------------
    def _structs(self, prof, emissivity, precision, memory, jacobian_f32, check_input, keep, n_streams=0,
                 overlap=0, include_scattering=True, skip_profile=None, use_antenna_correction=False,
                 apply_nlte_correction=True):
        def arr(x, dt):
            if x is None:
                return None
            if memory == MEM_DEVICE:
                return x
            a = np.ascontiguousarray(x, dt)
            keep.append(a)
            return a
        atm = CAtmosphere()
        atm.n_profiles = prof.n_profiles
        host_nl = getattr(prof, "n_layers_host", None)
        if host_nl is None and memory == MEM_HOST:
            host_nl = np.asarray(prof.n_layers)
        if host_nl is not None and len(host_nl):
            atm.max_layers = int(np.max(host_nl))
            atm.total_layers = int(np.sum(host_nl))
        atm.n_layers = _ptr(arr(prof.n_layers, np.int32))
        atm.layer_offset = _ptr(arr(prof.layer_offset, np.int64))
        atm.n_absorbers = int(prof.absorber.shape[0])
        ids = np.ascontiguousarray(np.asarray(prof.absorber_id if not _is_torch(prof.absorber_id)
                                              else prof.absorber_id.cpu()), np.int32)
        units = np.ascontiguousarray(np.asarray(prof.absorber_units if not _is_torch(prof.absorber_units)
                                                else prof.absorber_units.cpu()), np.int32)
        keep.extend([ids, units])
        atm.absorber_id = _ptr(ids)
        atm.absorber_units = _ptr(units)
        atm.level_pressure = _ptr(arr(prof.level_pressure, np.float64))
        atm.pressure = _ptr(arr(prof.pressure, np.float64))
        atm.temperature = _ptr(arr(prof.temperature, np.float64))
        atm.absorber = _ptr(arr(prof.absorber, np.float64))
        clim = getattr(prof, "climatology", None)
        atm.climatology = _ptr(arr(clim, np.int32)) if clim is not None else None
        ctype = getattr(prof, "cloud_type", None)
        if ctype is not None and int(ctype.shape[1]) > 0:
            atm.n_clouds = int(ctype.shape[1])
            atm.cloud_type = _ptr(arr(ctype, np.int32))
            atm.cloud_water_content = _ptr(arr(prof.cloud_water_content, np.float64))
            atm.cloud_effective_radius = _ptr(arr(prof.cloud_effective_radius, np.float64))
            cf = getattr(prof, "cloud_fraction", None)
            atm.cloud_fraction = _ptr(arr(cf, np.float64)) if cf is not None else None
        atype = getattr(prof, "aerosol_type", None)
        if atype is not None and int(atype.shape[1]) > 0:
            atm.n_aerosols = int(atype.shape[1])
            atm.aerosol_type = _ptr(arr(atype, np.int32))
            atm.aerosol_concentration = _ptr(arr(prof.aerosol_concentration, np.float64))
            atm.aerosol_effective_radius = _ptr(arr(prof.aerosol_effective_radius, np.float64))
        sfc = CSurface()
        for f in ("land_coverage", "water_coverage", "snow_coverage", "ice_coverage", "land_temperature",
                  "water_temperature", "snow_temperature", "ice_temperature"):
            setattr(sfc, f, _ptr(arr(getattr(prof, f), np.float64)))
        for f in ("wind_speed", "wind_direction", "salinity"):   # optional (reference defaults)
            v = getattr(prof, f, None)
            setattr(sfc, f, _ptr(arr(v, np.float64)) if v is not None else None)
        geo = CGeometry()
        geo.sensor_zenith_angle = _ptr(arr(prof.sensor_zenith, np.float64))
        geo.sensor_scan_angle = _ptr(arr(prof.sensor_scan, np.float64))
        geo.source_zenith_angle = _ptr(arr(prof.source_zenith, np.float64))
        az = getattr(prof, "sensor_azimuth", None)
        geo.sensor_azimuth_angle = _ptr(arr(az, np.float64)) if az is not None else None
        ifov = getattr(prof, "ifov", None)
        geo.ifov = _ptr(arr(ifov, np.int32)) if ifov is not None else None
        opt = COptions()
        opt.use_antenna_correction = int(bool(use_antenna_correction))
        opt.no_nlte_correction = 0 if apply_nlte_correction else 1
        # emissivity None: Options%Use_Emissivity = .FALSE., the surface optics are computed
        opt.use_emissivity = 0 if emissivity is None else 1
        opt.emissivity = _ptr(arr(emissivity, np.float64))
        opt.precision = precision
        opt.memory = memory
        opt.jacobian_f32 = int(jacobian_f32)
        opt.check_input = int(check_input)
        if np.ndim(n_streams) == 0 and not _is_torch(n_streams):
            opt.n_streams = int(n_streams)
        else:   # Options(m)%n_Streams per profile: host memory unless given as a device tensor
            if _is_torch(n_streams) and n_streams.is_cuda:
                import torch
                ns = n_streams.to(torch.int32).contiguous()
            else:
                ns = np.ascontiguousarray(np.asarray(n_streams.cpu() if _is_torch(n_streams) else n_streams),
                                          np.int32)
            keep.append(ns)
            if ns.shape[0] != prof.n_profiles:
                raise CRTMError(f"n_streams: {ns.shape[0]} values for {prof.n_profiles} profiles")
            opt.n_streams = 0
            opt.n_streams_profile = _ptr(ns)
        opt.overlap_id = int(overlap)
        opt.no_scattering = 0 if include_scattering else 1
        if skip_profile is not None:   # Options(m)%Skip_Profile, in the call's memory
            if _is_torch(skip_profile):
                import torch
                sk = skip_profile.to(torch.int32).contiguous()
            else:
                sk = np.ascontiguousarray(np.asarray(skip_profile), np.int32)
            if sk.shape[0] != prof.n_profiles:
                raise CRTMError(f"skip_profile: {sk.shape[0]} values for {prof.n_profiles} profiles")
            keep.append(sk)
            opt.skip_profile = _ptr(sk)
        return atm, sfc, geo, opt

    def _alloc(self, shape, dtype, memory, like=None):
        if memory == MEM_DEVICE:
            import torch
            tdt = {np.float64: torch.float64, np.float32: torch.float32, np.int32: torch.int32}[dtype]
            return torch.zeros(shape, dtype=tdt, device=like.device)
        return np.zeros(shape, dtype)

    def forward(self, prof, emissivity, precision=PREC_FP64, extras=False, check_input=True, n_streams=0,
                overlap=0, include_scattering=True, skip_profile=None, use_antenna_correction=False,
                apply_nlte_correction=True):
        """CRTM_Forward over a ProfileSet (host numpy) or a device batch (torch).

        Profiles carrying clouds take the all-sky path; n_streams is
        Options%n_Streams (0: chosen from the Mie parameter), or one value per
        profile (Options(m)%n_Streams), and overlap Options%Overlap_Id
        (0: average).  include_scattering=False is Options%Include_Scattering
        = .FALSE. (clouds and aerosols absorb only); skip_profile [M] holds
        Options(m)%Skip_Profile (skipped profiles: zero outputs, SUCCESS)."""
        memory = MEM_DEVICE if _is_torch(prof.temperature) else MEM_HOST
        if memory == MEM_DEVICE:
            self._use_torch_stream()
        keep = []
        atm, sfc, geo, opt = self._structs(prof, emissivity, precision, memory, False, check_input, keep,
                                           n_streams, overlap, include_scattering, skip_profile,
                                           use_antenna_correction, apply_nlte_correction)
        M, L = prof.n_profiles, self.n_channels
        out = {k: self._alloc((M, L), np.float64, memory, prof.temperature)
               for k in ("radiance", "brightness_temperature")}
        if extras:
            for k in ("up_radiance", "down_radiance", "surface_planck_radiance", "surface_emissivity",
                      "surface_reflectivity"):
                out[k] = self._alloc((M, L), np.float64, memory, prof.temperature)
            if atm.n_clouds > 0:
                for k in ("r_clear", "tb_clear", "total_cloud_cover"):
                    out[k] = self._alloc((M, L), np.float64, memory, prof.temperature)
        out["status"] = self._alloc((M,), np.int32, memory, prof.temperature)
        rts = CRTSolution(**{k: _ptr(v) for k, v in out.items()})
        st = lib().crtm_forward(self._h, atm, sfc, geo, opt, rts)
        if st != SUCCESS:
            self._err("crtm_forward")
        return out

    def k_matrix(self, prof, emissivity, bt_k=None, rad_k=None, precision=PREC_FP64,
                 jacobian_f32=False, check_input=True, out=None, n_streams=0, overlap=0, extras=False,
                 no_alloc=False, include_scattering=True, skip_profile=None, use_antenna_correction=False,
                 apply_nlte_correction=True, _tune=None):
        """CRTM_K_Matrix: returns (RTSolution dict, K dict).

        K layout: temperature_k [NT, L] (profile m rows layer_offset[m]:...),
        absorber_k [n_absorbers, NT, L]; surface temperature K per type [M, L];
        with clouds also cloud_water_content_k and cloud_effective_radius_k
        [n_clouds, NT, L] and cloud_fraction_k [NT, L].
        bt_k / rad_k are RTSolution_K inputs (BT_K defaults to 1, Radiance_K
        to 0).  `out` = (rts, k) dicts from a previous call are reused as the
        output buffers (every element is overwritten).  n_streams / overlap as
        in forward(); extras adds R_Clear, Tb_Clear and the cloud cover of a
        cloudy call to the RTSolution dict.  no_alloc calls
        crtm_k_matrix_noalloc (clear-sky device batches: the library's work
        space comes from its fixed scratch array, nothing is allocated).
        include_scattering / skip_profile as in forward(); a skipped profile's
        Jacobians are zero and its radiance_k is left as given."""
        memory = MEM_DEVICE if _is_torch(prof.temperature) else MEM_HOST
        if memory == MEM_DEVICE:
            self._use_torch_stream()
        keep = []
        atm, sfc, geo, opt = self._structs(prof, emissivity, precision, memory, jacobian_f32,
                                           check_input, keep, n_streams, overlap, include_scattering,
                                           skip_profile, use_antenna_correction, apply_nlte_correction)
        nc, nae = int(atm.n_clouds), int(atm.n_aerosols)
        M, L = prof.n_profiles, self.n_channels
        host_nl = getattr(prof, "n_layers_host", None)
        NT = int(np.sum(host_nl if host_nl is not None else np.asarray(prof.n_layers)))
        na = int(prof.absorber.shape[0])
        like = prof.temperature
        kdt = np.float32 if jacobian_f32 else np.float64
        if out is not None:
            rts_d, kd = out
        else:
            rts_d = {k: self._alloc((M, L), np.float64, memory, like)
                     for k in ("radiance", "brightness_temperature")}
            rts_d["status"] = self._alloc((M,), np.int32, memory, like)
            kd = {"surface_emissivity_k": self._alloc((M, L), np.float64, memory, like),
                  "temperature_k": self._alloc((NT, L), kdt, memory, like),
                  "absorber_k": self._alloc((na, NT, L), kdt, memory, like)}
            for f in ("land_temperature_k", "water_temperature_k", "snow_temperature_k", "ice_temperature_k"):
                kd[f] = self._alloc((M, L), np.float64, memory, like)
            if emissivity is None:   # computed surface optics: their input sensitivities
                for f in ("salinity_k", "wind_speed_k", "wind_direction_k"):
                    kd[f] = self._alloc((M, L), np.float64, memory, like)
            if nae > 0:
                kd["aerosol_concentration_k"] = self._alloc((nae, NT, L), kdt, memory, like)
                kd["aerosol_effective_radius_k"] = self._alloc((nae, NT, L), kdt, memory, like)
            if nc > 0:
                kd["cloud_water_content_k"] = self._alloc((nc, NT, L), kdt, memory, like)
                kd["cloud_effective_radius_k"] = self._alloc((nc, NT, L), kdt, memory, like)
                kd["cloud_fraction_k"] = self._alloc((NT, L), kdt, memory, like)
                if extras:
                    for k in ("r_clear", "tb_clear", "total_cloud_cover"):
                        rts_d[k] = self._alloc((M, L), np.float64, memory, like)
        if bt_k is None:
            bt_k = self._alloc((M, L), np.float64, memory, like) + 1.0
        if memory == MEM_HOST:
            bt_k = np.ascontiguousarray(bt_k, np.float64)
            keep.append(bt_k)
        if rad_k is not None:
            kd["radiance_k"] = rad_k.clone() if _is_torch(rad_k) else np.array(rad_k, np.float64)
        rts = CRTSolution(**{k: _ptr(v) for k, v in rts_d.items()})
        ko = CKOutput(brightness_temperature_k=_ptr(bt_k), **{k: _ptr(v) for k, v in kd.items()})
        if _tune is not None:   # autotune(): the same call, run per tile
            entry = "crtm_autotune_k_matrix"
            reps, ms, best = _tune
            st = lib().crtm_autotune_k_matrix(self._h, atm, sfc, geo, opt, rts, ko, reps,
                                              ms.ctypes.data_as(_dp), best.ctypes.data_as(_ip))
        else:
            entry = "crtm_k_matrix_noalloc" if no_alloc else "crtm_k_matrix"
            st = getattr(lib(), entry)(self._h, atm, sfc, geo, opt, rts, ko)
        if st != SUCCESS:
            self._err(entry)
        return rts_d, kd

    def adjoint(self, prof, emissivity, bt_ad, precision=PREC_FP64, n_streams=0, overlap=0,
                include_scattering=True, skip_profile=None, max_work_bytes=16 << 30, progress=None,
                use_antenna_correction=False, apply_nlte_correction=True):
        """CRTM_Adjoint (src/CRTM_Adjoint_Module.f90:299-325) through
        crtm_adjoint: the channel-summed adjoint of the brightness
        temperatures seeded with bt_ad [M, L] (RTSolution_AD%
        Brightness_Temperature).  The kernels sum over channels before any
        per-profile row is written; no per-channel Jacobian is formed.

        Returns (bt [M, L], ad) with ad holding temperature_ad [NT],
        absorber_ad [n_absorbers, NT], surface temperature adjoints per type
        [M], surface_emissivity_ad [M, L] (per channel) and, with clouds /
        aerosols, cloud_water_content_ad, cloud_effective_radius_ad
        [n_clouds, NT], cloud_fraction_ad [NT], aerosol_concentration_ad,
        aerosol_effective_radius_ad [n_aerosols, NT].  Host (numpy) or device
        (torch) batches.  A batch with clouds or aerosols runs in profile
        chunks whose scattering work space (per layer and channel: the optical
        depth, Planck and temperature seeds of the scattering adjoint and the
        cloudy layers' optics and tapes, estimated at 64 bytes FP32-split and
        128 FP64) stays within max_work_bytes; a clear-sky batch is one call.  n_streams may
        be one value or one per profile; progress(message) is called after
        each chunk."""
        memory = MEM_DEVICE if _is_torch(prof.temperature) else MEM_HOST
        M, L = prof.n_profiles, self.n_channels
        nl = np.asarray(getattr(prof, "n_layers_host", None) if memory == MEM_DEVICE else prof.n_layers, np.int64)
        loh = np.concatenate([[0], np.cumsum(nl)])
        NT = int(loh[-1])
        na = int(prof.absorber.shape[0])
        nc = 0 if prof.cloud_type is None else int(prof.cloud_type.shape[1])
        nae = 0 if prof.aerosol_type is None else int(prof.aerosol_type.shape[1])
        like = prof.temperature
        ad = {"temperature_ad": self._alloc((NT,), np.float64, memory, like),
              "absorber_ad": self._alloc((na, NT), np.float64, memory, like),
              "surface_emissivity_ad": self._alloc((M, L), np.float64, memory, like)}
        for f in ("land", "water", "snow", "ice"):
            ad[f + "_temperature_ad"] = self._alloc((M,), np.float64, memory, like)
        if emissivity is None:
            for f in ("salinity_ad", "wind_speed_ad", "wind_direction_ad"):
                ad[f] = self._alloc((M,), np.float64, memory, like)
        if nc:
            ad["cloud_water_content_ad"] = self._alloc((nc, NT), np.float64, memory, like)
            ad["cloud_effective_radius_ad"] = self._alloc((nc, NT), np.float64, memory, like)
            ad["cloud_fraction_ad"] = self._alloc((NT,), np.float64, memory, like)
        if nae:
            ad["aerosol_concentration_ad"] = self._alloc((nae, NT), np.float64, memory, like)
            ad["aerosol_effective_radius_ad"] = self._alloc((nae, NT), np.float64, memory, like)
        bt = self._alloc((M, L), np.float64, memory, like)
        scattering = nc > 0 or nae > 0
        # scattering work per layer and channel: three seeds and the gas OD in
        # the call's precision, plus the cloudy layers' optics, R/T matrices
        # and tapes, taken here as a few percent of the layers at up to ~2 KB
        per_layer = L * 64 * (1 if precision == PREC_FP32_SPLIT else 2)
        per_profile_streams = not (np.ndim(n_streams) == 0 and not _is_torch(n_streams))
        m0 = 0
        while m0 < M:
            m1 = M
            if scattering:
                m1 = m0 + 1
                while m1 < M and (loh[m1 + 1] - loh[m0]) * per_layer <= max_work_bytes:
                    m1 += 1
            a, b = int(loh[m0]), int(loh[m1])
            whole = m0 == 0 and m1 == M
            sub = prof if whole else profile_slice(prof, m0, m1)
            ns = n_streams[m0:m1] if (per_profile_streams and not whole) else n_streams
            sk = skip_profile if (skip_profile is None or whole) else skip_profile[m0:m1]
            part = {k: (v[..., a:b] if k in _LAYER_AD else v[m0:m1]) for k, v in ad.items()}
            self._adjoint_call(sub, emissivity if (whole or emissivity is None) else emissivity[m0:m1],
                               bt_ad if whole else bt_ad[m0:m1],
                               precision, ns, overlap, include_scattering, sk, bt[m0:m1], part, m0,
                               use_antenna_correction, apply_nlte_correction)
            if progress:
                progress(f"adjoint profiles {m0}-{m1} of {M} done")
            m0 = m1
        return bt, ad

    def _adjoint_call(self, prof, emissivity, bt_ad, precision, n_streams, overlap, include_scattering,
                      skip_profile, bt_out, ad, m0, use_antenna_correction=False,
                      apply_nlte_correction=True):
        memory = MEM_DEVICE if _is_torch(prof.temperature) else MEM_HOST
        if memory == MEM_DEVICE:
            self._use_torch_stream()
        keep = []
        if emissivity is not None:
            emissivity = emissivity.contiguous() if memory == MEM_DEVICE else np.ascontiguousarray(emissivity, np.float64)
        bt_ad = bt_ad.contiguous() if memory == MEM_DEVICE else np.ascontiguousarray(bt_ad, np.float64)
        atm, sfc, geo, opt = self._structs(prof, emissivity, precision, memory, False, False, keep, n_streams,
                                           overlap, include_scattering, skip_profile, use_antenna_correction,
                                           apply_nlte_correction)
        M, L = prof.n_profiles, self.n_channels
        rts_d = {"radiance": self._alloc((M, L), np.float64, memory, prof.temperature),
                 "brightness_temperature": bt_out, "status": self._alloc((M,), np.int32, memory, prof.temperature)}
        # a chunk's layer slice of an [n, NT] array is strided: the library
        # accumulates into a contiguous copy, copied back below
        cont = {}
        for k, v in ad.items():
            if _is_torch(v):
                cont[k] = v if v.is_contiguous() else v.contiguous()
            else:
                cont[k] = v if v.flags.c_contiguous else np.ascontiguousarray(v)
        rts = CRTSolution(**{k: _ptr(v) for k, v in rts_d.items()})
        out = CADOutput(brightness_temperature_ad=_ptr(bt_ad), **{k: _ptr(v) for k, v in cont.items()})
        keep.append(bt_ad)
        st = lib().crtm_adjoint(self._h, atm, sfc, geo, opt, rts, out)
        if st != SUCCESS:
            self._err("crtm_adjoint")
        for k, v in ad.items():
            if cont[k] is not v:
                v[...] = cont[k]
        if memory == MEM_DEVICE:
            _raise_on_failed(rts_d, m0, "adjoint")

    def tangent_linear(self, prof, emissivity, tl, precision=PREC_FP64, n_streams=0, overlap=0,
                       jacobian_f32=None, max_jacobian_bytes=8 << 30):
        """CRTM_Tangent_Linear (src/CRTM_Tangent_Linear_Module.f90:271-296) for a
        device batch: brightness-temperature perturbations [M, L] for the
        state perturbations in `tl`, a dict with any of temperature [NT],
        absorber [n_absorbers, NT], land/water/snow/ice_temperature [M],
        surface_emissivity [M, L], cloud_water_content /
        cloud_effective_radius [n_clouds, NT], cloud_fraction [NT],
        aerosol_concentration / aerosol_effective_radius [n_aerosols, NT].

        Computed as K-matrix (RTSolution_K = 1) contracted with the
        perturbations over layers, in profile chunks; returns (bt, bt_tl)."""
        import torch
        if not _is_torch(prof.temperature):
            raise CRTMError("tangent_linear runs on device batches (api.device_profiles)")
        if jacobian_f32 is None:
            jacobian_f32 = precision == PREC_FP32_SPLIT
        M, L = prof.n_profiles, self.n_channels
        nl = np.asarray(prof.n_layers_host, np.int64)
        loh = np.concatenate([[0], np.cumsum(nl)])
        na = int(prof.absorber.shape[0])
        nc = 0 if prof.cloud_type is None else int(prof.cloud_type.shape[1])
        nae = 0 if prof.aerosol_type is None else int(prof.aerosol_type.shape[1])
        per_layer = L * (2 + na + (3 * nc + 1 if nc else 0) + 2 * nae) * (4 if jacobian_f32 else 8)
        dev = prof.temperature.device
        bt = torch.empty(M, L, dtype=torch.float64, device=dev)
        bt_tl = torch.zeros(M, L, dtype=torch.float64, device=dev)
        per_profile_streams = not (np.ndim(n_streams) == 0 and not _is_torch(n_streams))
        layer_fields = {"temperature": ("temperature_k", None), "cloud_fraction": ("cloud_fraction_k", None),
                        "absorber": ("absorber_k", 1), "cloud_water_content": ("cloud_water_content_k", 1),
                        "cloud_effective_radius": ("cloud_effective_radius_k", 1),
                        "aerosol_concentration": ("aerosol_concentration_k", 1),
                        "aerosol_effective_radius": ("aerosol_effective_radius_k", 1)}
        m0 = 0
        while m0 < M:
            m1 = m0 + 1
            while m1 < M and (loh[m1 + 1] - loh[m0]) * per_layer <= max_jacobian_bytes:
                m1 += 1
            a, b = int(loh[m0]), int(loh[m1])
            sub = profile_slice(prof, m0, m1)
            ns = n_streams[m0:m1] if per_profile_streams else n_streams
            rts, kd = self.k_matrix(sub, emissivity[m0:m1].contiguous(), precision=precision,
                                    jacobian_f32=jacobian_f32, check_input=False, n_streams=ns, overlap=overlap)
            _raise_on_failed(rts, m0, "tangent_linear")
            bt[m0:m1] = rts["brightness_temperature"]
            acc = bt_tl[m0:m1]
            # layer fields: profile m's rows [loff[m], loff[m+1]) of the chunk
            seg = torch.as_tensor(np.repeat(np.arange(m1 - m0), nl[m0:m1]), device=dev)
            for name, (kname, stacked) in layer_fields.items():
                x = tl.get(name)
                if x is None or kname not in kd:
                    continue
                k = kd[kname].to(torch.float64)
                if stacked is None:
                    contrib = k * x[a:b, None]                         # [NTc, L]
                else:
                    contrib = (k * x[:, a:b, None]).sum(dim=0)         # [NTc, L]
                acc.index_add_(0, seg, contrib)
            for f in ("land", "water", "snow", "ice"):
                x = tl.get(f + "_temperature")
                if x is not None:
                    acc += kd[f + "_temperature_k"] * x[m0:m1, None]
            if tl.get("surface_emissivity") is not None:
                acc += kd["surface_emissivity_k"] * tl["surface_emissivity"][m0:m1]
            del rts, kd
            m0 = m1
        return bt, bt_tl

    def profile_enable(self, on=True):
        if lib().crtm_profile_enable(self._h, int(on)) != SUCCESS:
            self._err("crtm_profile_enable")

    def profile_query_kinds(self):
        """Milliseconds and launches per kernel kind since the last enable/query:
        ([prep, forward/K, scattering, surface optics], [n_prep, n_main, n_scattering, n_surface])."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_int * 4)()
        if lib().crtm_profile_query_kinds(self._h, ms, n, 4) != SUCCESS:
            self._err("crtm_profile_query_kinds")
        return list(ms), list(n)

    def profile_query(self):
        """(prep_ms, main_ms, n_launches) summed since the last enable/query."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        if lib().crtm_profile_query(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)) != SUCCESS:
            self._err("crtm_profile_query")
        return a.value, b.value, n.value

    def synchronize(self):
        if lib().crtm_synchronize(self._h) != SUCCESS:
            self._err("crtm_synchronize")


def _raise_on_failed(rts, m0, what):
    """A device-memory call reports per-profile failures only in its status
    array; a failed profile has no Jacobian, so no adjoint or gradient may be
    built from the call (the reference returns FAILURE for it,
    src/CRTM_Adjoint_Module.f90:404-423)."""
    st = rts["status"]
    st = st.cpu().numpy() if _is_torch(st) else np.asarray(st)
    bad = np.flatnonzero(st)
    if bad.size:
        raise CRTMError(f"{what}: profile #{m0 + int(bad[0]) + 1} failed (status {int(st[bad[0]])})")


def crtm_init(sensor_ids: Sequence[str], file_path: str) -> Context:
    """CRTM_Init (src/CRTM_LifeCycle.f90:520)."""
    L = lib()
    ids = (ctypes.c_char_p * len(sensor_ids))(*[s.encode() for s in sensor_ids])
    h = _vp()
    st = L.crtm_init(len(sensor_ids), ids, file_path.encode(), ctypes.byref(h))
    if st != SUCCESS:
        raise CRTMError(f"crtm_init failed for {list(sensor_ids)} in {file_path}", st)
    infos = []
    off = 0
    for i in range(L.crtm_n_sensors(h)):
        n, t = ctypes.c_int(), ctypes.c_int()
        L.crtm_channel_info(h, i, ctypes.byref(n), ctypes.byref(t), None, None)
        ch = np.zeros(n.value, np.int32)
        sid = ctypes.create_string_buffer(32)
        L.crtm_channel_info(h, i, None, None, ch.ctypes.data_as(_ip), sid)
        infos.append(ChannelInfo(sid.value.decode(), t.value, n.value, ch, off))
        off += n.value
    return Context(h, infos)


def profile_slice(prof, m0, m1):
    """Profiles [m0, m1) of a batch (host numpy or device torch) as a batch of
    its own: per-profile arrays sliced, per-layer arrays cut at the layer
    offsets, offsets rebased to the first profile."""
    host_nl = getattr(prof, "n_layers_host", None)
    nl = np.asarray(host_nl if host_nl is not None else prof.n_layers, np.int64)
    loh = np.concatenate([[0], np.cumsum(nl)])
    a, b = int(loh[m0]), int(loh[m1])
    torch_batch = _is_torch(prof.temperature)

    def cut(v):
        return v.contiguous() if torch_batch else np.ascontiguousarray(v)
    per_layer = ("pressure", "temperature", "cloud_fraction")
    per_layer_2d = ("absorber", "cloud_water_content", "cloud_effective_radius", "aerosol_concentration",
                    "aerosol_effective_radius")
    shared = ("absorber_id", "absorber_units")
    d = {}
    for f in dataclasses.fields(prof):
        v = getattr(prof, f.name)
        if v is None or f.name in shared:
            continue
        if f.name == "layer_offset":
            d[f.name] = cut(v[m0:m1 + 1] - v[m0])
        elif f.name == "level_pressure":
            d[f.name] = cut(v[a + m0:b + m1])
        elif f.name in per_layer:
            d[f.name] = cut(v[a:b])
        elif f.name in per_layer_2d:
            d[f.name] = cut(v[:, a:b])
        else:   # per profile (and the cloud / aerosol type rows)
            d[f.name] = cut(v[m0:m1])
    out = dataclasses.replace(prof, **d)
    if host_nl is not None:
        out.n_layers_host = np.asarray(host_nl)[m0:m1]
    return out


def device_profiles(prof, device="cuda"):
    """Copy a synthetic.ProfileSet to torch tensors on the GPU (device-memory calls)."""
    import torch
    d = {}
    for f in dataclasses.fields(prof):
        v = getattr(prof, f.name)
        if f.name in ("absorber_id", "absorber_units"):   # read by the host side of the call
            d[f.name] = None if v is None else np.ascontiguousarray(v)
            continue
        d[f.name] = None if v is None else torch.as_tensor(np.ascontiguousarray(v)).to(device)
    out = dataclasses.replace(prof, **d)
    # host copy of the layer counts: lets device calls size their work without a sync
    out.n_layers_host = np.asarray(prof.n_layers)
    return out
