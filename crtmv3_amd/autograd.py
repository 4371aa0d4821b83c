"""torch.autograd.Function over the C-ABI: brightness temperatures as a
differentiable function of the atmospheric state, clouds and aerosols.

Forward is CRTM_Forward (src/CRTM_Forward_Module.f90:229-245); backward is
CRTM_Adjoint (src/CRTM_Adjoint_Module.f90:299-325) seeded with the incoming
gradient as RTSolution_AD%Brightness_Temperature: the channel-summed adjoint
of crtm_adjoint (Context.adjoint), which never holds a per-channel Jacobian.
Both run on the GPU in device memory on torch's current stream.

    bt = crtm_bt(ctx, profiles, temperature, absorber, water_temperature,
                 emissivity, precision=api.PREC_FP32_SPLIT,
                 cloud_water_content=wc, cloud_effective_radius=re,
                 aerosol_concentration=ac, n_streams=ns)
    J = (((y - bt) / sigma) ** 2).sum(); J.backward()

`profiles` is a device batch from api.device_profiles (it supplies the
pressures, layer counts, coverages, cloud / aerosol types and geometry); the
differentiable inputs replace its temperature [NT], absorber [n_absorbers,
NT] and water temperature [M], and, when given, its cloud water content and
effective radius [n_clouds, NT] and aerosol concentration and effective
radius [n_aerosols, NT].  Gradients flow to those and to the emissivity
[M, L].  With emissivity None the water surface optics are computed
(Options%Use_Emissivity = .FALSE.), and wind_speed, wind_direction and
salinity [M], when given, are differentiable inputs too.  n_streams is
Options%n_Streams, one value or one per profile.
"""
from __future__ import annotations

import dataclasses

import torch

from . import api

_FIELDS = ("temperature", "absorber", "water_temperature", "cloud_water_content", "cloud_effective_radius",
           "aerosol_concentration", "aerosol_effective_radius", "wind_speed", "wind_direction", "salinity")
_ADJOINT = ("temperature_ad", "absorber_ad", "water_temperature_ad", "cloud_water_content_ad",
            "cloud_effective_radius_ad", "aerosol_concentration_ad", "aerosol_effective_radius_ad",
            "wind_speed_ad", "wind_direction_ad", "salinity_ad")


class CRTMBrightnessTemperature(torch.autograd.Function):
    @staticmethod
    def forward(ctx, crtm_ctx, profiles, options, emissivity, *state):
        precision, n_streams, overlap, max_bytes, chunk, progress = options
        prof = _with_state(profiles, state)
        e = None if emissivity is None else emissivity.detach().to(torch.float64).contiguous()
        M = prof.n_profiles
        if not chunk or chunk >= M:
            bt = crtm_ctx.forward(prof, e, precision=precision, n_streams=n_streams,
                                  overlap=overlap)["brightness_temperature"]
        else:   # profile chunks bound the scattering work space of one call
            bt = torch.empty(M, crtm_ctx.n_channels, dtype=torch.float64, device=prof.temperature.device)
            per_profile = not (isinstance(n_streams, int) or getattr(n_streams, "ndim", 1) == 0)
            for m0 in range(0, M, chunk):
                m1 = min(M, m0 + chunk)
                ns = n_streams[m0:m1] if per_profile else n_streams
                bt[m0:m1] = crtm_ctx.forward(api.profile_slice(prof, m0, m1),
                                             None if e is None else e[m0:m1].contiguous(),
                                             precision=precision, n_streams=ns,
                                             overlap=overlap)["brightness_temperature"]
                if progress:
                    progress(f"forward profiles {m0}-{m1} of {M} issued")
        ctx.has_emis = emissivity is not None
        ctx.save_for_backward(*([emissivity] if ctx.has_emis else []), *[s for s in state if s is not None])
        ctx.present = tuple(s is not None for s in state)
        ctx.crtm = (crtm_ctx, profiles, options)
        return bt

    @staticmethod
    def backward(ctx, grad_bt):
        saved = list(ctx.saved_tensors)
        emissivity = saved.pop(0) if ctx.has_emis else None
        state = [saved.pop(0) if p else None for p in ctx.present]
        crtm_ctx, profiles, (precision, n_streams, overlap, max_bytes, chunk, progress) = ctx.crtm
        prof = _with_state(profiles, state)
        seed = grad_bt.detach().to(torch.float64).contiguous()
        e = None if emissivity is None else emissivity.detach().to(torch.float64).contiguous()
        _, ad = crtm_ctx.adjoint(prof, e, seed,
                                 precision=precision, n_streams=n_streams, overlap=overlap,
                                 max_work_bytes=max_bytes, progress=progress)
        grads = [None if s is None else ad[k].to(s.dtype) for s, k in zip(state, _ADJOINT)]
        ge = None if emissivity is None else ad["surface_emissivity_ad"].to(emissivity.dtype)
        return (None, None, None, ge, *grads)


def _with_state(profiles, state):
    repl = {f: s.detach().to(torch.float64).contiguous() for f, s in zip(_FIELDS, state) if s is not None}
    prof = dataclasses.replace(profiles, **repl)
    prof.n_layers_host = profiles.n_layers_host
    return prof


def crtm_bt(crtm_ctx, profiles, temperature, absorber, water_temperature, emissivity,
            precision=api.PREC_FP32_SPLIT, cloud_water_content=None, cloud_effective_radius=None,
            aerosol_concentration=None, aerosol_effective_radius=None, n_streams=0, overlap=0,
            max_work_bytes=16 << 30, forward_chunk=0, progress=None, wind_speed=None, wind_direction=None,
            salinity=None):
    """Differentiable CRTM brightness temperatures [M, L].  forward_chunk > 0
    runs the forward model in chunks of that many profiles; the backward is
    crtm_adjoint, whose scattering work space per call max_work_bytes
    bounds.  progress(message) is called after each chunk is issued."""
    if not temperature.is_cuda:
        raise api.CRTMError("crtm_bt runs on the GPU: pass device tensors (api.device_profiles)")
    if emissivity is not None and any(x is not None for x in (wind_speed, wind_direction, salinity)):
        raise api.CRTMError("wind_speed / wind_direction / salinity feed the computed surface optics: "
                            "pass emissivity=None")
    state = (temperature, absorber, water_temperature, cloud_water_content, cloud_effective_radius,
             aerosol_concentration, aerosol_effective_radius, wind_speed, wind_direction, salinity)
    return CRTMBrightnessTemperature.apply(crtm_ctx, profiles,
                                           (precision, n_streams, overlap, max_work_bytes, forward_chunk, progress),
                                           emissivity, *state)
