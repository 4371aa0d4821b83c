"""torch.autograd.Function over the C-ABI: brightness temperatures as a
differentiable function of the atmospheric state.

Forward is CRTM_Forward (src/CRTM_Forward_Module.f90:229-245); backward is
the adjoint of the same model, obtained from CRTM_K_Matrix
(src/CRTM_K_Matrix_Module.f90:300-325) seeded with the incoming gradient as
RTSolution_K%Brightness_Temperature and summed over channels -- the
channel-summed adjoint that CRTM_Adjoint (src/CRTM_Adjoint_Module.f90:299)
returns.  Both run on the GPU in device memory on torch's current stream.

    bt = crtm_bt(ctx, profiles, temperature, absorber, water_temperature,
                 emissivity, precision=api.PREC_FP32_SPLIT)
    J = (((y - bt) / sigma) ** 2).sum(); J.backward()

`profiles` is a device batch from api.device_profiles (it supplies the
pressures, layer counts, coverages and geometry); the differentiable inputs
replace its temperature [NT], absorber [n_absorbers, NT] and water
temperature [M].  Gradients flow to those and to the emissivity [M, L].
"""
from __future__ import annotations

import dataclasses

import torch

from . import api


class CRTMBrightnessTemperature(torch.autograd.Function):
    @staticmethod
    def forward(ctx, temperature, absorber, water_temperature, emissivity, crtm_ctx, profiles, precision):
        prof = _with_state(profiles, temperature, absorber, water_temperature)
        rts = crtm_ctx.forward(prof, emissivity, precision=precision)
        ctx.save_for_backward(temperature, absorber, water_temperature, emissivity)
        ctx.crtm = (crtm_ctx, profiles, precision)
        return rts["brightness_temperature"]

    @staticmethod
    def backward(ctx, grad_bt):
        temperature, absorber, water_temperature, emissivity = ctx.saved_tensors
        crtm_ctx, profiles, precision = ctx.crtm
        prof = _with_state(profiles, temperature, absorber, water_temperature)
        seed = grad_bt.detach().to(torch.float64).contiguous()
        _, k = crtm_ctx.k_matrix(prof, emissivity, bt_k=seed, precision=precision, jacobian_f32=False)
        g_t = k["temperature_k"].sum(dim=1)
        g_a = k["absorber_k"].sum(dim=2)
        g_ts = k["water_temperature_k"].sum(dim=1)
        g_e = k["surface_emissivity_k"]
        return (g_t.to(temperature.dtype), g_a.to(absorber.dtype), g_ts.to(water_temperature.dtype),
                g_e.to(emissivity.dtype), None, None, None)


def _with_state(profiles, temperature, absorber, water_temperature):
    prof = dataclasses.replace(profiles, temperature=temperature.detach().to(torch.float64).contiguous(),
                               absorber=absorber.detach().to(torch.float64).contiguous(),
                               water_temperature=water_temperature.detach().to(torch.float64).contiguous())
    prof.n_layers_host = profiles.n_layers_host
    return prof


def crtm_bt(crtm_ctx, profiles, temperature, absorber, water_temperature, emissivity,
            precision=api.PREC_FP32_SPLIT):
    """Differentiable CRTM brightness temperatures [M, L]."""
    if not temperature.is_cuda:
        raise api.CRTMError("crtm_bt runs on the GPU: pass device tensors (api.device_profiles)")
    return CRTMBrightnessTemperature.apply(temperature, absorber, water_temperature, emissivity, crtm_ctx,
                                           profiles, precision)
