"""Seeded synthetic sensors and atmospheric profiles.

No CRTM coefficient data can be fetched offline, so every run (tests, bench,
smoke) writes synthetic SpcCoeff / ODPS TauCoeff files in the reference binary
formats (crtmv3_amd/coeffio.py) and reads them back through the native loader,
exactly as a user would load real fix/ files.

Sensor layouts ("like" = same channel count, spectral range and ODPS group):
  amsua_syn   15 ch  MW  ODPS group 3   (AMSU-A-like)
  atms_syn    22 ch  MW  ODPS group 3   (ATMS-like)
  cris_syn  2211 ch  IR  ODPS group 1   (CrIS full-spectral-resolution-like, OPTRAN WLO)
  iasi_syn  8461 ch  IR  ODPS group 1   (IASI-like, OPTRAN WLO)
  hirs_syn    19 ch  IR  ODPS group 2   (small group-2 test sensor)

ODPS reference grid: the 101-level model pressure grid of
src/Atmosphere/CRTM_Model_Profiles.f90:80-97 with the US-standard climatology
as the reference profile.  Regression coefficients are built so that, at the
reference state, each component's layer optical depth equals a smooth target
profile; perturbations then move OD through every predictor the component
uses.

Profiles follow the construction of the benchmark configurations: 100 layers
on pressure levels log-spaced 0.005-1100 hPa, US-standard temperature plus a
vertically correlated N(0, 5 K) field (1 km correlation length in
hypsometric height), log-normal water vapour (sigma 0.3 in ln), ozone x
U(0.8, 1.2), ocean surface at U(270, 305) K, zenith U(0, 60 deg).

Clouds (add_clouds): liquid, ice, rain and snow water contents log-normal
around 0.05 g/m^3 (sigma 1 in ln) in 0-10 contiguous layers per type, liquid
radii U(5, 20) um, ice radii U(20, 80) um.  The synthetic CloudCoeff tables
(make_cloud_coeff) are smooth Rayleigh-to-Mie-like functions of frequency,
radius and temperature with Henyey-Greenstein Legendre coefficients.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import coeffio

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "model_profiles.json")

# Planck radiation constants for radiance in mW/(m^2 sr cm^-1), wavenumber in cm^-1
PLANCK_C1 = 1.191042972e-05
PLANCK_C2 = 1.4387752
COSMIC_T = 2.7253
GHZ_TO_INV_CM = 1.0 / 29.9792458

# CRTM absorber ids (src/Atmosphere/CRTM_Atmosphere_Define.f90)
H2O_ID, CO2_ID, O3_ID, N2O_ID, CO_ID, CH4_ID = 1, 2, 3, 4, 5, 6

# ODPS group layouts (src/AtmAbsorption/ODPS/ODPS_Predictor.f90:85-207)
GROUPS = {
    1: dict(n_pred=[7, 18, 7, 11, 11, 14, 10, 11], comp_id=[7, 101, 15, 114, 121, 120, 119, 118],
            abs_id=[H2O_ID, O3_ID, CO2_ID, N2O_ID, CO_ID, CH4_ID]),
    2: dict(n_pred=[7, 15, 7, 11, 10], comp_id=[20, 101, 15, 114, 121],
            abs_id=[H2O_ID, O3_ID, CO2_ID]),
    3: dict(n_pred=[7, 14], comp_id=[113, 12], abs_id=[H2O_ID]),
}

DT_SCALE = 5.0   # typical |T - Tref| used to scale coefficients of DT-type predictors


def climatology() -> dict:
    with open(_DATA) as f:
        return json.load(f)


def _layer_mean(levels_p, level_vals):
    """Layer value as the ln(p)-weighted mean of the bounding level values."""
    return 0.5 * (np.asarray(level_vals)[1:] + np.asarray(level_vals)[:-1])


def reference_profile():
    """ODPS reference grid: 101 model levels, US-standard layer profile."""
    clim = climatology()
    p = np.array(clim["level_pressure"])
    us = clim["models"].index("US_STANDARD_ATMOSPHERE")
    t = _layer_mean(p, clim["temperature"][us])
    ab = np.stack([_layer_mean(p, clim["absorber"][us][i]) for i in range(6)], axis=1)
    lay_p = (p[1:] - p[:-1]) / np.log(p[1:] / p[:-1])
    # absorber columns in CRTM absorber-id order 1..6 -> clim order H2O,O3,CO2,N2O,CO,CH4
    by_id = {H2O_ID: ab[:, 0], O3_ID: ab[:, 1], CO2_ID: ab[:, 2], N2O_ID: ab[:, 3],
             CO_ID: ab[:, 4], CH4_ID: ab[:, 5]}
    return p, lay_p, t, by_id


# ----------------------------------------------------------------------------
# Nominal predictor magnitudes at the reference state (T = Tref, all absorber
# ratios and integrated ratios = 1), as functions of the secant s.  A value of
# the form ("dt", v) marks a predictor that vanishes at the reference state
# because it carries a factor DT or DT*|DT|; those get small signed weights.
# ----------------------------------------------------------------------------
def _nominal(group: int, comp: int, s: float):
    r = np.sqrt(s)
    q = s ** 0.25
    if group in (1, 2):
        tab = {
            0: [s, s, s, 1, s * s, 1, 1],
            1: [s, ("dt", s), s * s, ("dt2", s), q, s ** 3, r, ("dt", r), s ** 4, 1, r, s * s, s, s, s,
                s, ("dt", s * s), s],
            2: [s] * 7,
            3: [s, ("dt", s), s, s * s, s, s * r, ("dt", r), r, r, s, s * s],
            4: [s, s, 1, 1, s, s, s, s * s, 1, s, s],
            5: [("dt", s), r, ("dt", 1), q, s, s, s, s * s, 1, r, s, s, s, s * s],
            6: [s, ("dt", s), q, ("dt", r), s * s, r, ("dt2", s), s, r, s],
            7: [("dt", s), r, s * s, s, ("dt", 1), s, s * s, q, 1, s, r],
        }
    else:
        tab = {
            0: [s, s, s, 1, s * s, 1, 1],
            1: [s, s, s, s, s, s, s, ("dt", s), s * s, s, s, s ** 3, s ** 4, 1],
        }
    return tab[comp]


@dataclasses.dataclass
class SensorDef:
    sensor_id: str
    sensor_type: int
    group: int
    frequency: np.ndarray       # GHz (MW) or 0
    wavenumber: np.ndarray      # cm^-1
    strength: np.ndarray        # [J, L] column OD of each component at the reference state
    optran_fraction: float = 0.0


def _smooth_noise(rng, n, corr, sigma):
    x = rng.standard_normal(n + 4 * corr)
    k = np.exp(-0.5 * (np.arange(-2 * corr, 2 * corr + 1) / max(corr, 1)) ** 2)
    y = np.convolve(x, k / np.sqrt((k * k).sum()), mode="same")[2 * corr:2 * corr + n]
    return sigma * y


def _bump(nu, c, w):
    return np.exp(-0.5 * ((nu - c) / w) ** 2)


def _ir_strengths(nu, group, rng):
    L = nu.size
    J = len(GROUPS[group]["n_pred"])
    lines = np.exp(rng.standard_normal((J, L)) * 0.7)
    A = np.zeros((J, L))
    A[0] = (0.02 + 0.06 * _bump(nu, 1600, 150) + 0.05 * _bump(nu, 2350, 60)) * lines[0]
    A[1] = (0.02 + 25.0 * _bump(nu, 1560, 130) + 0.3 * _bump(nu, 1300, 60)) * lines[1]
    A[2] = (0.03 + 0.25 * _bump(nu, 1000, 250) + 0.1 * _bump(nu, 1500, 200)) * lines[2]
    A[3] = (0.002 + 4.0 * _bump(nu, 1040, 25) + 0.3 * _bump(nu, 2110, 20)) * lines[3]
    A[4] = (0.01 + 60.0 * _bump(nu, 667, 25) + 8.0 * _bump(nu, 720, 35) + 40.0 * _bump(nu, 2330, 40)
            + 0.5 * _bump(nu, 960, 15)) * lines[4]
    if J > 5:
        A[5] = (0.002 + 1.2 * _bump(nu, 1280, 25) + 2.0 * _bump(nu, 2220, 25)) * lines[5]
        A[6] = (0.001 + 1.5 * _bump(nu, 2150, 30)) * lines[6]
        A[7] = (0.002 + 1.5 * _bump(nu, 1305, 35) + 0.3 * _bump(nu, 3000, 50)) * lines[7]
    return A


# sensors whose SpcCoeff files carry correction data: the channel set of a
# base sensor plus antenna-correction (microwave) or non-LTE-correction
# (infrared) blocks
_DERIVED = {"amsua_ac_syn": "amsua_syn", "cris_nlte_syn": "cris_small_syn"}
N_FOVS_AC = 30   # fields of view of the antenna correction (AMSU-A has 30)


def sensor_def(name: str, seed: int = 20261203) -> SensorDef:
    rng = np.random.default_rng(seed + sum(map(ord, name)))
    if name in _DERIVED:
        return dataclasses.replace(sensor_def(_DERIVED[name], seed), sensor_id=name)
    if name == "amsua_syn":
        f = np.array([23.8, 31.4, 50.3, 52.8, 53.596, 54.4, 54.94, 55.5,
                      57.29, 57.29, 57.29, 57.29, 57.29, 57.29, 89.0])
        peak = np.array([0, 0, 1000, 900, 600, 400, 250, 150, 90, 50, 25, 10, 5, 2.5, 0.0])
        dry = np.where(peak > 0, 1100.0 / np.maximum(peak, 1e-9), [0.05, 0.03] + [0] * 12 + [0.1])
        wet = np.array([0.35, 0.12, 0.06, 0.04, 0.02, 0.01, 0.005, 0.003, 0.002, 0.002, 0.002, 0.002,
                        0.002, 0.002, 0.5])
        return SensorDef(name, coeffio.MICROWAVE_SENSOR, 3, f, f * GHZ_TO_INV_CM, np.stack([dry, wet]))
    if name == "atms_syn":
        f = np.array([23.8, 31.4, 50.3, 51.76, 52.8, 53.596, 54.4, 54.94, 55.5,
                      57.29, 57.29, 57.29, 57.29, 57.29, 57.29, 88.2, 165.5,
                      183.31 + 7.0, 183.31 + 4.5, 183.31 + 3.0, 183.31 + 1.8, 183.31 + 1.0])
        peak = np.array([0, 0, 1000, 950, 900, 600, 400, 250, 150, 90, 50, 25, 10, 5, 2.5,
                         0, 0, 0, 0, 0, 0, 0], float)
        dry = np.where(peak > 0, 1100.0 / np.maximum(peak, 1e-9), 0.05)
        wet = np.array([0.35, 0.12, 0.06, 0.05, 0.04, 0.02, 0.01, 0.005, 0.003, 0.002, 0.002, 0.002,
                        0.002, 0.002, 0.002, 0.5, 1.5, 4.0, 8.0, 16.0, 30.0, 60.0])
        return SensorDef(name, coeffio.MICROWAVE_SENSOR, 3, f, f * GHZ_TO_INV_CM, np.stack([dry, wet]))
    if name == "cris_syn":
        nu = np.concatenate([650 + 0.625 * np.arange(713), 1210 + 0.625 * np.arange(865),
                             2155 + 0.625 * np.arange(633)])
        return SensorDef(name, coeffio.INFRARED_SENSOR, 1, np.zeros_like(nu), nu,
                         _ir_strengths(nu, 1, rng), optran_fraction=0.35)
    if name == "iasi_syn":
        nu = 645 + 0.25 * np.arange(8461)
        return SensorDef(name, coeffio.INFRARED_SENSOR, 1, np.zeros_like(nu), nu,
                         _ir_strengths(nu, 1, rng), optran_fraction=0.35)
    if name == "hirs_syn":
        nu = np.array([669.0, 680, 690, 703, 716, 733, 749, 900, 1030, 1345, 1365, 1533,
                       2188, 2210, 2235, 2245, 2420, 2515, 2660])
        return SensorDef(name, coeffio.INFRARED_SENSOR, 2, np.zeros_like(nu), nu,
                         _ir_strengths(nu, 2, rng), optran_fraction=0.0)
    if name == "cris_small_syn":
        # 96-channel subset spanning all three CrIS bands; fast for CPU oracle tests
        full = np.concatenate([650 + 0.625 * np.arange(713), 1210 + 0.625 * np.arange(865),
                               2155 + 0.625 * np.arange(633)])
        nu = full[np.linspace(0, full.size - 1, 96).astype(int)]
        return SensorDef(name, coeffio.INFRARED_SENSOR, 1, np.zeros_like(nu), nu,
                         _ir_strengths(nu, 1, rng), optran_fraction=0.5)
    raise KeyError(name)


SENSORS = ("amsua_syn", "atms_syn", "cris_syn", "iasi_syn", "hirs_syn", "cris_small_syn", "amsua_ac_syn",
           "cris_nlte_syn")


def _layer_shapes(group, lvl_p, lay_p, ref_abs):
    """Normalised vertical distribution of each component's reference OD."""
    dp = np.diff(lvl_p)
    x = lay_p / 1100.0
    q = ref_abs[H2O_ID]
    if group == 3:
        shapes = [dp * (0.5 + 0.5 * x), q * dp * x ** 0.3]
    else:
        shapes = [dp * x ** 0.5, q * dp * x ** 0.3, q * q * dp * x, ref_abs[O3_ID] * dp,
                  dp * x ** 0.7]
        if group == 1:
            shapes += [ref_abs[N2O_ID] * dp, ref_abs[CO_ID] * dp, ref_abs[CH4_ID] * dp]
    return [s / s.sum() for s in shapes]


MAX_LAYER_OD = 12.0   # stays below the ODPS clamp MAX_OD = 20


def make_tau_coeff(sd: SensorDef, seed: int = 20261203) -> coeffio.ODPS:
    rng = np.random.default_rng(seed + 7 + sum(map(ord, sd.sensor_id)))
    grp = GROUPS[sd.group]
    lvl_p, lay_p, ref_t, ref_by_id = reference_profile()
    Ko = lay_p.size
    J = len(grp["n_pred"])
    L = sd.wavenumber.size
    abs_ids = grp["abs_id"]
    ref_abs = np.stack([ref_by_id[a] for a in abs_ids], axis=1)
    shapes = _layer_shapes(sd.group, lvl_p, lay_p, ref_by_id)
    s_nom = 1.25
    n_pred = np.zeros((J, L), np.int32)
    pos = np.zeros((J, L), np.int32)
    chunks = []
    nc = 0
    for l in range(L):
        for j in range(J):
            A = sd.strength[j, l]
            if j == 0:
                npj = grp["n_pred"][0]
            elif A < 0.004:
                npj = 0
            else:
                npj = int(rng.integers((grp["n_pred"][j] + 1) // 2, grp["n_pred"][j] + 1))
            n_pred[j, l] = npj
            if npj == 0:
                pos[j, l] = nc + 1
                continue
            A = min(A, MAX_LAYER_OD / shapes[j].max())
            tau = A * shapes[j]                      # [Ko]
            nom = _nominal(sd.group, j, s_nom)[:npj]
            w = rng.dirichlet(np.ones(npj))
            coef = np.zeros((npj, Ko))
            for i, v in enumerate(nom):
                if isinstance(v, tuple):
                    scale = DT_SCALE if v[0] == "dt" else DT_SCALE * DT_SCALE
                    coef[i] = 0.05 * rng.standard_normal() * tau / (scale * v[1])
                else:
                    coef[i] = w[i] * tau / v
            pos[j, l] = nc + 1
            chunks.append(coef.astype(np.float32).ravel())   # (i-1)*Ko + k
            nc += npj * Ko
    c = np.concatenate(chunks) if chunks else np.zeros(0, np.float32)
    t = coeffio.ODPS(
        group_index=sd.group, sensor_id=sd.sensor_id, wmo_satellite_id=1023, wmo_sensor_id=2047,
        sensor_type=sd.sensor_type, sensor_channel=np.arange(1, L + 1, dtype=np.int32),
        component_id=np.array(grp["comp_id"], np.int32), absorber_id=np.array(abs_ids, np.int32),
        ref_level_pressure=lvl_p, ref_pressure=lay_p, ref_temperature=ref_t, ref_absorber=ref_abs,
        min_absorber=ref_abs * 1e-3, max_absorber=ref_abs * 1e3, n_predictors=n_pred,
        pos_index=pos, c=c)
    if sd.optran_fraction > 0:
        _add_optran(t, sd, rng, shapes[1], lvl_p, lay_p, ref_t, ref_by_id[H2O_ID])
    return t


def _add_optran(t, sd, rng, wlo_shape, lvl_p, lay_p, ref_t, ref_q):
    """OPTRAN water-vapour-line block for a subset of strong WLO channels.

    For OPTRAN channels the WLO layer OD is chi*dA with
    ln(chi) = b0(Ap) + sum_i b_i(Ap) OX_i   (ODPS_AtmAbsorption.f90:664-727).
    b0 is a least-squares polynomial in Ap1 fitted so that chi*dA reproduces
    the channel's WLO target OD at the reference state.
    """
    L = t.n_channels
    order_max = 4
    g_inv = 1.0 / 980.665
    s_nom = 1.25
    dA = g_inv * np.diff(lvl_p) * ref_q * s_nom
    intv = np.cumsum(dA)
    aveA = 0.5 * (intv + np.concatenate([[0.0], intv[:-1]]))
    alpha, c1, c2 = 9.0, 1.0e-9, 0.0
    ap1 = np.log((aveA - c2) / c1) / alpha
    ox_ref = {1: ref_t, 2: lay_p, 10: ref_q, 14: np.full_like(lay_p, s_nom)}
    osig = np.zeros(L, np.int32)
    order = np.zeros(L, np.int32)
    op_index = np.zeros((coeffio.N_PREDICTOR_USED_OPTRAN + 1, L), np.int32)
    opos = np.ones(L, np.int32)
    oc = []
    n = 0
    for l in range(L):
        A = sd.strength[1, l]
        if t.n_predictors[1, l] == 0 or A < 0.5 or rng.random() > sd.optran_fraction:
            continue
        A = min(A, MAX_LAYER_OD / wlo_shape.max())
        tau = A * wlo_shape
        preds = [1, 2, 10, 14][: int(rng.integers(1, 5))]
        b = {1: 2.0e-3 * rng.standard_normal(), 2: 2.0e-4 * rng.standard_normal(),
             10: 0.02 * rng.standard_normal(), 14: 0.1 * rng.standard_normal()}
        resid = np.log(tau / dA) - sum(b[i] * ox_ref[i] for i in preds)
        coef0 = np.polynomial.polynomial.polyfit(ap1, resid, order_max)
        osig[l] = 1
        order[l] = order_max
        op_index[0, l] = len(preds)
        op_index[1:1 + len(preds), l] = preds
        opos[l] = n + 1
        block = [coef0] + [np.r_[b[i], 1.0e-4 * rng.standard_normal(order_max)] for i in preds]
        for row in block:
            oc.append(row)
            n += order_max + 1
    if n == 0:
        return
    t.osignificance, t.order, t.op_index, t.opos_index = osig, order, op_index, opos
    t.oc = np.concatenate(oc).astype(np.float64)
    t.alpha, t.alpha_c1, t.alpha_c2 = alpha, c1, c2
    t.ocomponent_index = 2


# polarisation flags (src/InstrumentInfo/SensorInfo/SensorInfo_Parameters.f90:92-108)
UNPOLARIZED, SECOND_STOKES_COMPONENT, THIRD_STOKES_COMPONENT, FOURTH_STOKES_COMPONENT = 1, 2, 3, 4
VL_POLARIZATION, HL_POLARIZATION, PLUS45L_POLARIZATION, MINUS45L_POLARIZATION = 5, 6, 7, 8
VL_MIXED_POLARIZATION, HL_MIXED_POLARIZATION, RC_POLARIZATION, LC_POLARIZATION = 9, 10, 11, 12
CONST_MIXED_POLARIZATION, PRA_POLARIZATION = 13, 14
# ATMS-like: quasi-V at 23.8/31.4/88.2 GHz, quasi-H elsewhere; the AMSU-A-like
# set cycles through every flag the surface optics combine
_MW_POLARIZATION = {
    "atms_syn": (9, 9, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10, 9, 10, 10, 10, 10, 10, 10),
    "amsua_syn": (9, 10, 5, 6, 1, 2, 3, 4, 7, 8, 11, 12, 13, 14, 9),
}
# bumped when the content of the synthetic sensor files changes
_SENSOR_FILE_REV = 2


def make_spc_coeff(sd: SensorDef, seed: int = 20261203) -> coeffio.SpcCoeff:
    rng = np.random.default_rng(seed + 11 + sum(map(ord, sd.sensor_id)))
    L = sd.wavenumber.size
    nu = sd.wavenumber
    c1 = PLANCK_C1 * nu ** 3
    c2 = PLANCK_C2 * nu
    if sd.sensor_type == coeffio.MICROWAVE_SENSOR:
        bc1 = np.zeros(L)
        bc2 = np.ones(L)
    else:
        bc1 = rng.uniform(-0.05, 0.05, L)
        bc2 = 1.0 + rng.uniform(-2e-4, 2e-4, L)
    if sd.sensor_type == coeffio.MICROWAVE_SENSOR:
        cosmic = c1 / (np.exp(c2 / (bc1 + bc2 * COSMIC_T)) - 1.0)
    else:
        cosmic = np.zeros(L)
    pol, pol_angle = np.ones(L, np.int32), np.zeros(L)
    if sd.sensor_type == coeffio.MICROWAVE_SENSOR:
        pattern = _MW_POLARIZATION.get(sd.sensor_id, (UNPOLARIZED,))
        pol = np.resize(np.asarray(pattern, np.int32), L)
        pol_angle = np.where(np.isin(pol, (CONST_MIXED_POLARIZATION, PRA_POLARIZATION)),
                             20.0 + 3.0 * np.arange(L), 0.0)
    sc = coeffio.SpcCoeff(
        sensor_id=sd.sensor_id, sensor_type=sd.sensor_type, wmo_satellite_id=1023,
        wmo_sensor_id=2047, sensor_channel=np.arange(1, L + 1, dtype=np.int32),
        polarization=pol, pol_angle=pol_angle, channel_flag=np.zeros(L, np.int32),
        frequency=sd.frequency, wavenumber=nu, planck_c1=c1, planck_c2=c2, band_c1=bc1,
        band_c2=bc2, cosmic_background_radiance=cosmic, solar_irradiance=np.zeros(L))
    if sd.sensor_id == "amsua_ac_syn":
        sc.ac = make_ac_coeff(sc, N_FOVS_AC, rng)
    if sd.sensor_id == "cris_nlte_syn":
        sc.nc = make_nlte_coeff(sc, rng)
    return sc


def make_ac_coeff(sc: coeffio.SpcCoeff, n_fovs: int, rng) -> coeffio.ACCoeff:
    """Antenna efficiencies of a cross-track scanner: the space and platform
    fractions grow towards the scan edges, the Earth fraction takes the rest."""
    L = sc.n_channels
    edge = (np.abs(np.arange(n_fovs) - (n_fovs - 1) / 2) / ((n_fovs - 1) / 2)) ** 2
    a_space = (0.004 + 0.012 * edge)[:, None] * rng.uniform(0.6, 1.4, (1, L))
    a_platform = (0.001 + 0.004 * edge)[:, None] * rng.uniform(0.5, 1.5, (1, L))
    a_earth = 1.0 - a_space - a_platform + rng.uniform(-5e-4, 5e-4, (n_fovs, L))
    return coeffio.ACCoeff(sc.sensor_id, sc.wmo_satellite_id, sc.wmo_sensor_id, sc.sensor_channel.copy(),
                           a_earth, a_space, a_platform)


def make_nlte_coeff(sc: coeffio.SpcCoeff, rng) -> coeffio.NLTECoeff:
    """Non-LTE radiance excess of the 4.3 um band (channels above 2200 cm^-1):
    daytime excess growing with the sun's height and with the mean
    temperatures of the 0.005-0.2 hPa and 0.2-52 hPa layers (the pressure
    levels of src/NLTE/NLTE_Parameters.f90)."""
    L = sc.n_channels
    nlte = np.nonzero(sc.wavenumber > 2200.0)[0]
    n = nlte.size
    ssen = 1.0 / np.cos(np.deg2rad(np.array([0.0, 20.0, 35.0, 45.0, 55.0, 65.0])))
    ssol = 1.0 / np.cos(np.deg2rad(np.array([0.0, 30.0, 50.0, 65.0, 75.0, 80.0, 85.0])))
    c = np.zeros((3, ssen.size, ssol.size, n))
    sun = 1.0 / ssol                                   # cosine of the solar zenith
    view = 1.0 / ssen
    band = np.exp(-((sc.wavenumber[nlte] - 2330.0) / 60.0) ** 2)   # strongest in the band core
    scale = 0.1 * (0.3 + band) * rng.uniform(0.8, 1.2, n)
    for i, (t0, w) in enumerate(((210.0, 1.0), (240.0, 0.5))):
        c[1 + i] = scale * w * 1e-3 * (sun[None, :, None] * (0.7 + 0.3 * view[:, None, None]))
        c[0] -= c[1 + i] * (t0 - 40.0)                 # excess at the layer temperature t0: 40 K x slope
    c_index = np.zeros(L, np.int32)
    c_index[nlte] = np.arange(1, n + 1, dtype=np.int32)
    return coeffio.NLTECoeff(sc.sensor_id, sc.wmo_satellite_id, sc.wmo_sensor_id, sc.sensor_channel.copy(),
                             upper_plevel=np.array([0.005, 0.2]), lower_plevel=np.array([0.2, 52.0]),
                             min_tm=np.array([170.0, 200.0]), max_tm=np.array([300.0, 290.0]),
                             mean_tm=np.array([215.0, 245.0]), nlte_channel=sc.sensor_channel[nlte].copy(),
                             secant_sensor_zenith=ssen, secant_solar_zenith=ssol, c_index=c_index, c=c)


def write_sensor(dirpath: str, name: str, seed: int = 20261203) -> str:
    """Write <name>.SpcCoeff.bin and <name>.TauCoeff.bin; returns dirpath."""
    os.makedirs(dirpath, exist_ok=True)
    spc = os.path.join(dirpath, f"{name}.SpcCoeff.bin")
    tau = os.path.join(dirpath, f"{name}.TauCoeff.bin")
    stamp = os.path.join(dirpath, f"{name}.seed")
    if os.path.exists(spc) and os.path.exists(tau) and os.path.exists(stamp) \
            and open(stamp).read().strip() == f"{seed}:{_SENSOR_FILE_REV}":
        return dirpath
    sd = sensor_def(name, seed)
    coeffio.write_spccoeff(spc + ".tmp", make_spc_coeff(sd, seed))
    coeffio.write_odps(tau + ".tmp", make_tau_coeff(sd, seed))
    os.replace(spc + ".tmp", spc)
    os.replace(tau + ".tmp", tau)
    with open(stamp, "w") as f:
        f.write(f"{seed}:{_SENSOR_FILE_REV}")
    return dirpath


def default_coeff_dir() -> str:
    d = os.environ.get("CRTM_SYNTH_DIR")
    if d:
        return d
    return os.path.join(os.environ.get("TMPDIR", "/tmp"), "crtmv3_amd_coeffs")


# ----------------------------------------------------------------------------
# Profiles
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class ProfileSet:
    """Flat, layer-major profile batch (top-down layers, CRTM ordering).

    Per-profile layer counts may differ; layer_offset[m] indexes the first
    layer of profile m in the per-layer arrays, level arrays use
    layer_offset[m] + m.
    """
    n_layers: np.ndarray                 # int32 [M]
    layer_offset: np.ndarray             # int64 [M+1]
    absorber_id: np.ndarray              # int32 [nabs]
    absorber_units: np.ndarray           # int32 [nabs]
    level_pressure: np.ndarray           # [sum(K)+M]
    pressure: np.ndarray                 # [sum K]
    temperature: np.ndarray              # [sum K]
    absorber: np.ndarray                 # [nabs, sum K]
    # surface
    land_coverage: np.ndarray
    water_coverage: np.ndarray
    snow_coverage: np.ndarray
    ice_coverage: np.ndarray
    land_temperature: np.ndarray
    water_temperature: np.ndarray
    snow_temperature: np.ndarray
    ice_temperature: np.ndarray
    # geometry
    sensor_zenith: np.ndarray
    sensor_scan: np.ndarray
    source_zenith: np.ndarray
    # clouds: per profile up to n_clouds slots; water_content[n_clouds, sum K]
    cloud_type: Optional[np.ndarray] = None          # int32 [M, nc]
    cloud_water_content: Optional[np.ndarray] = None # kg/m^2 [nc, sum K]
    cloud_effective_radius: Optional[np.ndarray] = None
    cloud_fraction: Optional[np.ndarray] = None      # [sum K]
    aerosol_type: Optional[np.ndarray] = None        # int32 [M, na]
    aerosol_concentration: Optional[np.ndarray] = None
    aerosol_effective_radius: Optional[np.ndarray] = None
    climatology: Optional[np.ndarray] = None         # int32 [M], 1..6 (None: US standard)
    # water surface and view azimuth, inputs of the computed surface optics
    # (None: the reference defaults 5 m/s, 0 degrees, 33 ppt, azimuth off)
    wind_speed: Optional[np.ndarray] = None          # m/s [M]
    wind_direction: Optional[np.ndarray] = None      # degrees [M]
    salinity: Optional[np.ndarray] = None            # ppt [M]
    sensor_azimuth: Optional[np.ndarray] = None      # degrees [M]
    ifov: Optional[np.ndarray] = None                # int32 [M] Geometry%iFOV (None: 0, no antenna correction)

    @property
    def n_profiles(self) -> int:
        return int(len(self.n_layers))

    def profile(self, m: int) -> Dict[str, np.ndarray]:
        a, b = int(self.layer_offset[m]), int(self.layer_offset[m + 1])
        return dict(level_pressure=self.level_pressure[a + m:b + m + 1], pressure=self.pressure[a:b],
                    temperature=self.temperature[a:b], absorber=self.absorber[:, a:b])


def _log_levels(K):
    lp = np.exp(np.log(0.005) + np.arange(K + 1) / K * (np.log(1100.0) - np.log(0.005)))
    # exact end points: the top must be TOA_PRESSURE itself, not a rounding
    # above it (which would make CRTM add a vanishing climatology layer)
    lp[0], lp[-1] = 0.005, 1100.0
    return lp


def _gp_factor(lay_p, t_clim, corr_km=1.0):
    """Cholesky factor of the squared-exponential covariance in height."""
    H = 0.001 * 287.05 * t_clim / 9.80665         # scale height, km
    z = np.concatenate([[0.0], np.cumsum(-H[:-1] * np.diff(np.log(lay_p)))])
    z = z[-1] - z                                   # increasing upwards from the bottom layer
    d = z[:, None] - z[None, :]
    cov = np.exp(-0.5 * (d / corr_km) ** 2) + 1e-8 * np.eye(z.size)
    return np.linalg.cholesky(cov)


def make_profiles(n_profiles: int, n_layers: int | Sequence[int] = 100, seed: int = 20261203,
                  ozone: bool = True, sensor_zenith_max: float = 60.0) -> ProfileSet:
    rng = np.random.default_rng(seed)
    clim = climatology()
    us = clim["models"].index("US_STANDARD_ATMOSPHERE")
    cp = np.log(np.array(clim["level_pressure"]))
    if np.isscalar(n_layers):
        ks = np.full(n_profiles, int(n_layers), np.int32)
    else:
        ks = np.asarray(n_layers, np.int32)
        assert ks.size == n_profiles
    off = np.zeros(n_profiles + 1, np.int64)
    off[1:] = np.cumsum(ks)
    ntot = int(off[-1])
    lvl = np.zeros(ntot + n_profiles)
    lay = np.zeros(ntot)
    tt = np.zeros(ntot)
    nabs = 2 if ozone else 1
    ab = np.zeros((nabs, ntot))
    cache = {}
    for m in range(n_profiles):
        K = int(ks[m])
        if K not in cache:
            lp = _log_levels(K)
            p = (lp[1:] - lp[:-1]) / np.log(lp[1:] / lp[:-1])
            t0 = np.interp(np.log(p), cp, clim["temperature"][us])
            q0 = np.exp(np.interp(np.log(p), cp, np.log(clim["absorber"][us][0])))
            o0 = np.exp(np.interp(np.log(p), cp, np.log(clim["absorber"][us][1])))
            cache[K] = (lp, p, t0, q0, o0, _gp_factor(p, t0))
        lp, p, t0, q0, o0, Lc = cache[K]
        a, b = int(off[m]), int(off[m + 1])
        lvl[a + m:b + m + 1] = lp
        lay[a:b] = p
        tt[a:b] = t0 + 5.0 * (Lc @ rng.standard_normal(K))
        ab[0, a:b] = q0 * np.exp(0.3 * (Lc @ rng.standard_normal(K)))
        if ozone:
            ab[1, a:b] = o0 * rng.uniform(0.8, 1.2)
    M = n_profiles
    return ProfileSet(
        n_layers=ks, layer_offset=off,
        absorber_id=np.array([H2O_ID, O3_ID][:nabs], np.int32),
        absorber_units=np.array([3, 1][:nabs], np.int32),   # MASS_MIXING_RATIO (g/kg), VOLUME_MIXING_RATIO (ppmv)
        level_pressure=lvl, pressure=lay, temperature=tt, absorber=ab,
        land_coverage=np.zeros(M), water_coverage=np.ones(M), snow_coverage=np.zeros(M),
        ice_coverage=np.zeros(M), land_temperature=np.full(M, 280.0),
        water_temperature=rng.uniform(270.0, 305.0, M), snow_temperature=np.full(M, 260.0),
        ice_temperature=np.full(M, 260.0), sensor_zenith=rng.uniform(0.0, sensor_zenith_max, M),
        sensor_scan=np.zeros(M), source_zenith=np.full(M, 100.0))


def default_emissivity(sensor_type: int, n_channels: int) -> np.ndarray:
    return np.full(n_channels, 0.55 if sensor_type == coeffio.MICROWAVE_SENSOR else 0.98)


# ----------------------------------------------------------------------------
# CloudCoeff (Mie/TAMU-layout tables) and cloudy profiles
# ----------------------------------------------------------------------------
WATER_CLOUD, ICE_CLOUD, RAIN_CLOUD, SNOW_CLOUD, GRAUPEL_CLOUD, HAIL_CLOUD = 1, 2, 3, 4, 5, 6
CLOUDCOEFF_FILE = "CloudCoeff.bin"
# stream sets and their Legendre offsets (src/Coefficients/CloudCoeff/CloudCoeff_Define.f90:83-87)
STREAM_SETS = ((2, 0), (4, 0), (6, 5), (8, 12), (16, 21))
N_LEGENDRE_FILE = 38


def _hg_slots(g, n_slots, one_based=False):
    """Legendre coefficient slots for every stream set: slot offset+l holds
    (2l+1)/2 g^l, l = 1..n (g <= 0.5 keeps the delta truncation 1 - f w > 0)."""
    g = np.asarray(g)
    out = np.zeros(g.shape + (n_slots,))
    if not one_based:
        out[..., 0] = 0.5
    for n, off in STREAM_SETS:
        for l in range(1, n + 1):
            out[..., off + l - (1 if one_based else 0)] = (2 * l + 1) / 2.0 * g ** l
    return out


def make_cloud_coeff() -> coeffio.CloudCoeff:
    f_mw = np.array([10.65, 18.7, 23.8, 31.4, 36.5, 50.3, 52.8, 55.5, 57.29, 60.0, 70.0, 80.0, 89.0, 100.0,
                     118.75, 130.0, 150.0, 165.5, 175.0, 183.31, 190.0, 200.0, 220.0])
    r_mw = np.array([5.0, 10.0, 20.0, 40.0, 80.0, 150.0, 300.0, 500.0, 800.0, 1200.0, 2000.0])
    temp = np.array([243.15, 253.15, 263.15, 273.15, 283.15, 293.15, 303.15])
    dens = np.array([100.0, 400.0, 917.0])
    f_ir = np.geomspace(100.0, 3500.0, 48)
    r_ir = np.array([2.0, 5.0, 10.0, 20.0, 40.0, 60.0, 80.0, 100.0, 150.0, 200.0])
    nL1 = N_LEGENDRE_FILE + 1
    # microwave size parameter x = 2 pi r / lambda
    lam_mm = 299.792458 / f_mw
    x = 2.0 * np.pi * (r_mw[None, :] * 1e-3) / lam_mm[:, None]            # (f, r)
    x3 = x ** 3
    g_fr = 0.5 * x ** 2 / (1.0 + x ** 2)
    # liquid (f, r, T)
    kab_l = (0.3 * (f_mw / 23.8) ** 1.7 / (1.0 + (f_mw / 120.0) ** 1.2))[:, None, None] * \
        np.exp(0.012 * (283.15 - temp))[None, None, :]
    ksc_l = (8.0 * x3 / (1.0 + x3) * np.sqrt(f_mw / 23.8)[:, None])[:, :, None] * np.ones(temp.size)
    ke_l = kab_l + ksc_l
    w_l = ksc_l / ke_l
    g_l = np.broadcast_to(g_fr[:, :, None], ke_l.shape).copy()
    pc_l = _hg_slots(g_l, nL1)[..., None]                                   # (f, r, T, 0:38, 1)
    # solid (f, r, D)
    rho = dens / 917.0
    kab_s = (0.004 * (f_mw / 23.8))[:, None, None] * rho[None, None, :] * np.ones((1, r_mw.size, 1))
    ksc_s = (6.0 * x3 / (1.0 + x3) * np.sqrt(f_mw / 23.8)[:, None])[:, :, None] * rho[None, None, :] ** 0.7
    ke_s = kab_s + ksc_s
    w_s = ksc_s / ke_s
    g_s = np.broadcast_to(g_fr[:, :, None], ke_s.shape).copy()
    pc_s = _hg_slots(g_s, nL1)[..., None]
    # infrared (f, r, D+1): index 0 liquid, 1..3 frozen
    lam_um = 1.0e4 / f_ir
    xi = 2.0 * np.pi * r_ir[None, :] / lam_um[:, None]
    q = 2.0 * xi ** 2 / (1.0 + xi ** 2)
    rho_ir = np.concatenate([[1000.0], dens])
    ke_ir = (1500.0 * q / r_ir[None, :])[:, :, None] * (1000.0 / rho_ir)[None, None, :]
    w_ir = (0.55 + 0.35 * np.sin(f_ir / 700.0)[:, None] * xi / (1.0 + xi))[:, :, None] * \
        np.ones(rho_ir.size)
    g_ir = np.broadcast_to((0.5 * xi ** 2 / (1.0 + xi ** 2))[:, :, None], ke_ir.shape).copy()
    pc_ir = _hg_slots(g_ir, N_LEGENDRE_FILE, one_based=True)[..., None]
    return coeffio.CloudCoeff(
        frequency_mw=f_mw, frequency_ir=f_ir, reff_mw=r_mw, reff_ir=r_ir, temperature=temp, density_mw=dens,
        ke_l_mw=ke_l, w_l_mw=w_l, g_l_mw=g_l, pcoeff_l_mw=pc_l,
        ke_s_mw=ke_s, w_s_mw=w_s, g_s_mw=g_s, pcoeff_s_mw=pc_s,
        ke_ir=ke_ir, w_ir=w_ir, g_ir=g_ir, pcoeff_ir=pc_ir)


def write_cloud_coeff(dirpath: str) -> str:
    """Write CloudCoeff.bin (idempotent); returns its path."""
    os.makedirs(dirpath, exist_ok=True)
    path = os.path.join(dirpath, CLOUDCOEFF_FILE)
    if not os.path.exists(path):
        coeffio.write_cloudcoeff(path + ".tmp", make_cloud_coeff())
        os.replace(path + ".tmp", path)
    return path


# pressure band (hPa) of each cloud type's block and its radius range (um)
_CLOUD_BANDS = {WATER_CLOUD: ((600.0, 950.0), (5.0, 20.0)), ICE_CLOUD: ((150.0, 450.0), (20.0, 80.0)),
                RAIN_CLOUD: ((700.0, 1000.0), (5.0, 20.0)), SNOW_CLOUD: ((350.0, 750.0), (20.0, 80.0))}


def add_clouds(ps: ProfileSet, seed: int = 20261204,
               types: Sequence[int] = (WATER_CLOUD, ICE_CLOUD, RAIN_CLOUD, SNOW_CLOUD),
               max_layers: int = 10, median_gm3: float = 0.05, sigma: float = 1.0,
               cloudy_fraction: float = 1.0, cover: str = "overcast",
               reff_scale: float = 1.0) -> ProfileSet:
    """Cloudy copy of ps: per profile and type a block of U{0..max_layers}
    contiguous layers in the type's pressure band, water content
    median_gm3 * exp(N(0, sigma)) g/m^3 times the hypsometric layer thickness
    (kg/m^2).  cloudy_fraction of the profiles get clouds.  cover: 'overcast'
    (cloud fraction 1 in cloudy layers) or 'random' (U(0.2, 1))."""
    rng = np.random.default_rng(seed)
    M, nc = ps.n_profiles, len(types)
    ntot = int(ps.layer_offset[-1])
    wc = np.zeros((nc, ntot))
    re_ = np.zeros((nc, ntot))
    cf = np.zeros(ntot)
    for m in range(M):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        if rng.uniform() >= cloudy_fraction:
            continue
        lp = ps.level_pressure[a + m:b + m + 1]
        p = ps.pressure[a:b]
        dz = 29.2712 * ps.temperature[a:b] * np.log(lp[1:] / lp[:-1])      # m (R_d / g0)
        for n, t in enumerate(types):
            (plo, phi), (rlo, rhi) = _CLOUD_BANDS[t]
            n_lay = int(rng.integers(0, max_layers + 1))
            band = np.nonzero((p >= plo) & (p <= phi))[0]
            if n_lay == 0 or band.size == 0:
                continue
            n_lay = min(n_lay, band.size)
            k0 = band[0] + int(rng.integers(0, band.size - n_lay + 1))
            ks = np.arange(k0, k0 + n_lay)
            wc[n, a + ks] = median_gm3 * 1e-3 * np.exp(sigma * rng.standard_normal(n_lay)) * dz[ks]
            re_[n, a + ks] = reff_scale * rng.uniform(rlo, rhi, n_lay)
            cf[a + ks] = 1.0 if cover == "overcast" else rng.uniform(0.2, 1.0, n_lay)
    out = dataclasses.replace(ps)
    out.cloud_type = np.tile(np.asarray(types, np.int32), (M, 1))
    out.cloud_water_content = wc
    out.cloud_effective_radius = re_
    out.cloud_fraction = cf
    return out


# --------------------------------------------------------------------------
# Aerosols
# --------------------------------------------------------------------------
AEROSOLCOEFF_FILE = "AerosolCoeff.bin"
DUST_AEROSOL, SEASALT_SSAM_AEROSOL, SEASALT_SSCM1_AEROSOL, SEASALT_SSCM2_AEROSOL, SEASALT_SSCM3_AEROSOL, \
    ORGANIC_CARBON_AEROSOL, BLACK_CARBON_AEROSOL, SULFATE_AEROSOL = range(1, 9)
_AEROSOL_TYPES = [  # id, name, particle density kg/m^3, radius table (um), single-scatter albedo level
    (DUST_AEROSOL, "Dust", 2600.0, (0.1, 10.0), 0.80),
    (SEASALT_SSAM_AEROSOL, "Sea salt (SSAM)", 2200.0, (0.2, 20.0), 0.95),
    (SEASALT_SSCM1_AEROSOL, "Sea salt (SSCM1)", 2200.0, (0.2, 20.0), 0.95),
    (SEASALT_SSCM2_AEROSOL, "Sea salt (SSCM2)", 2200.0, (0.2, 20.0), 0.95),
    (SEASALT_SSCM3_AEROSOL, "Sea salt (SSCM3)", 2200.0, (0.2, 20.0), 0.95),
    (ORGANIC_CARBON_AEROSOL, "Organic carbon", 1800.0, (0.02, 2.0), 0.75),
    (BLACK_CARBON_AEROSOL, "Black carbon", 1000.0, (0.02, 2.0), 0.25),
    (SULFATE_AEROSOL, "Sulfate", 1700.0, (0.02, 3.0), 0.90),
]


def make_aerosol_coeff() -> coeffio.AerosolCoeff:
    """Smooth synthetic aerosol tables in the 'CRTM' scheme layout (one RH
    and one size-variance entry): extinction from a size-parameter
    efficiency 2x^2/(1+x^2) over 4/3 rho r, type-dependent albedo, HG-like
    Legendre coefficients with asymmetry g = x^2/2(1+x^2) <= 1/2."""
    wl = np.geomspace(50.0, 2.0, 40)                     # microns, descending: frequency ascends
    nR, nT = 12, len(_AEROSOL_TYPES)
    nL1 = N_LEGENDRE_FILE + 1
    reff = np.zeros((nR, nT))
    shp = (wl.size, 1, nR, 1, nT)
    ke, w, g = np.zeros(shp), np.zeros(shp), np.zeros(shp)
    for t, (_, _, rho, (r0, r1), w0) in enumerate(_AEROSOL_TYPES):
        r = np.geomspace(r0, r1, nR)
        reff[:, t] = r
        x = 2.0 * np.pi * r[None, :] / wl[:, None]                       # (wl, r)
        q = 2.0 * x ** 2 / (1.0 + x ** 2)
        ke[:, 0, :, 0, t] = 7.5e5 * q / (rho * r[None, :]) + 5.0
        w[:, 0, :, 0, t] = w0 * (0.6 + 0.4 * x / (1.0 + x)) * (0.9 + 0.1 * np.sin(wl / 3.0))[:, None]
        g[:, 0, :, 0, t] = 0.5 * x ** 2 / (1.0 + x ** 2)
    pc = _hg_slots(g, nL1)[..., None]                                     # (wl, 1, r, 1, t, 0:38, 1)
    return coeffio.AerosolCoeff(
        type_id=np.array([a[0] for a in _AEROSOL_TYPES], np.int32), type_name=[a[1] for a in _AEROSOL_TYPES],
        wavelength=wl, reff=reff, rsig=np.ones((1, nT)), rh=np.zeros(1), ke=ke, w=w, g=g, kb=ke * (1.0 - g),
        pcoeff=pc)


def write_aerosol_coeff(dirpath: str) -> str:
    """Write AerosolCoeff.bin (idempotent); returns its path."""
    os.makedirs(dirpath, exist_ok=True)
    path = os.path.join(dirpath, AEROSOLCOEFF_FILE)
    if not os.path.exists(path):
        coeffio.write_aerosolcoeff(path + ".tmp", make_aerosol_coeff())
        os.replace(path + ".tmp", path)
    return path


_AEROSOL_RADII = {DUST_AEROSOL: (0.5, 4.0), SEASALT_SSAM_AEROSOL: (0.5, 5.0), SULFATE_AEROSOL: (0.05, 0.5),
                  ORGANIC_CARBON_AEROSOL: (0.05, 0.5), BLACK_CARBON_AEROSOL: (0.05, 0.3)}


def add_aerosols(ps: ProfileSet, seed: int = 20261205,
                 types: Sequence[int] = (DUST_AEROSOL, SEASALT_SSAM_AEROSOL, SULFATE_AEROSOL),
                 n_layers: int = 20, median_kgm2: float = 2.0e-6, sigma: float = 1.0,
                 aerosol_fraction: float = 1.0) -> ProfileSet:
    """Copy of ps with aerosols: each type log-normal in mass concentration
    (median_kgm2 * exp(N(0, sigma)) kg/m^2 per layer) in the lowest n_layers
    layers, effective radius uniform in the type's range (microns).  A
    fraction aerosol_fraction of the profiles carries them; the others get
    unused slots (type 0: no aerosols, n_Aerosols = 0)."""
    rng = np.random.default_rng(seed)
    M, na = ps.n_profiles, len(types)
    ntot = int(ps.layer_offset[-1])
    conc = np.zeros((na, ntot))
    re_ = np.zeros((na, ntot))
    atype = np.tile(np.asarray(types, np.int32), (M, 1))
    for m in range(M):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        if rng.uniform() >= aerosol_fraction:
            atype[m] = 0
            continue
        k0 = max(a, b - n_layers)
        for n, t in enumerate(types):
            lo, hi = _AEROSOL_RADII.get(t, (0.1, 1.0))
            conc[n, k0:b] = median_kgm2 * np.exp(sigma * rng.standard_normal(b - k0))
            re_[n, k0:b] = rng.uniform(lo, hi, b - k0)
    out = dataclasses.replace(ps)
    out.aerosol_type = atype
    out.aerosol_concentration = conc
    out.aerosol_effective_radius = re_
    return out


# ----------------------------------------------------------------------------
# Surface emissivity model coefficients: FASTEM6 (microwave water) and the
# IR sea-surface emissivity table.  Synthetic values of physically plausible
# size, so that every term of both models is exercised.
# ----------------------------------------------------------------------------
MWWATERCOEFF_FILE = "FASTEM6.MWwater.EmisCoeff.bin"
IRWATERCOEFF_FILE = "Nalli.IRwater.EmisCoeff.bin"


def make_mwwater_coeff(seed: int = 20261206) -> coeffio.MWwaterCoeff:
    rng = np.random.default_rng(seed)
    u = lambda *shape: rng.uniform(0.6, 1.4, shape)
    fc = np.array([2.95e-6, 3.52])                         # Monahan-O'Muircheartaigh whitecap law
    fr = np.array([0.93, -2.0e-4, -2.0e-5, 1.0e-8, 1.0, -2.0e-3])
    # reflection correction: each of the 21 terms (predictors 1, var, var/cos,
    # 1/cos, (var/cos)^2, 1/cos^2, var^2 at their largest -- slope variance
    # var <= 0.2 up to 35 m/s, cos >= 0.5 -- times 1, x, x^2 with |x| =
    # |ln(od)| <= 7) adds at most 0.02 to the roughness exponent, which so
    # stays within [0.58, 1.42]
    zmag = np.array([1.0, 0.2, 0.4, 2.0, 0.16, 4.0, 0.04])
    rc = np.empty((3, 7, 2))
    for t, xmax in enumerate((1.0, 7.0, 49.0)):
        rc[t] = 0.02 / (xmax * zmag)[:, None] * u(7, 2) * rng.choice((-1.0, 1.0), (7, 2))
    az = np.zeros((6, 6, 2))
    fscale = u(6)
    az[0, :, 0] = 1.0
    az[1:4, :, 0] = np.array([-3.0e-4, 2.0e-5, -1.0e-6])[:, None] * fscale
    az[4, :, 0] = 0.02
    az[5, :, 0] = 1.0e-4 * fscale
    az[0, :, 1] = 1.2e-4 * fscale
    az[1, :, 1] = 1.0
    az[2:5, :, 1] = np.array([-2.5e-4, 1.5e-5, -8.0e-7])[:, None] * fscale
    az[5, :, 1] = 0.025
    ssc = np.array([1.0e-5, 1.0e-8, 1.0e-6, -1.0e-9, 1.0e-3, 1.0e-2, 1.0e-3, 1.0e-4]) * u(8)
    z0 = np.array([[-1.0e-2, 1.2e-2, -3.0e-3, -2.0e-4, 5.0e-6, 1.0e-4],
                   [-1.2e-2, 1.5e-2, -3.5e-3, 2.5e-4, -4.0e-6, -1.5e-4]]).T      # (term, pol)
    lsc = np.stack([z0, 2.0e-3 * z0, -1.0e-5 * z0]) * u(3, 6, 2)
    return coeffio.MWwaterCoeff(fc=fc, fr=fr, rc=rc, az=az, ssc=ssc, lsc=lsc)


def make_irwater_coeff(with_temperature: bool = True) -> coeffio.IRwaterCoeff:
    ang = np.array([0.0, 10.0, 20.0, 30.0, 40.0, 50.0, 55.0, 60.0, 65.0, 70.0, 75.0, 80.0, 85.0])
    nu = np.linspace(600.0, 3000.0, 49)
    v = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 15.0])
    t = np.array([271.0, 280.0, 290.0, 300.0, 310.0]) if with_temperature else np.array([300.0])
    A, F, V, T = np.meshgrid(ang, nu, v, t, indexing="ij")
    base = 0.976 - 0.010 * np.sin(F / 310.0) - 0.004 * np.cos(F / 90.0)
    fall = (1.0 - np.cos(np.radians(A))) ** 3 * (0.62 - 0.012 * V)
    e = base * (1.0 - fall) + 1.5e-4 * (T - 290.0) * (1.0 + 0.5 * np.cos(F / 500.0))
    return coeffio.IRwaterCoeff("Nalli2" if with_temperature else "Nalli", ang, nu, v, t, e)


def _write_if_changed(path: str, writer) -> str:
    tmp = f"{path}.{os.getpid()}.tmp"
    writer(tmp)
    if os.path.exists(path) and open(path, "rb").read() == open(tmp, "rb").read():
        os.remove(tmp)
    else:
        os.replace(tmp, path)
    return path


def write_mwwater_coeff(dirpath: str) -> str:
    os.makedirs(dirpath, exist_ok=True)
    return _write_if_changed(os.path.join(dirpath, MWWATERCOEFF_FILE),
                             lambda p: coeffio.write_mwwatercoeff(p, make_mwwater_coeff()))


def write_irwater_coeff(dirpath: str, with_temperature: bool = True, name: str = IRWATERCOEFF_FILE) -> str:
    os.makedirs(dirpath, exist_ok=True)
    return _write_if_changed(os.path.join(dirpath, name),
                             lambda p: coeffio.write_irwatercoeff(p, make_irwater_coeff(with_temperature)))


def ocean_surface(ps: ProfileSet, seed: int = 20261207, wind_range=(0.5, 22.0)) -> ProfileSet:
    """All-water surfaces with random wind, salinity and view azimuth (the
    inputs of the computed surface optics); scan angles consistent with the
    zenith angles of an 833 km orbit."""
    rng = np.random.default_rng(seed)
    M = ps.n_profiles
    zen = np.asarray(ps.sensor_zenith, np.float64)
    # scan angle from the zenith angle: sin(scan) = sin(zenith) R / (R + h)
    scan = np.degrees(np.arcsin(np.sin(np.radians(zen)) * 6370.0 / (6370.0 + 833.0)))
    return dataclasses.replace(
        ps, land_coverage=np.zeros(M), water_coverage=np.ones(M), snow_coverage=np.zeros(M),
        ice_coverage=np.zeros(M), sensor_scan=scan * rng.choice((-1.0, 1.0), M),
        wind_speed=rng.uniform(*wind_range, M), wind_direction=rng.uniform(0.0, 360.0, M),
        salinity=rng.uniform(30.0, 37.0, M), sensor_azimuth=rng.uniform(0.0, 360.0, M))
