"""Shared fixtures.  Tests marked `gpu` need an MI355X and the built
libcrtm_mi355x.so; everything else runs on the CPU (oracle, coefficient
I/O, host-side logic, library exports)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# small sensors used by most tests: (name, ODPS group, OPTRAN)
SMALL_SENSORS = ("amsua_syn", "hirs_syn", "cris_small_syn")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def coeff_dir(tmp_path_factory):
    """Synthetic SpcCoeff/ODPS files in CRTM binary format for the small sensors."""
    from crtmv3_amd import synthetic
    d = os.environ.get("CRTM_SYNTH_DIR") or str(tmp_path_factory.mktemp("coeffs"))
    for name in SMALL_SENSORS:
        synthetic.write_sensor(d, name)
    synthetic.write_cloud_coeff(d)
    synthetic.write_aerosol_coeff(d)
    return d


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as orc
    orc.lib()
    return orc


def load_sensor(coeff_dir, name):
    """(ODPS, SpcCoeff) read back with the package reader."""
    from crtmv3_amd import coeffio, synthetic
    synthetic.write_sensor(coeff_dir, name)
    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{name}.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{name}.SpcCoeff.bin"))
    return tau, spc


def profiles_from_arrays(level_pressure, temperature, absorber, water_t, zenith, absorber_id=(1, 3),
                         units=(3, 1), pressure=None):
    """ProfileSet from explicit per-profile arrays (lists over profiles)."""
    from crtmv3_amd.synthetic import ProfileSet
    M = len(temperature)
    ks = np.array([len(t) for t in temperature], np.int32)
    off = np.zeros(M + 1, np.int64)
    off[1:] = np.cumsum(ks)
    lvl = np.concatenate([np.asarray(x, float) for x in level_pressure])
    if pressure is None:
        pressure = [(np.asarray(lp[1:]) - np.asarray(lp[:-1])) / np.log(np.asarray(lp[1:]) / np.asarray(lp[:-1]))
                    for lp in level_pressure]
    lay = np.concatenate([np.asarray(p, float) for p in pressure])
    na = len(absorber_id)
    ab = np.concatenate([np.asarray(a, float).reshape(na, -1) for a in absorber], axis=1)
    return ProfileSet(
        n_layers=ks, layer_offset=off, absorber_id=np.array(absorber_id, np.int32),
        absorber_units=np.array(units, np.int32), level_pressure=lvl, pressure=lay,
        temperature=np.concatenate([np.asarray(t, float) for t in temperature]), absorber=ab,
        land_coverage=np.zeros(M), water_coverage=np.ones(M), snow_coverage=np.zeros(M),
        ice_coverage=np.zeros(M), land_temperature=np.full(M, 280.0),
        water_temperature=np.asarray(water_t, float), snow_temperature=np.full(M, 260.0),
        ice_temperature=np.full(M, 260.0), sensor_zenith=np.asarray(zenith, float),
        sensor_scan=np.zeros(M), source_zenith=np.full(M, 100.0))


# Element-wise Jacobian bar of every GPU K-matrix test, besides the
# block-relative bar each test states: every element of a block (one
# Jacobian variable of one profile, all channels and layers) within
#     |got - ref| <= rtol * |ref| + atol * max|ref of the block|.
# FP64: rtol 1e-7, atol 1e-10.  FP32-split: rtol 1e-3, atol 1e-5 for the
# gas and surface variables, 2e-4 for the scattering-path variables (cloud
# water content, radius and fraction, aerosol concentration and radius).
# Where terms cancel, FP32 leaves those an absolute error of up to about
# 1.1e-4 of the block maximum.  The worst cases seen are cloud-fraction
# Jacobians: overcast, where the cover term (R_cloudy - R_clear) nearly
# cancels the water-content rescaling term (6e-5), and absorption-only
# clouds (Options%Include_Scattering = .FALSE.) in microwave channels that
# the cloud barely affects, where R_cloudy - R_clear is ~1e-6 of R (1.1e-4).
# Elements above about 0.1 of the block maximum are within 1e-3 relative.
SCATTERING_VARIABLES = ("water content", "radius", "fraction", "aerosol")
JAC_ELEM = {0: dict(rtol=1e-7, atol=1e-10, atol_scattering=1e-10),
            1: dict(rtol=1e-3, atol=1e-5, atol_scattering=2e-4)}


def assert_jacobian_close(got, want, rtol_block, what, precision):
    """(block-relative error, worst element-wise relative error among the
    elements where the relative part of the bar dominates)."""
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.abs(want).max() if want.size else 0.0
    if scale == 0.0:
        err = np.abs(got).max() if got.size else 0.0
        assert err <= rtol_block, f"{what}: reference zero, got max {err:.3e}"
        return err, 0.0
    diff = np.abs(got - want)
    blk = diff.max() / scale
    assert blk <= rtol_block, f"{what}: block-relative error {blk:.3e}"
    bar = JAC_ELEM[int(precision)]
    scattering = any(v in what for v in SCATTERING_VARIABLES)
    atol = (bar["atol_scattering"] if scattering else bar["atol"]) * scale
    allowed = bar["rtol"] * np.abs(want) + atol
    if (diff > allowed).any():
        i = np.unravel_index(np.argmax(diff - allowed), diff.shape)
        raise AssertionError(f"{what}: element {i} off by {diff[i]:.3e} = {diff[i] / abs(want[i]):.3e} relative "
                             f"(element {abs(want[i]) / scale:.3e} of the block max); bar rtol {bar['rtol']:g}, "
                             f"atol {atol / scale:g} of the block max")
    rel_regime = bar["rtol"] * np.abs(want) >= atol
    elem = float((diff[rel_regime] / np.abs(want[rel_regime])).max()) if rel_regime.any() else 0.0
    return blk, elem


def subset_sensor(tau, spc, idx):
    """ODPS and SpcCoeff copies holding only channels idx (the coefficient
    arrays are kept whole: the per-channel start indices still point into
    them), so the oracle can check a wide sensor on a channel sample."""
    import copy
    idx = np.asarray(idx)
    t = copy.deepcopy(tau)
    t.sensor_channel = tau.sensor_channel[idx].copy()
    t.n_predictors = tau.n_predictors[:, idx].copy()
    t.pos_index = tau.pos_index[:, idx].copy()
    if tau.osignificance is not None:
        t.osignificance = tau.osignificance[idx].copy()
        t.order = tau.order[idx].copy()
        t.op_index = tau.op_index[:, idx].copy()
        t.opos_index = tau.opos_index[idx].copy()
    s = copy.deepcopy(spc)
    for f in ("sensor_channel", "polarization", "pol_angle", "channel_flag", "frequency", "wavenumber",
              "planck_c1", "planck_c2", "band_c1", "band_c2", "cosmic_background_radiance", "solar_irradiance"):
        setattr(s, f, getattr(spc, f)[idx].copy())
    return t, s
