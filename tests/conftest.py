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
