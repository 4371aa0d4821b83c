#!/usr/bin/env python3
"""Write the golden fixtures of the forward / K-matrix path.

Inputs (regression_inputs.json): the two profiles and surfaces of the
reference's own regression test, parsed from
test/mains/regression/incsrc/Load_Atm_Data.inc and Load_Sfc_Data.inc
(92 layers from 0.714 hPa, so CRTM's AddLayers extends them with the
US-standard and tropical climatologies; H2O g/kg and O3 ppmv; mixed
land/water/snow/ice coverage).  The reference ships no outputs for them
(its regression "saved" files are produced on first run), so the expected
values (regression_expected.npz) are the FP64 oracle's, computed on the
synthetic sensors whose coefficient files are fingerprinted in the JSON.

    python tests/golden/make_golden.py [/path/to/reference]
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SENSORS = ("amsua_syn", "hirs_syn", "cris_small_syn")
ZENITH = (30.0, 52.5)          # sensor zenith of profiles 1, 2
CLIMATOLOGY = {"US_STANDARD_ATMOSPHERE": 6, "TROPICAL": 1, "MIDLATITUDE_SUMMER": 2,
               "MIDLATITUDE_WINTER": 3, "SUBARCTIC_SUMMER": 4, "SUBARCTIC_WINTER": 5}


def _array(txt, lhs):
    m = re.search(re.escape(lhs) + r"\s*=\s*&?\s*\(/(.*?)/\)", txt, re.S)
    if not m:
        raise ValueError(lhs)
    body = re.sub(r"!.*", "", m.group(1)).replace("&", " ")
    return [float(v.replace("_fp", "")) for v in re.findall(r"[-+0-9.Ee]+_fp", body)]


def _scalar(txt, lhs):
    m = re.search(re.escape(lhs) + r"\s*=\s*([-+0-9.Ee]+)_fp", txt)
    return float(m.group(1))


def parse(ref):
    atm = open(os.path.join(ref, "test/mains/regression/incsrc/Load_Atm_Data.inc")).read()
    sfc = open(os.path.join(ref, "test/mains/regression/incsrc/Load_Sfc_Data.inc")).read()
    profs = []
    for m in (1, 2):
        clim = re.search(rf"atm\({m}\)%Climatology\s*=\s*(\w+)", atm).group(1)
        p = {"climatology": CLIMATOLOGY[clim],
             "level_pressure": _array(atm, f"atm({m})%Level_Pressure"),
             "pressure": _array(atm, f"atm({m})%Pressure"),
             "temperature": _array(atm, f"atm({m})%Temperature"),
             "h2o": _array(atm, f"atm({m})%Absorber(:,1)"),
             "o3": _array(atm, f"atm({m})%Absorber(:,2)")}
        for kind in ("Land", "Water", "Snow", "Ice"):
            for q in ("Coverage", "Temperature"):
                lhs = f"sfc({m})%{kind}_{q}"
                p[f"{kind.lower()}_{q.lower()}"] = _scalar(sfc, lhs) if re.search(re.escape(lhs) + r"\s*=", sfc) else (
                    0.0 if q == "Coverage" else 280.0)
        p["sensor_zenith"] = ZENITH[m - 1]
        profs.append(p)
    return profs


def profile_set(profs):
    from crtmv3_amd.synthetic import ProfileSet
    M = len(profs)
    ks = np.array([len(p["temperature"]) for p in profs], np.int32)
    off = np.zeros(M + 1, np.int64)
    off[1:] = np.cumsum(ks)
    g = lambda key: np.array([p[key] for p in profs], float)
    return ProfileSet(
        n_layers=ks, layer_offset=off, absorber_id=np.array([1, 3], np.int32),
        absorber_units=np.array([3, 1], np.int32),
        level_pressure=np.concatenate([p["level_pressure"] for p in profs]),
        pressure=np.concatenate([p["pressure"] for p in profs]),
        temperature=np.concatenate([p["temperature"] for p in profs]),
        absorber=np.stack([np.concatenate([p["h2o"] for p in profs]), np.concatenate([p["o3"] for p in profs])]),
        land_coverage=g("land_coverage"), water_coverage=g("water_coverage"), snow_coverage=g("snow_coverage"),
        ice_coverage=g("ice_coverage"), land_temperature=g("land_temperature"),
        water_temperature=g("water_temperature"), snow_temperature=g("snow_temperature"),
        ice_temperature=g("ice_temperature"), sensor_zenith=g("sensor_zenith"), sensor_scan=g("sensor_zenith"),
        source_zenith=np.full(M, 100.0), climatology=np.array([p["climatology"] for p in profs], np.int32))


def fingerprint(coeff_dir, name):
    h = hashlib.sha256()
    for suffix in ("SpcCoeff.bin", "TauCoeff.bin"):
        h.update(open(os.path.join(coeff_dir, f"{name}.{suffix}"), "rb").read())
    return h.hexdigest()


def emissivity(sensor_type, L):
    return np.linspace(0.6, 0.95, L) if sensor_type == 1 else np.linspace(0.93, 0.99, L)


def expected(coeff_dir, ps):
    from crtmv3_amd import coeffio
    from oracle import oracle as orc
    out = {}
    for name in SENSORS:
        tau = coeffio.read_odps(os.path.join(coeff_dir, f"{name}.TauCoeff.bin"))
        spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{name}.SpcCoeff.bin"))
        s = orc.Sensor(tau, spc)
        eps = emissivity(spc.sensor_type, tau.n_channels)
        for m in range(ps.n_profiles):
            r = orc.run_profile(s, ps, m, eps, "k")
            for key in ("bt", "radiance", "t_k", "abs_k", "sfc_t_k", "emis_k"):
                out[f"{name}/{m}/{key}"] = r[key]
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    from crtmv3_amd import synthetic
    profs = parse(ref)
    coeff_dir = synthetic.default_coeff_dir()
    for name in SENSORS:
        synthetic.write_sensor(coeff_dir, name)
    doc = {"source": "test/mains/regression/incsrc/Load_Atm_Data.inc, Load_Sfc_Data.inc",
           "zenith": ZENITH, "profiles": profs,
           "coefficients": {name: fingerprint(coeff_dir, name) for name in SENSORS}}
    with open(os.path.join(HERE, "regression_inputs.json"), "w") as f:
        json.dump(doc, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "regression_expected.npz"), **expected(coeff_dir, profile_set(profs)))
    print("wrote", os.path.join(HERE, "regression_inputs.json"), "and regression_expected.npz")


if __name__ == "__main__":
    main()
