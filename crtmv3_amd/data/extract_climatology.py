"""Extract the six 101-level model climatologies into model_profiles.json.

Source: the DATA statements of src/Atmosphere/CRTM_Model_Profiles.f90
(MODEL_LEVEL_PRESSURE at :80-97, Level_Temperature from :108,
Level_Absorber from :242).  The JSON is the data table the synthetic profile
generator and the AddLayers climatology use; the extraction is rerun only
when the reference tree is present.

Usage: python crtmv3_amd/data/extract_climatology.py [reference_root]
"""
import json
import os
import re
import sys

MODELS = ["TROPICAL", "MIDLATITUDE_SUMMER", "MIDLATITUDE_WINTER",
          "SUBARCTIC_SUMMER", "SUBARCTIC_WINTER", "US_STANDARD_ATMOSPHERE"]
ABSORBERS = ["H2O", "O3", "CO2", "N2O", "CO", "CH4"]   # CRTM_Model_Profiles.f90:56-61


def _numbers(block: str):
    return [float(x.replace("_fp", "").replace("D", "E").replace("d", "e"))
            for x in re.findall(r"[-+]?\d+\.\d*(?:[EeDd][-+]?\d+)?_fp", block)]


def extract(root: str) -> dict:
    path = os.path.join(root, "src", "Atmosphere", "CRTM_Model_Profiles.f90")
    txt = open(path).read()
    m = re.search(r"MODEL_LEVEL_PRESSURE\(0:N_MODEL_LAYERS\)\s*=\s*&(.*?)/\)", txt, re.S)
    pressure = _numbers(m.group(1))
    temp, absorber = {}, {}
    for blk in re.finditer(r"DATA\s+Level_(Temperature|Absorber)\(:,([\w,]+)\)\s*/(.*?)/", txt, re.S):
        kind, idx, body = blk.group(1), blk.group(2).split(","), blk.group(3)
        vals = _numbers(body)
        if kind == "Temperature":
            temp[idx[0]] = vals
        else:
            absorber.setdefault(idx[1], {})[idx[0]] = vals
    out = {"level_pressure": pressure,
           "models": MODELS,
           "absorbers": ABSORBERS,
           "units": {"H2O": "g/kg", "O3": "ppmv", "CO2": "ppmv", "N2O": "ppmv",
                     "CO": "ppmv", "CH4": "ppmv", "temperature": "K", "pressure": "hPa"},
           "temperature": [temp[mm] for mm in MODELS],
           "absorber": [[absorber[mm][a] for a in ABSORBERS] for mm in MODELS]}
    assert len(pressure) == 101
    for t in out["temperature"]:
        assert len(t) == 101
    for mm in out["absorber"]:
        for a in mm:
            assert len(a) == 101
    return out


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(root)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model_profiles.json")
    with open(dst, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", dst)
