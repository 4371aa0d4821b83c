"""Build libcrtm_mi355x.so (HIP kernels + C-ABI) for gfx950 in-tree.

    python -m crtmv3_amd.build            # compile what changed
    python -m crtmv3_amd.build --force    # recompile everything

Every HIP translation unit is compiled with packed-FP32 instruction
selection disabled (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libcrtm_mi355x.so")
# clear_fwdk_pw0/pw2.hip instantiate the fused clear-sky kernel's other two
# tile sets from clear_fwdk.hip, so the three compile in parallel
SOURCES = ["odps_prep.hip", "clear_fwdk.hip", "clear_fwdk_pw0.hip", "clear_fwdk_pw2.hip", "sfc_optics.hip",
           "allsky.hip", "allsky_ad.hip", "allsky_wide.hip", "allsky_wide_ad.hip", "crtm_abi.cpp"]
HEADERS = ["crtm_device.h", "sfc_optics.h", "allsky_common.h", "allsky_wide.h", "pred_adjoint.h", "clear_fwdk.hip",
           os.path.join("..", "..", "include", "crtm_mi355x.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CRTM_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


CLIM_JSON = os.path.join(HERE, "data", "model_profiles.json")
CLIM_INC = os.path.join(OBJ, "model_profiles.inc")


def _write_climatology_table() -> None:
    """C table of the six model climatologies (from data/model_profiles.json,
    extracted from src/Atmosphere/CRTM_Model_Profiles.f90) for AddLayers:
    level pressure, temperature per model, absorbers per model in absorber-id
    order H2O, CO2, O3, N2O, CO, CH4."""
    import json
    d = json.load(open(CLIM_JSON))
    order = [d["absorbers"].index(n) for n in ("H2O", "CO2", "O3", "N2O", "CO", "CH4")]
    vals = list(d["level_pressure"])
    for mdl in range(6):
        vals += d["temperature"][mdl]
    for mdl in range(6):
        for j in order:
            vals += d["absorber"][mdl][j]
    body = ",\n".join(", ".join(repr(float(v)) for v in vals[i:i + 6]) for i in range(0, len(vals), 6))
    txt = ("// generated by crtmv3_amd/build.py from data/model_profiles.json\n"
           f"static const double CLIM_TABLE[{len(vals)}] = {{\n{body}}};\n")
    if not os.path.exists(CLIM_INC) or open(CLIM_INC).read() != txt:
        with open(CLIM_INC, "w") as f:
            f.write(txt)


def _stale(obj: str, src: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src, CLIM_INC] + [os.path.join(CSRC, h) for h in HEADERS]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if force or _stale(obj, path):
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-I" + OBJ, "-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


FORTRAN = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
F_SOURCES = ["crtm_mi355x_module.f90", "crtm_module.f90", "crtm_mi355x_driver.f90"]
DRIVER = os.path.join(HERE, "crtm_mi355x_driver")


def build_fortran() -> str:
    """Fortran host interface (CRTM_MI355X_Module, re-exported as
    CRTM_Module) and its example driver, linked against libcrtm_mi355x.so."""
    fdir = os.path.join(OBJ, "fortran")
    os.makedirs(fdir, exist_ok=True)
    objs = []
    for src in F_SOURCES:
        obj = os.path.join(fdir, src.replace(".f90", ".o"))
        cmd = [FORTRAN, "-O2", "-c", os.path.join(HERE, "fortran", src), "-o", obj, "-module-dir", fdir]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=fdir)
        if r.returncode != 0:
            raise RuntimeError(f"{FORTRAN} failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(obj)
    cmd = [FORTRAN, "-o", DRIVER] + objs[::-1] + ["-L" + HERE, "-lcrtm_mi355x", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stderr[-4000:]}")
    return DRIVER


def build(force: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    _write_climatology_table()
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", LIB + ".tmp"] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
        os.replace(LIB + ".tmp", LIB)
    if os.path.exists(FORTRAN):
        build_fortran()
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
