#!/usr/bin/env python3
"""Forward + K-matrix throughput of the MI355X CRTM hot path.

Workload (one step): clear-sky forward + K-matrix for a 2211-channel CrIS-like
sensor (ODPS group 1 with OPTRAN water-vapour lines) over 4096 synthetic
100-layer profiles, Jacobians dBT/dT, dBT/dq, dBT/dO3 per layer.  Inputs and
outputs are resident in HBM for the timed region (device-memory C-ABI calls).

Reported: FP32-split profile-channels/s (value), the plain FP64 GPU path
measured on the same inputs, the BT / Jacobian differences between them, the
roofline of the fused forward/K kernel measured with HIP events on the
context stream, and the CPU baseline (the FP64 restatement in oracle/, timed
on this host's cores on a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "forward+K-matrix profile-channels/s, FP32-split vs FP64 GPU, at 0.01 K BT err"
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 (vector incl. dual issue / MFMA)
JAC_SAMPLE = 32              # profiles whose Jacobians are compared with the oracle
PEAK_HBM_GBS = 8000.0
C_PAD = 250                  # adjoint flops per ODPS layer of the group-1 predictor chain


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_flops_per_pc(pbar, ko, k):
    """F_fwd = 2 P Ko + 10 Ko + 25 K;  F_K = 2 P Ko + c_pAD Ko + 70 K."""
    return 4.0 * pbar * ko + (10 + C_PAD) * ko + 95.0 * k


def algorithmic_bytes_per_pc(k, n_abs, kbytes):
    # Jacobians (T + absorbers) + BT, radiance, emissivity-K, water-temperature-K
    return (1 + n_abs) * k * kbytes + 4 * 8


def host_threads():
    """CPUs this process may run on (its affinity), capped by OMP_NUM_THREADS
    when that is set (a shared host gives each GPU a share of its cores)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def jacobian_errors(gpu, ref, ps, n, L, na):
    """Block-relative and element-wise relative errors (elements above 1e-3
    of the block maximum) of T and absorber Jacobians of the first n
    profiles: gpu = (t_k [NT, L], a_k [na, NT, L]) host arrays, ref = oracle
    batch output (per profile [L, K] and [L, na, K])."""
    off = ps.layer_offset
    res = {"T": [0.0, 0.0], "absorbers": [0.0, 0.0]}
    for m in range(n):
        a, b = int(off[m]), int(off[m + 1])
        K = b - a
        rt = ref["t_k"][a * L:b * L].reshape(L, K)
        ra = ref["abs_k"][a * L * na:b * L * na].reshape(L, na, K)
        for key, g, r in (("T", gpu[0][a:b].T, rt), ("absorbers", gpu[1][:, a:b, :].transpose(2, 0, 1), ra)):
            g = np.asarray(g, np.float64)
            scale = np.abs(r).max()
            d = np.abs(g - r)
            big = np.abs(r) > 1e-3 * scale
            res[key][0] = max(res[key][0], float(d.max() / scale))
            res[key][1] = max(res[key][1], float((d[big] / np.abs(r[big])).max()))
    return {k: {"block_relative": v[0], "element_relative_above_1e-3_of_max": v[1]} for k, v in res.items()}


def cpu_baseline(sensor_name, coeff_dir, ps, emissivity, bt_gpu, seconds, threads):
    """Time the oracle (FP64 CPU restatement, OpenMP over profiles) on a
    sample; returns the baseline record, the BT error of the GPU split path
    on the sample, and the oracle's outputs for it."""
    from crtmv3_amd import coeffio, synthetic
    from oracle import oracle as orc
    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{sensor_name}.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{sensor_name}.SpcCoeff.bin"))
    s = orc.Sensor(tau, spc)
    L = tau.n_channels
    # calibrate on a small chunk, then size the sample to ~`seconds`
    chunk = max(threads, 16)
    sub = _subset(ps, 0, chunk)
    t0 = time.perf_counter()
    out = orc.k_matrix_batch(s, sub, emissivity[:chunk], n_threads=threads)
    dt = time.perf_counter() - t0
    n = min(ps.n_profiles, max(chunk, int(chunk * seconds / max(dt, 1e-6))))
    sub = _subset(ps, 0, n)
    t0 = time.perf_counter()
    out = orc.k_matrix_batch(s, sub, emissivity[:n], n_threads=threads)
    dt = time.perf_counter() - t0
    bt_err = float(np.max(np.abs(out["bt"] - bt_gpu[:n])))
    return {"value": n * L / dt, "unit": "profile-channels/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"{n} of the {ps.n_profiles} profiles x {L} channels, forward+K-matrix, "
                      f"FP64 oracle (oracle_k_matrix_batch), {dt:.1f} s"}, bt_err, out, n


def _subset(ps, a, b):
    import dataclasses
    off = ps.layer_offset
    la, lb = int(off[a]), int(off[b])
    d = {}
    for f in dataclasses.fields(ps):
        v = getattr(ps, f.name)
        if v is None:
            d[f.name] = None
        elif f.name == "layer_offset":
            d[f.name] = off[a:b + 1] - off[a]
        elif f.name == "level_pressure":
            d[f.name] = v[la + a:lb + b]
        elif f.name in ("pressure", "temperature", "cloud_fraction"):
            d[f.name] = v[la:lb]
        elif f.name in ("absorber", "cloud_water_content", "cloud_effective_radius"):
            d[f.name] = v[:, la:lb]
        elif f.name in ("absorber_id", "absorber_units"):
            d[f.name] = v
        else:
            d[f.name] = v[a:b]
    return type(ps)(**d)


COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "clear_fwdk_cris2211x4096_traffic.json")


def copy_calibration():
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "launches": 3}), flush=True)


KERNEL_SOURCES = [os.path.join(ROOT, "crtmv3_amd", "csrc", f)
                  for f in ("clear_fwdk.hip", "odps_prep.hip", "crtm_device.h", "pred_adjoint.h")]


def kernel_fingerprint():
    """sha256 of the fused kernel's sources: the traffic summary applies to
    the build it was measured on only."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def measured_traffic(workload):
    """HBM bytes per launch of the fused kernel from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by a plain copy), when
    it was taken on this workload with these kernel sources; else None."""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("kernel_sources_sha256") != kernel_fingerprint():
        return None
    return t["traffic_bytes"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--profiles", type=int, default=4096)
    ap.add_argument("--sensor", default="cris_syn")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--split-only", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="clear-ir: keep the default tile of the fused kernel instead of autotuning it first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (3 times), the reference for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--streams", type=int, default=8, help="allsky-ir: ADA stream count (8 or 16)")
    ap.add_argument("--wavefront-angles", type=int, default=0,
                    help="all-sky: angle counts >= this run on the wavefront (matrix-core) kernels (6, 9)")
    ap.add_argument("--workload", choices=("clear-ir", "allsky-mw", "allsky-ir", "batched", "sfc"), default="clear-ir",
                    help="clear-ir: the headline (CrIS-like forward + K-matrix); allsky-mw: cloudy ATMS-like "
                         "forward + K-matrix with the ADA scattering solver at 4 and 8 streams; allsky-ir: "
                         "IASI-like forward + K-matrix, 60/40 clear/cloudy, 8 streams; batched: three sensors, "
                         "ragged profiles with clouds and aerosols, per-profile stream counts, gradient of a "
                         "variational cost through torch.autograd; sfc: clear-sky K-matrix over an ocean with "
                         "the surface optics computed (FASTEM6 for ATMS-like, the IR emissivity table for "
                         "CrIS-like) against the same call with user emissivity")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload in ("allsky-mw", "allsky-ir", "batched", "sfc"):
        if args.gpus != 1:
            print("bench.py measures one GPU. Use --gpus 1.", file=sys.stderr)
            sys.exit(2)
        {"allsky-mw": main_allsky, "allsky-ir": main_allsky_ir, "batched": main_batched,
         "sfc": main_sfc}[args.workload](args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the MI355X path runs on a single device (profiles shard "
              "across GPUs with no exchange; see DESIGN.md). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from crtmv3_amd import api, synthetic

    dev = torch.device("cuda:0")
    coeff_dir = synthetic.default_coeff_dir()
    synthetic.write_sensor(coeff_dir, args.sensor)
    ctx = api.crtm_init([args.sensor], coeff_dir)
    L = ctx.n_channels
    M = args.profiles
    K = 100
    log(f"sensor {args.sensor}: {L} channels; generating {M} profiles")
    ps = synthetic.make_profiles(M, K, seed=20261203 + 2)
    dps = api.device_profiles(ps, dev)
    st = ctx.channel_info[0].sensor_type
    emis_h = np.tile(synthetic.default_emissivity(st, L), (M, 1))
    emis = torch.as_tensor(emis_h, device=dev)
    bt_k = torch.ones((M, L), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def run(prec, f32, out):
        return ctx.k_matrix(dps, emis, bt_k=bt_k, precision=prec, jacobian_f32=f32, out=out)

    # ---- tile shapes of the fused kernel, chosen by the autotuner (untimed) ----
    tiles = {}
    if not args.no_autotune:
        for prec, name, f32 in ((api.PREC_FP32_SPLIT, "split", True), (api.PREC_FP64, "fp64", False)):
            if name == "fp64" and args.split_only:
                continue
            rep = ctx.autotune(dps, emis, precision=prec, repeats=2, jacobian_f32=f32)
            tiles[name] = {"profiles_per_workgroup": rep["profiles_per_workgroup"][0],
                           "kernel_ms_per_tile": {str(api.tile_profiles(prec, t)): round(v, 3)
                                                  for t, v in enumerate(rep["ms"][0])}}
            log(f"autotune {name}: {tiles[name]}")
        torch.cuda.empty_cache()

    # ---- FP32-split (the headline) ----
    split_pw = api.tile_profiles(api.PREC_FP32_SPLIT, ctx.get_tile(0, api.PREC_FP32_SPLIT))
    out = run(api.PREC_FP32_SPLIT, True, None)
    for _ in range(args.warmup):
        out = run(api.PREC_FP32_SPLIT, True, out)
    ctx.synchronize()
    torch.cuda.synchronize()
    ctx.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = run(api.PREC_FP32_SPLIT, True, out)
    ctx.synchronize()
    torch.cuda.synchronize()
    t_split = (time.perf_counter() - t0) / args.steps
    prep_ms, main_ms, nl = ctx.profile_query()
    ctx.profile_enable(False)
    main_ms /= max(nl, 1)
    prep_ms /= max(nl, 1)
    bt_split = out[0]["brightness_temperature"].cpu().numpy()
    nt_j = int(ps.layer_offset[JAC_SAMPLE])
    jac_split = (out[1]["temperature_k"][:nt_j].cpu().numpy(), out[1]["absorber_k"][:, :nt_j].cpu().numpy())
    jac64 = None
    log(f"split: {t_split * 1e3:.2f} ms/step (fwd/K kernel {main_ms:.2f} ms, prep {prep_ms:.2f} ms)")

    result = {}
    if not args.split_only:
        out64 = run(api.PREC_FP64, False, None)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            out64 = run(api.PREC_FP64, False, out64)
        ctx.synchronize()
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / args.fp64_steps
        bt64 = out64[0]["brightness_temperature"].cpu().numpy()
        bt_err = float(np.max(np.abs(bt_split - bt64)))
        jac = {}
        for key, name in (("temperature_k", "dT"), ("absorber_k", "dabs")):
            a = out[1][key].double()
            b = out64[1][key]
            jac[name] = float((a - b).abs().max().item() / b.abs().max().item())
        result["fp64"] = {"ms_per_step": t64 * 1e3, "profile_channels_per_s": M * L / t64,
                          "speedup_split_over_fp64": t64 / t_split}
        result["bt_err_split_vs_fp64_K"] = bt_err
        result["jacobian_relerr_split_vs_fp64"] = jac
        jac64 = (out64[1]["temperature_k"][:nt_j].cpu().numpy(), out64[1]["absorber_k"][:, :nt_j].cpu().numpy())
        result["bt_err_fp64_vs_oracle_K"] = None
        bt64_sample = bt64
        log(f"fp64: {t64 * 1e3:.2f} ms/step; BT err {bt_err:.2e} K; jac {jac}")
        del out64
        torch.cuda.empty_cache()

    # ---- roofline of the fused forward/K kernel ----
    from crtmv3_amd import coeffio
    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{args.sensor}.TauCoeff.bin"))
    pbar = float(tau.n_predictors.sum(0).mean())
    fpc = algorithmic_flops_per_pc(pbar, tau.n_layers, K)
    bpc = algorithmic_bytes_per_pc(K, ps.absorber.shape[0], 4)
    achieved = fpc * M * L / (main_ms * 1e-3) / 1e12
    workload = (f"clear-sky forward+K-matrix, CrIS-like 2211-ch ODPS group 1, {M} profiles x {K} layers, "
                "dBT/dT,dq,dO3")
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(workload),
            "kernel": f"clear_fwdk_kernel<1,float,float,true,false,{split_pw}>", "kernel_ms": main_ms,
            "flops_per_pc": fpc, "bytes_per_pc": bpc,
            "hbm_achieved_GBs": bpc * M * L / (main_ms * 1e-3) / 1e9}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            threads = host_threads()
            cpu, bt_oracle_err, ref, n_ref = cpu_baseline(args.sensor, coeff_dir, ps, emis_h, bt_split,
                                                          args.cpu_seconds, threads)
            result["bt_err_split_vs_oracle_K"] = bt_oracle_err
            nj = min(JAC_SAMPLE, n_ref)
            na = ps.absorber.shape[0]
            result["jacobian_relerr_split_vs_oracle"] = jacobian_errors(jac_split, ref, ps, nj, L, na)
            if jac64 is not None:
                result["bt_err_fp64_vs_oracle_K"] = float(np.max(np.abs(bt64_sample[:n_ref] - ref["bt"])))
                result["jacobian_relerr_fp64_vs_oracle"] = jacobian_errors(jac64, ref, ps, nj, L, na)
            result["jacobian_oracle_sample_profiles"] = nj
            log(f"cpu baseline: {cpu['value']:.3e} pc/s on {threads} threads; split-vs-oracle BT err "
                f"{bt_oracle_err:.2e} K; Jacobians vs oracle: split {result['jacobian_relerr_split_vs_oracle']}")
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")

    line = {"metric": METRIC, "value": M * L / t_split, "unit": "profile-channels/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_split * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64-split",
            "data": "synthetic (seeded profiles; synthetic ODPS/SpcCoeff files in CRTM binary format)",
            "config": {"workload": workload,
                       "n_profiles": M, "n_channels": L, "n_layers": K, "precision": "fp32-split",
                       "parallelism": "single GPU",
                       "tile": tiles or "default (8 split / 4 FP64 profiles per workgroup)"},
            "roofline": roof, "cpu_baseline": cpu}
    line.update(result)
    print(json.dumps(line), flush=True)
    ctx.destroy()


def main_sfc(args):
    """Computed surface optics: clear-sky K-matrix over an all-water surface
    with random wind, salinity and azimuths, ATMS-like (FASTEM6, 16384
    profiles) and CrIS-like (IR emissivity table, 4096 profiles), FP32-split
    and FP64; the same call with user emissivity for the overhead; the optics
    kernels timed with HIP events; BT and Jacobian errors against the oracle
    on a sample."""
    import torch
    from crtmv3_amd import api, coeffio, synthetic
    from oracle import oracle as orc
    dev = torch.device("cuda:0")
    coeff_dir = synthetic.default_coeff_dir()
    synthetic.write_mwwater_coeff(coeff_dir)
    synthetic.write_irwater_coeff(coeff_dir)
    mw = orc.MWwater(coeffio.read_mwwatercoeff(os.path.join(coeff_dir, synthetic.MWWATERCOEFF_FILE)))
    ir = orc.IRwater(coeffio.read_irwatercoeff(os.path.join(coeff_dir, synthetic.IRWATERCOEFF_FILE)))
    K = 100
    results = {}
    for sensor, M in (("atms_syn", 16384), ("cris_syn", 4096)):
        synthetic.write_sensor(coeff_dir, sensor)
        ctx = api.crtm_init([sensor], coeff_dir)
        L = ctx.n_channels
        st = ctx.channel_info[0].sensor_type
        ps = synthetic.ocean_surface(synthetic.make_profiles(M, K, seed=20261203 + 9), seed=20261203 + 10)
        dps = api.device_profiles(ps, dev)
        emis = torch.as_tensor(np.tile(synthetic.default_emissivity(st, L), (M, 1)), device=dev)
        btk = torch.ones((M, L), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        r = {}
        for prec, pname, steps in ((api.PREC_FP32_SPLIT, "split", args.steps), (api.PREC_FP64, "fp64", args.fp64_steps)):
            for mode, e in (("computed", None), ("user", emis)):
                prev = ctx.k_matrix(dps, e, bt_k=btk, precision=prec, jacobian_f32=prec == api.PREC_FP32_SPLIT)
                for _ in range(args.warmup):
                    prev = ctx.k_matrix(dps, e, bt_k=btk, precision=prec, jacobian_f32=prec == api.PREC_FP32_SPLIT,
                                        out=prev)
                ctx.synchronize()
                ctx.profile_enable(True)
                t0 = time.perf_counter()
                for _ in range(steps):
                    prev = ctx.k_matrix(dps, e, bt_k=btk, precision=prec, jacobian_f32=prec == api.PREC_FP32_SPLIT,
                                        out=prev)
                ctx.synchronize()
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) / steps
                ms, _ = ctx.profile_query_kinds()
                ctx.profile_enable(False)
                r[f"{pname}_{mode}"] = {"ms_per_step": t * 1e3, "profile_channels_per_s": M * L / t,
                                        "kernel_ms": {"prep": ms[0] / steps, "clear": ms[1] / steps,
                                                      "surface_optics": ms[3] / steps}}
                log(f"  {sensor} {pname} {mode}: {t * 1e3:.2f} ms/step, kernels {r[f'{pname}_{mode}']['kernel_ms']}")
                if mode == "computed":
                    n = 8
                    rts, kd = prev
                    bt = rts["brightness_temperature"][:n].cpu().numpy()
                    tk = kd["temperature_k"].double().cpu().numpy()
                    wk = kd["wind_speed_k"][:n].cpu().numpy()
                    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{sensor}.TauCoeff.bin"))
                    spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{sensor}.SpcCoeff.bin"))
                    S = orc.Sensor(tau, spc)
                    bt_err = tk_err = wk_err = 0.0
                    for m in range(n):
                        so = orc.SurfaceOptics(spc, mw=mw if st == 1 else None, ir=ir if st != 1 else None,
                                               salinity=ps.salinity[m], wind_speed=ps.wind_speed[m],
                                               wind_direction=ps.wind_direction[m], sensor_azimuth=ps.sensor_azimuth[m])
                        o = orc.run_profile_sfc(S, ps, m, so, "k", bt_k=np.ones(L))
                        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
                        bt_err = max(bt_err, float(np.abs(bt[m] - o["bt"]).max()))
                        tk_err = max(tk_err, float(np.abs(tk[a:b].T - o["t_k"]).max() / np.abs(o["t_k"]).max()))
                        sw = np.abs(o["sfc_k"][:, 1]).max()
                        if sw > 0:
                            wk_err = max(wk_err, float(np.abs(wk[m] - o["sfc_k"][:, 1]).max() / sw))
                    r[f"{pname}_vs_oracle"] = {"bt_err_K": bt_err, "temperature_k_rel": tk_err,
                                               "wind_speed_k_rel": wk_err, "sample_profiles": n}
                del prev
                torch.cuda.empty_cache()
        for pname in ("split", "fp64"):
            r[f"{pname}_overhead"] = r[f"{pname}_computed"]["ms_per_step"] / r[f"{pname}_user"]["ms_per_step"]
        results[sensor] = {"n_profiles": M, "n_channels": L, **r}
        ctx.destroy()
    a = results["atms_syn"]
    line = {"metric": "clear-sky K-matrix profile-channels/s with computed water surface optics", "value":
            a["split_computed"]["profile_channels_per_s"], "unit": "profile-channels/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": a["split_computed"]["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64-split",
            "data": "synthetic (seeded ocean profiles; synthetic ODPS/SpcCoeff/EmisCoeff files)",
            "config": {"workload": "clear-sky K-matrix over an ocean, FASTEM6 (ATMS-like 22 ch x 16384 profiles) and "
                                   "the IR sea-surface emissivity table (CrIS-like 2211 ch x 4096 profiles), 100 layers",
                       "parallelism": "single GPU"},
            "by_sensor": results, "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def main_allsky(args):
    """Cloudy microwave forward + K-matrix: 22-channel ATMS-like sensor, 8192
    profiles x 100 layers with liquid, ice, rain and snow in 0-10 contiguous
    layers each (overcast), ADA at 4 and 8 streams; FP32-split vs FP64 GPU, BT
    and Jacobian errors against the FP64 oracle on a sample, oracle timed on the
    host cores."""
    import torch
    from crtmv3_amd import api, coeffio, synthetic
    from oracle import oracle as orc
    dev = torch.device("cuda:0")
    coeff_dir = synthetic.default_coeff_dir()
    sensor = "atms_syn"
    synthetic.write_sensor(coeff_dir, sensor)
    synthetic.write_cloud_coeff(coeff_dir)
    ctx = api.crtm_init([sensor], coeff_dir)
    if args.wavefront_angles:
        ctx.set_wavefront_angles(args.wavefront_angles)
    L = ctx.n_channels
    M = args.profiles if args.profiles != 4096 else 8192
    K = 100
    log(f"allsky: {sensor} {L} channels, {M} cloudy profiles")
    ps = synthetic.add_clouds(synthetic.make_profiles(M, K, seed=20261203 + 3), seed=20261203 + 4)
    dps = api.device_profiles(ps, dev)
    emis_h = np.tile(synthetic.default_emissivity(1, L), (M, 1))
    emis = torch.as_tensor(emis_h, device=dev)
    btk = torch.ones((M, L), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    by_streams = {}
    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{sensor}.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{sensor}.SpcCoeff.bin"))
    S = orc.Sensor(tau, spc)
    C = orc.CloudCoeff(coeffio.read_cloudcoeff(os.path.join(coeff_dir, synthetic.CLOUDCOEFF_FILE)))
    threads = host_threads()
    n_sample = 256     # forward BT check
    n_ksample = 32     # K-matrix check (hand-written oracle adjoint)
    sub = _subset(ps, 0, n_sample)
    ksub = _subset(ps, 0, n_ksample)
    for ns in (4, 8):
        res = {}
        for prec, name, steps in ((api.PREC_FP32_SPLIT, "split", args.steps), (api.PREC_FP64, "fp64", args.fp64_steps)):
            r = {}
            for mode in ("forward", "k_matrix"):
                if mode == "forward":
                    def run():
                        return ctx.forward(dps, emis, precision=prec, n_streams=ns), None
                else:
                    prev = ctx.k_matrix(dps, emis, bt_k=btk, precision=prec, jacobian_f32=prec == api.PREC_FP32_SPLIT,
                                        n_streams=ns)

                    def run():
                        return ctx.k_matrix(dps, emis, bt_k=btk, precision=prec,
                                            jacobian_f32=prec == api.PREC_FP32_SPLIT, n_streams=ns, out=prev)
                out = run()
                for _ in range(args.warmup):
                    out = run()
                ctx.synchronize()
                ctx.profile_enable(True)
                t0 = time.perf_counter()
                for _ in range(steps):
                    out = run()
                ctx.synchronize()
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) / steps
                ms, nl = ctx.profile_query_kinds()
                ctx.profile_enable(False)
                r[mode] = {"ms_per_step": t * 1e3, "profile_channels_per_s": M * L / t,
                           "kernel_ms": {"prep": ms[0] / steps, "clear": ms[1] / steps, "scattering": ms[2] / steps}}
                rts, kd = out
                r[mode + "_bt"] = rts["brightness_temperature"].cpu().numpy()
                if kd is not None:
                    r["k"] = {k: kd[k].double().cpu().numpy() for k in
                              ("temperature_k", "cloud_water_content_k", "cloud_effective_radius_k")}
                log(f"  {ns} streams {name} {mode}: {t * 1e3:.2f} ms/step  kernels {r[mode]['kernel_ms']}")
            res[name] = r
        _, bt_or = orc.forward_allsky_batch(S, C, sub, emis_h[:n_sample], n_streams=ns, n_threads=threads)
        kbt_or, tk_or, wck_or = orc.k_matrix_allsky_batch(S, C, ksub, emis_h[:n_ksample], n_streams=ns,
                                                         n_threads=threads)
        nk = int(ps.layer_offset[n_ksample])
        nc = ps.cloud_type.shape[1]

        def kerr(prec_name):   # largest relative error per profile and block, vs the oracle adjoint
            kk = res[prec_name]["k"]
            worst = {"T": 0.0, "water_content": 0.0}
            for m in range(n_ksample):
                a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
                want = tk_or[a * L:b * L].reshape(L, b - a)
                got = kk["temperature_k"][a:b].T
                worst["T"] = max(worst["T"], float(np.abs(got - want).max() / np.abs(want).max()))
                want = wck_or[a * nc * L:b * nc * L].reshape(L, nc, b - a)
                got = kk["cloud_water_content_k"][:, a:b, :].transpose(2, 0, 1)
                sc = np.abs(want).max()
                if sc > 0:
                    worst["water_content"] = max(worst["water_content"], float(np.abs(got - want).max() / sc))
            return worst
        jac_split_vs_fp64 = {}
        for key in ("temperature_k", "cloud_water_content_k", "cloud_effective_radius_k"):
            a_, b_ = res["split"]["k"][key], res["fp64"]["k"][key]
            jac_split_vs_fp64[key] = float(np.abs(a_ - b_).max() / max(np.abs(b_).max(), 1e-300))
        by_streams[ns] = {
            "split": {m_: res["split"][m_] for m_ in ("forward", "k_matrix")},
            "fp64": {m_: res["fp64"][m_] for m_ in ("forward", "k_matrix")},
            "speedup_split_over_fp64_k": res["fp64"]["k_matrix"]["ms_per_step"] / res["split"]["k_matrix"]["ms_per_step"],
            "speedup_split_over_fp64_forward": res["fp64"]["forward"]["ms_per_step"] / res["split"]["forward"]["ms_per_step"],
            "bt_err_split_vs_fp64_K": float(np.abs(res["split"]["forward_bt"] - res["fp64"]["forward_bt"]).max()),
            "bt_err_split_vs_oracle_K": float(np.abs(res["split"]["forward_bt"][:n_sample] - bt_or).max()),
            "bt_err_fp64_vs_oracle_K": float(np.abs(res["fp64"]["forward_bt"][:n_sample] - bt_or).max()),
            "k_bt_err_fp64_vs_oracle_K": float(np.abs(res["fp64"]["k_matrix_bt"][:n_ksample] - kbt_or).max()),
            "jacobian_rel_err_split_vs_fp64": jac_split_vs_fp64,
            "jacobian_rel_err_fp64_vs_oracle": kerr("fp64"),
            "jacobian_rel_err_split_vs_oracle": kerr("split"),
        }
        _ = nk
    cpu = None
    if not args.no_cpu_baseline:
        n = 32
        t0 = time.perf_counter()
        orc.k_matrix_allsky_batch(S, C, _subset(ps, 0, n), emis_h[:n], n_streams=8, n_threads=threads)
        dt = time.perf_counter() - t0
        n = min(M, max(n, int(n * args.cpu_seconds / max(dt, 1e-6))))
        t0 = time.perf_counter()
        orc.k_matrix_allsky_batch(S, C, _subset(ps, 0, n), emis_h[:n], n_streams=8, n_threads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": n * L / dt, "unit": "profile-channels/s", "cores": threads, "kind": "port",
               "host_cpus": os.cpu_count(),
               "sample": f"{n} of the {M} profiles x {L} channels, all-sky forward + K-matrix at 8 streams, FP64 "
                         f"oracle hand-written adjoint (oracle_k_matrix_allsky_ad_batch), {dt:.1f} s"}
    v = by_streams[8]["split"]["k_matrix"]["profile_channels_per_s"]
    line = {"metric": "all-sky forward+K-matrix profile-channels/s, FP32-split vs FP64 GPU", "value": v,
            "unit": "profile-channels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": by_streams[8]["split"]["k_matrix"]["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64-split",
            "data": "synthetic (seeded cloudy profiles; synthetic ODPS/SpcCoeff/CloudCoeff files)",
            "config": {"workload": f"all-sky microwave forward + K-matrix, ATMS-like 22-ch ODPS group 3, {M} profiles "
                                   f"x {K} layers, liquid/ice/rain/snow in 0-10 layers each, ADA 4 and 8 streams",
                       "n_profiles": M, "n_channels": L, "n_layers": K, "parallelism": "single GPU"},
            "by_streams": by_streams, "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    ctx.destroy()


def main_allsky_ir(args):
    """Infrared all-sky forward + K-matrix: 8461-channel IASI-like sensor,
    1024 profiles x 100 layers, 40 % of them with liquid and ice clouds as in
    the microwave workload (overcast), ADA at 8 streams; FP32-split vs FP64
    GPU (Jacobians stored in FP32 in both), BT and Jacobian errors against the
    FP64 oracle adjoint on a sample of cloudy profiles, oracle timed on the
    host cores."""
    import torch
    from crtmv3_amd import api, coeffio, synthetic
    from oracle import oracle as orc
    dev = torch.device("cuda:0")
    coeff_dir = synthetic.default_coeff_dir()
    sensor = "iasi_syn"
    synthetic.write_sensor(coeff_dir, sensor)
    synthetic.write_cloud_coeff(coeff_dir)
    ctx = api.crtm_init([sensor], coeff_dir)
    if args.wavefront_angles:
        ctx.set_wavefront_angles(args.wavefront_angles)
    L = ctx.n_channels
    M = args.profiles if args.profiles != 4096 else 1024
    K, ns = 100, args.streams
    ps = synthetic.add_clouds(synthetic.make_profiles(M, K, seed=20261203 + 5), seed=20261203 + 6,
                              types=(synthetic.WATER_CLOUD, synthetic.ICE_CLOUD), cloudy_fraction=0.4)
    n_cloudy = int(sum(ps.cloud_water_content[:, int(ps.layer_offset[m]):int(ps.layer_offset[m + 1])].max() > 0
                       for m in range(M)))
    log(f"allsky-ir: {sensor} {L} channels, {M} profiles ({n_cloudy} cloudy)")
    dps = api.device_profiles(ps, dev)
    emis_h = np.tile(synthetic.default_emissivity(1, L), (M, 1))
    emis = torch.as_tensor(emis_h, device=dev)
    btk = torch.ones((M, L), dtype=torch.float64, device=dev)
    res = {}
    for prec, name, steps in ((api.PREC_FP32_SPLIT, "split", args.steps), (api.PREC_FP64, "fp64", args.fp64_steps)):
        r = {}
        for mode in ("forward", "k_matrix"):
            if mode == "forward":
                def run():
                    return ctx.forward(dps, emis, precision=prec, n_streams=ns), None
            else:
                prev = ctx.k_matrix(dps, emis, bt_k=btk, precision=prec, jacobian_f32=True, n_streams=ns)

                def run():
                    return ctx.k_matrix(dps, emis, bt_k=btk, precision=prec, jacobian_f32=True, n_streams=ns,
                                        out=prev)
            out = run()
            for _ in range(args.warmup):
                out = run()
            ctx.synchronize()
            ctx.profile_enable(True)
            t0 = time.perf_counter()
            for _ in range(steps):
                out = run()
            ctx.synchronize()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / steps
            ms, _ = ctx.profile_query_kinds()
            ctx.profile_enable(False)
            r[mode] = {"ms_per_step": t * 1e3, "profile_channels_per_s": M * L / t,
                       "kernel_ms": {"prep": ms[0] / steps, "clear": ms[1] / steps, "scattering": ms[2] / steps}}
            rts, kd = out
            r[mode + "_bt"] = rts["brightness_temperature"].cpu().numpy()
            if kd is not None:
                sel = [m for m in range(M)
                       if ps.cloud_water_content[:, int(ps.layer_offset[m]):int(ps.layer_offset[m + 1])].max() > 0][:2]
                r["k_sel"] = sel
                r["k"] = {key: [kd[key][..., int(ps.layer_offset[m]):int(ps.layer_offset[m + 1]), :].double()
                                .cpu().numpy() for m in sel]
                          for key in ("temperature_k", "cloud_water_content_k")}
            log(f"  {name} {mode}: {t * 1e3:.2f} ms/step  kernels {r[mode]['kernel_ms']}")
            del out
        res[name] = r
    # oracle on a sample of cloudy profiles: forward BT and the hand-written adjoint
    tau = coeffio.read_odps(os.path.join(coeff_dir, f"{sensor}.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(coeff_dir, f"{sensor}.SpcCoeff.bin"))
    S = orc.Sensor(tau, spc)
    C = orc.CloudCoeff(coeffio.read_cloudcoeff(os.path.join(coeff_dir, synthetic.CLOUDCOEFF_FILE)))
    sel = res["fp64"]["k_sel"]
    worst = {"split": {"T": 0.0, "water_content": 0.0}, "fp64": {"T": 0.0, "water_content": 0.0}}
    bt_err = {"split": 0.0, "fp64": 0.0}
    for i, m in enumerate(sel):
        f, k = orc.k_matrix_allsky(S, C, ps, m, emis_h[m], ns, 0, np.ones(L), None, method="adjoint")
        for name in ("split", "fp64"):
            bt_err[name] = max(bt_err[name], float(np.abs(res[name]["forward_bt"][m] - f["bt"]).max()))
            got = res[name]["k"]["temperature_k"][i].T
            worst[name]["T"] = max(worst[name]["T"], float(np.abs(got - k["t_k"]).max() / np.abs(k["t_k"]).max()))
            got = res[name]["k"]["cloud_water_content_k"][i].transpose(2, 0, 1)
            sc = np.abs(k["wc_k"]).max()
            if sc > 0:
                worst[name]["water_content"] = max(worst[name]["water_content"],
                                                   float(np.abs(got - k["wc_k"]).max() / sc))
    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        n = 2
        t0 = time.perf_counter()
        orc.k_matrix_allsky_batch(S, C, _subset(ps, 0, n), emis_h[:n], n_streams=ns, n_threads=threads)
        dt = time.perf_counter() - t0
        n = min(M, max(n, int(n * args.cpu_seconds / max(dt, 1e-6))))
        t0 = time.perf_counter()
        orc.k_matrix_allsky_batch(S, C, _subset(ps, 0, n), emis_h[:n], n_streams=ns, n_threads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": n * L / dt, "unit": "profile-channels/s", "cores": threads, "kind": "port",
               "host_cpus": os.cpu_count(),
               "sample": f"{n} of the {M} profiles x {L} channels (the first ones; 40 % cloudy overall), "
                         f"all-sky forward + K-matrix at {ns} streams, FP64 oracle hand-written adjoint, {dt:.1f} s"}
    v = res["split"]["k_matrix"]["profile_channels_per_s"]
    line = {"metric": "all-sky IR forward+K-matrix profile-channels/s, FP32-split vs FP64 GPU", "value": v,
            "unit": "profile-channels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res["split"]["k_matrix"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32+fp64-split",
            "data": "synthetic (seeded profiles; synthetic ODPS/SpcCoeff/CloudCoeff files)",
            "config": {"workload": f"all-sky infrared forward + K-matrix, IASI-like {L}-ch ODPS group 1, {M} profiles "
                                   f"x {K} layers, {n_cloudy} with liquid and ice clouds in 0-10 layers each, "
                                   f"ADA {ns} streams, FP32 Jacobians",
                       "n_profiles": M, "n_channels": L, "n_layers": K, "n_streams": ns,
                       "wavefront_angles": args.wavefront_angles or 9, "parallelism": "single GPU"},
            "split": {m_: res["split"][m_] for m_ in ("forward", "k_matrix")},
            "fp64": {m_: res["fp64"][m_] for m_ in ("forward", "k_matrix")},
            "speedup_split_over_fp64_k": res["fp64"]["k_matrix"]["ms_per_step"] / res["split"]["k_matrix"]["ms_per_step"],
            "bt_err_split_vs_fp64_K": float(np.abs(res["split"]["forward_bt"] - res["fp64"]["forward_bt"]).max()),
            "bt_err_vs_oracle_K": bt_err, "jacobian_rel_err_vs_oracle": worst,
            "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    ctx.destroy()


def main_batched(args):
    """Batched all-sky variational gradient: the ATMS-, CrIS- and IASI-like
    sensors in one context (10694 channels), profiles of 60-137 layers,
    liquid / ice / rain / snow clouds in 40 % of them, dust / sea-salt /
    sulfate aerosols in the lowest 20 layers of the profiles, Options(m)%
    n_Streams of 2 / 4 / 8 / 16 by profile (2 only where there is no
    aerosol).  One step = J = sum((y - H(x))^2 / sigma^2) and its gradient
    with respect to temperature, absorbers, surface temperature, emissivity,
    cloud water content and aerosol concentration through torch.autograd
    (forward model + channel-summed adjoint), with y = H(x_truth) + N(0,
    0.2 K).  FP32-split arithmetic; the gradient is checked against a central
    difference of J along a random direction."""
    import torch
    from crtmv3_amd import api, synthetic
    from crtmv3_amd.autograd import crtm_bt
    dev = torch.device("cuda:0")
    coeff_dir = synthetic.default_coeff_dir()
    sensors = ["atms_syn", "cris_syn", "iasi_syn"]
    for s_ in sensors:
        synthetic.write_sensor(coeff_dir, s_)
    synthetic.write_cloud_coeff(coeff_dir)
    synthetic.write_aerosol_coeff(coeff_dir)
    ctx = api.crtm_init(sensors, coeff_dir)
    L = ctx.n_channels
    M = args.profiles if args.profiles != 4096 else 256
    rng = np.random.default_rng(20261203 + 9)
    ks = rng.integers(60, 138, M).tolist()
    ps = synthetic.add_clouds(synthetic.make_profiles(M, ks, seed=20261203 + 10), seed=20261203 + 11,
                              cloudy_fraction=0.4)
    ps = synthetic.add_aerosols(ps, seed=20261203 + 12, aerosol_fraction=0.75)
    has_aer = (ps.aerosol_type != 0).any(axis=1)
    ns = rng.choice([4, 8, 16], M)
    ns[~has_aer] = rng.choice([2, 4, 8, 16], int((~has_aer).sum()))
    ns = ns.astype(np.int32)
    dps = api.device_profiles(ps, dev)
    nL = [22, 2211, 8461]
    emis = torch.as_tensor(np.concatenate([np.full((M, nL[0]), 0.6), np.full((M, L - nL[0]), 0.98)], axis=1),
                           device=dev)
    prec = api.PREC_FP32_SPLIT
    # a progress line per profile chunk when a step takes long (full size)
    kw = dict(precision=prec, n_streams=ns, forward_chunk=256, max_work_bytes=32 << 30,
              progress=log if M > 2048 else None)
    # observations: H(x_truth) + N(0, 0.2 K), x_truth = x + 1 K temperature noise
    g = torch.Generator(device="cpu").manual_seed(11)
    with torch.no_grad():
        t_truth = dps.temperature + torch.randn(dps.temperature.shape, generator=g, dtype=torch.float64).to(dev)
        y = crtm_bt(ctx, dps, t_truth, dps.absorber, dps.water_temperature, emis, **kw)
        y = y + 0.2 * torch.randn(y.shape, generator=g, dtype=torch.float64).to(dev)
    log(f"batched: {L} channels x {M} profiles ({int(has_aer.sum())} with aerosols, "
        f"{int((ps.cloud_water_content.max(axis=0) > 0).sum())} cloudy layers), streams "
        f"{dict(zip(*np.unique(ns, return_counts=True)))}")

    def step():
        leaves = [dps.temperature.clone().requires_grad_(True), dps.absorber.clone().requires_grad_(True),
                  dps.water_temperature.clone().requires_grad_(True), emis.clone().requires_grad_(True),
                  dps.cloud_water_content.clone().requires_grad_(True),
                  dps.aerosol_concentration.clone().requires_grad_(True)]
        T, A, Ts, E, W, C = leaves
        bt = crtm_bt(ctx, dps, T, A, Ts, E, cloud_water_content=W, aerosol_concentration=C, **kw)
        J = (((y - bt) / 0.2) ** 2).sum()
        J.backward()
        return J.detach(), leaves
    J, leaves = step()
    for _ in range(args.warmup):
        J, leaves = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, leaves = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    # gradient check: directional derivative along a random temperature direction
    d = torch.randn(dps.temperature.shape, generator=g, dtype=torch.float64).to(dev)
    h = 0.05   # K: large against the FP32-split rounding of the BT, small against the curvature of J
    with torch.no_grad():
        def jcost(tt):
            bt = crtm_bt(ctx, dps, tt, dps.absorber, dps.water_temperature, emis,
                         cloud_water_content=dps.cloud_water_content,
                         aerosol_concentration=dps.aerosol_concentration, **kw)
            return float((((y - bt) / 0.2) ** 2).sum())
        fd = (jcost(dps.temperature + h * d) - jcost(dps.temperature - h * d)) / (2 * h)
    ad = float((leaves[0].grad * d).sum())
    line = {"metric": "batched all-sky gradient of a variational cost, profile-channels/s (forward + adjoint)",
            "value": M * L / t, "unit": "profile-channels/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32+fp64-split",
            "data": "synthetic (seeded profiles; synthetic ODPS/SpcCoeff/CloudCoeff/AerosolCoeff files)",
            "config": {"workload": f"three sensors ({L} channels), {M} profiles of 60-137 layers, 40 % cloudy, "
                                   f"aerosols in 75 %, streams 2/4/8/16 per profile, J and dJ/dx by autograd",
                       "n_profiles": M, "n_channels": L, "parallelism": "single GPU"},
            "J": float(J), "gradient_check": {"directional_adjoint": ad, "central_difference": fd,
                                              "rel_err": abs(ad - fd) / abs(fd)},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    ctx.destroy()


if __name__ == "__main__":
    main()
