"""Benchmark of the MI355X tile-matching path (one GPU).

A step is one plain mosaic job on a device-resident library and target:
cell Lab means -> K nearest materials per cell -> placement chain -> SA ->
resize + colour adjustment + composite (mh_generate_device), submitted and
waited for one at a time (the plain one-job-at-a-time GPU form).  The
workload is the 1920x1080 / 1,500-material / 80x60 job with 2,000 SA
proposals.  Prints one JSON line; per-phase kernel times come from HIP
events recorded on the library's stream during the timed region.
"""
import argparse
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mosaic-art-rust_amd")

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_work(phase, p, n):
    """(amount, unit, bound) of algorithmic work per launch of a phase."""
    cells = p["grid_w"] * p["grid_h"]
    tw, th = p["width"] // p["grid_w"], p["height"] // p["grid_h"]
    k = min(n, 100)
    if phase == "cell_lab":  # 3 B/px read + 12 B/cell written
        return cells * tw * th * 3 + cells * 12, "B", "hbm"
    if phase == "topk":  # 8 flops per (cell, material)
        return 8.0 * cells * n, "FLOP", "valu"
    if phase == "place":  # 16 B + 12*K B per tile placed
        return cells * (16 + 12 * k), "B", "hbm"
    if phase == "anneal":  # 104 B per proposal
        return p["iters"] * 104, "B", "hbm"
    if phase == "resize":  # material pixels read + tiles written
        return n * (p["msize"] ** 2 * 3 + tw * th * 3), "B", "hbm"
    if phase == "adjust":  # target read for cell averages + tile averages
        return cells * tw * th * 3 + n * tw * th * 3, "B", "hbm"
    if phase == "composite":  # 3 B tile read + 3 B written per output pixel
        return cells * tw * th * 6, "B", "hbm"
    if phase == "sim_matrix":
        return 4 * n * n, "B", "hbm"
    return 0, "B", "hbm"


def measured_traffic(phase, workload):
    """HBM bytes per launch of a phase's kernel from the committed counter
    summary (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch,
    each scaled by the factor a copy of known size reads), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["workloads"][workload][phase]
        return int(round((k["fetch_kib"] * t["fetch_factor"] + k["write_kib"] * t["write_factor"])
                         * 1024))
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(oracle, p, target, mats, budget_s):
    """The oracle (single-threaded C restatement of the reference's CPU path)
    running whole jobs of the same workload until ~budget_s has passed."""
    t0 = time.perf_counter()
    jobs = 0
    while True:
        oracle.generate(target, mats, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                        p["color"], True, p["iters"], seed=jobs)
        jobs += 1
        el = time.perf_counter() - t0
        if el >= budget_s or jobs >= 64:
            return jobs / el, jobs, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="medium", choices=["small", "medium", "large"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="also run a device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one mosaic job stream per GPU and has "
              "no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mh = _load("mosaic_hip", os.path.join(PKG, "mosaic_hip.py"))
    synth = _load("mosaic_synth", os.path.join(PKG, "synth.py"))

    p, target, mats = synth.make(args.workload)
    n = p["n"]
    cells = p["grid_w"] * p["grid_h"]
    tw, th = p["width"] // p["grid_w"], p["height"] // p["grid_h"]

    ctx = mh.Context(0)
    lib = ctx.library_from_pixels(mats)
    d_target = ctx.to_device(target)
    d_grid = ctx.device_alloc(cells * 4)
    d_image = ctx.device_alloc(p["grid_w"] * tw * p["grid_h"] * th * 3)
    prm = mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"], 1,
                       p["iters"], 100.0, 0.99995, 0)

    for _ in range(args.warmup):
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()

    ctx.set_timing(True)
    ctx.timing()  # clear
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    phases = ctx.timing()
    ctx.set_timing(False)
    res = ctx.generate_device_result()

    # similarity matrix of the resident library (build_similarities)
    ctx.set_timing(True)
    for _ in range(max(3, args.steps // 4)):
        lib.build_similarities()
    sim_phase = ctx.timing()["sim_matrix"]
    ctx.set_timing(False)

    if args.calibrate_copy > 0:
        src = ctx.device_alloc(args.calibrate_copy)
        dst = ctx.device_alloc(args.calibrate_copy)
        for _ in range(3):
            ctx.copy_on_device(dst, src, args.calibrate_copy)
        ctx.synchronize()
        ctx.device_free(src)
        ctx.device_free(dst)

    jobs_per_s = args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in phases.items()}
    place_ms = avg["place"] + avg["topk"]
    kernels = {}
    for name, ms in avg.items():
        if name == "sim_matrix" or ms <= 0:
            continue
        amount, unit, bound = algorithmic_work(name, p, n)
        rate = amount / (ms * 1e-3)
        kernels[name] = {"avg_ms": round(ms, 5), "algorithmic": amount, "unit": unit,
                         "rate": rate, "bound": bound}
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    d_amount, d_unit, d_bound = algorithmic_work(dominant, p, n)
    d_ms = kernels[dominant]["avg_ms"]
    traffic = measured_traffic(dominant, args.workload)
    if d_unit == "B":
        achieved = d_amount / (d_ms * 1e-3) / 1e9
        roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 3),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": traffic, "algorithmic_bytes": d_amount, "avg_ms": d_ms}
    else:
        achieved = d_amount / (d_ms * 1e-3) / 1e12
        roof = {"kernel": dominant, "bound": "mfma", "achieved": round(achieved, 4),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 6), "traffic": traffic, "avg_ms": d_ms}
    sim_ms = sim_phase[0] / max(sim_phase[1], 1)
    sim_gbs = 4.0 * n * n / (sim_ms * 1e-3) / 1e9

    out = {
        "metric": "jobs/s + p50/p99 latency served vs plain GPU; tiles placed/s, SA proposals/s",
        "value": round(jobs_per_s, 3),
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain GPU mosaic job ({args.workload}): {p['width']}x{p['height']} "
                               f"target, {n} materials {p['msize']}x{p['msize']}, grid "
                               f"{p['grid_w']}x{p['grid_h']}, usage {p['max_usage']}, weight "
                               f"{p['weight']}, colour {p['color']}, {p['iters']} SA proposals",
                   "cells": cells, "materials": n, "sa_proposals": p["iters"],
                   "parallelism": "1 GPU, one job at a time"},
        "latency_ms": {"p50": round(1000 * float(np.percentile(lat, 50)), 4),
                       "p99": round(1000 * float(np.percentile(lat, 99)), 4)},
        "tiles_placed_per_s": round(cells * jobs_per_s, 1),
        "place_chain_tiles_per_s": round(cells / (place_ms * 1e-3), 1) if place_ms > 0 else None,
        "sa_proposals_per_s": round(p["iters"] / (avg["anneal"] * 1e-3), 1) if avg["anneal"] > 0 else None,
        "similarity_pairs_per_s": round(n * n / (sim_ms * 1e-3), 1),
        "similarity_hbm_frac": round(sim_gbs / HBM_PEAK_GBS, 4),
        "kernels": kernels,
        "job_stats": {"place": res["place"], "anneal_final_cost": res["anneal"]["final_cost"],
                      "anneal_rounds": res["anneal"]["rounds"]},
        "roofline": roof,
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.load()
        rate, jobs, el = cpu_baseline(oracle, p, target, mats, args.cpu_seconds)
        out["cpu_baseline"] = {"value": round(rate, 4), "unit": "jobs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{jobs} whole {args.workload} jobs ({el:.1f} s) of the C "
                                         f"restatement on one host core "
                                         f"({platform.processor() or platform.machine()})"}
    else:
        out["cpu_baseline"] = None

    ctx.device_free(d_target)
    ctx.device_free(d_grid)
    ctx.device_free(d_image)
    lib.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
