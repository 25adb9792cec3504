"""Benchmark of the MI355X tile-matching path (one GPU).

A step is one plain mosaic job on a device-resident library and target:
cell Lab means -> K nearest materials per cell -> placement chain -> SA ->
resize + colour adjustment + composite (mh_generate_device), submitted and
waited for one at a time (the plain one-job-at-a-time GPU form).  The
workload is the 1920x1080 / 1,500-material / 80x60 job with 2,000 SA
proposals.  Prints one JSON line.  The placement chain's kernel time (the
roofline's) comes from HIP events recorded on the library's stream during
the timed region; the other phases' times from a second pass with every
phase bracketed by events.
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


def cpu_baseline_all_cores(oracle, p, target, mats, budget_s, threads):
    """The same whole jobs on `threads` host threads, one job per thread at a
    time (jobs are independent; each is the single-threaded restatement, whose
    C calls release the GIL).  Returns (jobs/s, jobs, seconds)."""
    import concurrent.futures
    deadline = time.perf_counter() + budget_s
    t0 = time.perf_counter()

    def worker(w):
        done = 0
        while time.perf_counter() < deadline:
            oracle.generate(target, mats, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                            p["color"], True, p["iters"], seed=1000 * w + done)
            done += 1
        return done

    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        jobs = sum(ex.map(worker, range(threads)))
    el = time.perf_counter() - t0
    return jobs / el, jobs, el


def _pct(xs, q):
    return round(float(np.percentile(np.asarray(xs, np.float64), q)), 4) if len(xs) else None


def run_ingest(mh, synth, ctx, reps, oracle=None):
    """Material ingest: 10,000 materials of 64x64 (one size), then 10,000 of
    four sizes (48x48, 64x64, 96x64, 128x96) interleaved, into resident
    libraries (pixels + Lab means).  Device time is the Lab kernel's; the
    end-to-end rate includes host packing and the PCIe upload."""
    rng = np.random.default_rng(11)
    uniform = list(synth.flat_materials(10000, 64, rng))
    sizes = [(48, 48), (64, 64), (96, 64), (128, 96)]
    by_size = [synth.flat_materials(2500, 1, rng) for _ in sizes]  # colours only
    mixed = []
    for i in range(10000):
        w, h = sizes[i % 4]
        c = by_size[i % 4][i // 4, 0, 0].astype(np.int16)
        noise = rng.integers(-12, 13, (h, w, 3), dtype=np.int16)
        mixed.append(np.clip(c + noise, 0, 255).astype(np.uint8))
    out = {}
    for name, mats in (("uniform_64", uniform), ("mixed_4_sizes", mixed)):
        px = sum(m.shape[0] * m.shape[1] for m in mats)
        ctx.library_from_images(mats[:64]).close()  # warm
        ctx.set_timing(True)
        ctx.timing()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.library_from_images(mats).close()
        el = (time.perf_counter() - t0) / reps
        ph = ctx.timing()["cell_lab"]
        ctx.set_timing(False)
        k_ms = ph[0] / max(ph[1], 1)
        out[name] = {"materials": len(mats), "pixels": px,
                     "lab_kernel_ms": round(k_ms, 4),
                     "lab_pixels_per_s": round(px / (k_ms * 1e-3), 1),
                     "lab_hbm_gbs": round(3.0 * px / (k_ms * 1e-3) / 1e9, 2),
                     "end_to_end_ms": round(1000 * el, 2),
                     "materials_per_s_end_to_end": round(len(mats) / el, 1)}
    if oracle is not None:  # the C restatement of calculate_average_lab, one core
        sample = uniform[:300]
        t0 = time.perf_counter()
        for m in sample:
            oracle.average_lab(m)
        el = time.perf_counter() - t0
        out["cpu_port_materials_per_s"] = round(len(sample) / el, 1)
        out["cpu_sample"] = "300 materials of 64x64, one core"
    return out


def run_batched(mh, synth, ctx, steps, warmup):
    """64 independent jobs (own libraries, grids 20-160 x 15-120 at 16 px)
    plain one at a time, then as one fused batch replayed from a hipGraph."""
    jobs = synth.make_batch(64)
    libs, d_t, d_g, d_i, d_g2, d_i2, prm = [], [], [], [], [], [], []
    for k, (p, target, mats) in enumerate(jobs):
        libs.append(ctx.library_from_pixels(mats))
        d_t.append(ctx.to_device(target))
        cells = p["grid_w"] * p["grid_h"]
        img = p["width"] * p["height"] * 3
        d_g.append(ctx.device_alloc(cells * 4))
        d_i.append(ctx.device_alloc(img))
        d_g2.append(ctx.device_alloc(cells * 4))
        d_i2.append(ctx.device_alloc(img))
        prm.append(mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"], 1,
                                p["iters"], 100.0, 0.99995, k))

    def plain_pass(lat=None):
        for k, (p, _, _) in enumerate(jobs):
            t0 = time.perf_counter()
            ctx.generate_device(libs[k], d_t[k], p["width"], p["height"], prm[k], d_g[k], d_i[k])
            ctx.synchronize()
            if lat is not None:
                lat.append(1000.0 * (time.perf_counter() - t0))

    plain_pass()  # warm: library caches (resized tiles, term matrices)
    plain_lat = []
    t0 = time.perf_counter()
    plain_pass(plain_lat)
    plain_s = time.perf_counter() - t0

    b = ctx.batch([ctx.job(libs[k], d_t[k], p["width"], p["height"], prm[k], d_g2[k], d_i2[k])
                   for k, (p, _, _) in enumerate(jobs)])
    b.set_graph(True)
    default_shapes = b.tuning()
    tuned = b.tune(repeats=3)  # launch shapes timed on this batch; the graph captures them
    for _ in range(max(1, warmup)):
        b.run()
        ctx.synchronize()
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        b.run()
        ctx.synchronize()
        step_ms.append(1000.0 * (time.perf_counter() - t1))
    batch_s = time.perf_counter() - t0
    same = True
    for k, (p, _, _) in enumerate(jobs):
        cells = p["grid_w"] * p["grid_h"]
        same &= np.array_equal(ctx.to_host(d_g[k], cells, np.int32), ctx.to_host(d_g2[k], cells, np.int32))
        n_img = p["width"] * p["height"] * 3
        same &= np.array_equal(ctx.to_host(d_i[k], n_img, np.uint8), ctx.to_host(d_i2[k], n_img, np.uint8))
    b.close()
    cells = sum(p["grid_w"] * p["grid_h"] for p, _, _ in jobs)
    for ptr in d_t + d_g + d_i + d_g2 + d_i2:
        ctx.device_free(ptr)
    for lib in libs:
        lib.close()
    n = len(jobs)
    return {
        "workload": f"{n} independent jobs, own libraries of 256-4,096 materials 64x64, grids "
                    f"20-160 x 15-120 at 16 px ({cells} cells), usage 3-5, 1,000 SA proposals each",
        "jobs_per_s": round(n * steps / batch_s, 2),
        "step_ms": {"p50": _pct(step_ms, 50), "p99": _pct(step_ms, 99)},
        "tiles_placed_per_s": round(cells * steps / batch_s, 1),
        "plain_jobs_per_s": round(n / plain_s, 2),
        "plain_job_ms": {"p50": _pct(plain_lat, 50), "p99": _pct(plain_lat, 99)},
        "speedup_vs_plain": round((n * steps / batch_s) / (n / plain_s), 2),
        "outputs_identical_to_plain": bool(same),
        "graph": True,
        "tuning": {"default": {k: default_shapes[k] for k in ("topk_cells_per_wave", "composite_threads")},
                   "tuned": tuned},
    }


def run_served(mh, synth, ctx, n_jobs, rate, budget, streams, max_batch):
    """The Poisson job stream through the server, and the same stream through
    the plain path one job at a time (a job starts when it has arrived and the
    previous one is done).  Latency = completion - arrival."""
    st = synth.make_stream(n_jobs, rate)
    jobs = st["jobs"]
    W = st["canvas_size"][0]
    libs = [ctx.library_from_pixels(m) for m in st["libs"]]
    d_can = [ctx.to_device(c) for c in st["canvases"]]
    d_tg = []
    for j in jobs:  # each job's target: a contiguous crop of a canvas, on the device
        ptr = ctx.device_alloc(j["width"] * j["height"] * 3)
        ctx.copy_2d_on_device(ptr, j["width"] * 3, d_can[j["canvas"]] + (j["oy"] * W + j["ox"]) * 3,
                              W * 3, j["width"] * 3, j["height"])
        d_tg.append(ptr)
    ctx.synchronize()
    prm = [mh.JobParams(j["grid_w"], j["grid_h"], j["max_usage"], j["weight"], j["color"], 1,
                        j["iters"], 100.0, 0.99995, j["seed"]) for j in jobs]
    sample = list(range(0, n_jobs, max(1, n_jobs // 100)))
    d_gs = {k: ctx.device_alloc(jobs[k]["grid_w"] * jobs[k]["grid_h"] * 4) for k in sample}
    max_img = max(j["width"] * j["height"] * 3 for j in jobs)
    max_cells = max(j["grid_w"] * j["grid_h"] for j in jobs)
    d_img = ctx.device_alloc(max_img)
    d_grid = ctx.device_alloc(max_cells * 4)
    # warm the library caches (resized tiles, term matrices) once per library
    for li in range(len(libs)):
        k = next(i for i, j in enumerate(jobs) if j["lib"] == li)
        ctx.generate_device(libs[li], d_tg[k], jobs[k]["width"], jobs[k]["height"], prm[k], d_grid, d_img)
    ctx.synchronize()

    mj = [ctx.job(libs[j["lib"]], d_tg[k], j["width"], j["height"], prm[k], d_gs.get(k), None)
          for k, j in enumerate(jobs)]

    def serve(arrivals):
        """Submits job k at t0 + arrivals[k]; returns (times, stats)."""
        srv = ctx.server(memory_budget=budget, max_batch=max_batch, streams=streams)
        tickets = []
        t0 = time.perf_counter()
        for k in range(len(jobs)):
            due = t0 + arrivals[k]
            while time.perf_counter() < due:
                pass
            tickets.append(srv.submit(mj[k]))
        srv.drain()
        times = np.array([srv.times(t) for t in tickets])
        stats = srv.stats()
        srv.close()
        return times, stats

    # served at the stream's arrival rate, then with every job queued at once
    times, stats = serve([j["arrival"] for j in jobs])
    lat = times[:, 2] - times[:, 0]
    span = (times[:, 2].max() - times[:, 0].min()) / 1000.0
    ctimes, cstats = serve([0.0] * len(jobs))
    clat = ctimes[:, 2] - ctimes[:, 0]
    cspan = (ctimes[:, 2].max() - ctimes[:, 0].min()) / 1000.0
    served_grids = {k: ctx.to_host(d_gs[k], jobs[k]["grid_w"] * jobs[k]["grid_h"], np.int32)
                    for k in sample}

    # plain, same arrivals
    plain_lat = []
    same = True
    t0 = time.perf_counter()
    last = 0.0
    for k, j in enumerate(jobs):
        due = t0 + j["arrival"]
        while time.perf_counter() < due:
            pass
        ctx.generate_device(libs[j["lib"]], d_tg[k], j["width"], j["height"], prm[k], d_grid, d_img)
        ctx.synchronize()
        last = time.perf_counter()
        plain_lat.append(1000.0 * (last - due))
        if k in d_gs:
            cells = j["grid_w"] * j["grid_h"]
            same &= np.array_equal(ctx.to_host(d_grid, cells, np.int32), served_grids[k])
    plain_span = last - (t0 + jobs[0]["arrival"])
    for ptr in d_tg + list(d_gs.values()) + d_can + [d_img, d_grid]:
        ctx.device_free(ptr)
    for lib in libs:
        lib.close()
    cells = sum(j["grid_w"] * j["grid_h"] for j in jobs)
    served_jps = n_jobs / span
    capacity_jps = n_jobs / cspan
    plain_jps = n_jobs / plain_span
    return {
        "workload": f"{n_jobs} Poisson jobs at {rate:g}/s, cells log-uniform 300-24,000 "
                    f"({cells} total), 4:3 / 16:9 at 16 px, libraries 1k/2.5k/5k/10k of 64x64, "
                    f"usage 3-5, 1,000 SA proposals each",
        "jobs_per_s": round(served_jps, 2),
        "latency_ms": {"p50": _pct(lat, 50), "p99": _pct(lat, 99)},
        "capacity_jobs_per_s": round(capacity_jps, 2),
        "capacity_latency_ms": {"p50": _pct(clat, 50), "p99": _pct(clat, 99)},
        "plain_jobs_per_s": round(plain_jps, 2),
        "plain_latency_ms": {"p50": _pct(plain_lat, 50), "p99": _pct(plain_lat, 99)},
        "speedup_vs_plain": round(capacity_jps / plain_jps, 2),
        "speedup_at_rate": round(served_jps / plain_jps, 2),
        "tiles_placed_per_s": round(cells / cspan, 1),
        "placements_identical_to_plain": bool(same),
        "checked_jobs": len(sample),
        "batches": stats["batches"],
        "mean_jobs_per_batch": round(stats["submitted"] / max(stats["batches"], 1), 2),
        "max_jobs_per_batch": stats["max_batch_jobs"],
        "capacity_batches": cstats["batches"],
        "memory_peak_gib": round(max(stats["memory_peak"], cstats["memory_peak"]) / 2**30, 3),
        "memory_budget_gib": round(stats["memory_budget"] / 2**30, 3),
        "streams": streams,
    }


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
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job batched form")
    ap.add_argument("--no-serve", action="store_true", help="skip the served job stream")
    ap.add_argument("--no-ingest", action="store_true", help="skip the material-ingest leg")
    ap.add_argument("--serve-jobs", type=int, default=5000)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-max-batch", type=int, default=256)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
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

    # Timed region: only the dominant kernel (the placement chain) is
    # bracketed by HIP events, for the roofline; events around every phase
    # would add about 46 us to each 1.4 ms job.
    ctx.set_timing_phases(["place"])
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
    place_phase = ctx.timing()["place"]
    ctx.set_timing(False)
    res = ctx.generate_device_result()
    # the other phases' kernel times: a second pass with every phase timed
    ctx.set_timing(True)
    ctx.timing()
    for _ in range(max(5, args.steps // 2)):
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()
    phases = ctx.timing()
    ctx.set_timing(False)
    phases["place"] = place_phase  # the live figure from the timed region

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
        peak = HBM_PEAK_GBS * 1e9 if unit == "B" else FP32_PEAK_TFLOPS * 1e12
        kernels[name] = {"avg_ms": round(ms, 5), "algorithmic": amount, "unit": unit,
                         "rate": rate, "bound": bound, "roofline_frac": round(rate / peak, 6)}
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

    if not args.no_batch:
        out["batched"] = run_batched(mh, synth, ctx, max(3, args.steps // 4), 1)
    if not args.no_serve:
        out["served"] = run_served(mh, synth, ctx, args.serve_jobs, args.serve_rate,
                                   int(args.serve_budget_gib * 2**30), args.serve_streams,
                                   args.serve_max_batch)

    orc = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.load()
    if not args.no_ingest:
        out["ingest"] = run_ingest(mh, synth, ctx, 3, orc)

    if not args.no_cpu:
        oracle = orc
        rate, jobs, el = cpu_baseline(oracle, p, target, mats, args.cpu_seconds)
        out["cpu_baseline"] = {"value": round(rate, 4), "unit": "jobs/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{jobs} whole {args.workload} jobs ({el:.1f} s) of the C "
                                         f"restatement on one host core "
                                         f"({platform.processor() or platform.machine()})"}
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        rate_all, jobs_all, el_all = cpu_baseline_all_cores(orc, p, target, mats, args.cpu_seconds,
                                                            threads)
        out["cpu_baseline_all_cores"] = {
            "value": round(rate_all, 4), "unit": "jobs/s", "cores": threads, "kind": "port",
            "sample": f"{jobs_all} whole {args.workload} jobs ({el_all:.1f} s), one job per "
                      f"thread on {threads} host threads"}
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
