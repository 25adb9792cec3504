"""Benchmark of the MI355X tile-matching path (one GPU).

Headline (`value`): the served form on the 5,000-job Poisson stream
(cells log-uniform 300-24,000, four resident libraries of 1k-10k materials).
A step is one pass of the whole stream through the continuous-batching
server with every job queued at once; `value` is jobs/s over the K timed
steps.  Beside it: p50/p99 job latency with the jobs arriving at rate lambda,
and the plain one-job-at-a-time GPU path on the same jobs (throughput and
latency), with placements checked identical.  The roofline is the served
placement chain's (the dominant kernel), timed with HIP events on the
server's stream slots during the timed steps.

Other legs in the same JSON line: the plain medium job (1920x1080, 1,500
materials, 80x60), the large job (4000x2400, 10,000 materials, 200x120,
20,000 SA proposals, and its 400 MB similarity matrix), the 64-job batch,
material ingest, and the CPU baselines (the oracle on a library prepared
once, one job at a time on one core and one job per thread on the host's
cores).  Prints one JSON line.
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


def cpu_jobs(oracle, jobs, budget_s, threads=1):
    """The oracle (single-threaded C restatement of the reference's CPU path)
    on jobs = [(target, prepared library, params dict)], taken in order and
    cycled, one job per thread at a time, until ~budget_s has passed.  The
    libraries were prepared before (material Lab means, tiles resized to the
    cell size), as the GPU's resident libraries are.  Returns (jobs/s, jobs,
    seconds, per-phase seconds summed over the jobs)."""
    import concurrent.futures
    import threading
    deadline = time.perf_counter() + budget_s
    lock = threading.Lock()
    nxt = [0]
    phases = {}

    def worker(_):
        done = 0
        mine = {}
        while time.perf_counter() < deadline:
            with lock:
                k = nxt[0]
                nxt[0] += 1
            target, lib, p = jobs[k % len(jobs)]
            oracle.generate_prepared(target, lib, p["grid_w"], p["grid_h"], p["max_usage"],
                                     p["weight"], p["color"], True, p["iters"], 100.0, 0.99995,
                                     p.get("seed", k), times=mine)
            done += 1
        with lock:
            for key, v in mine.items():
                phases[key] = phases.get(key, 0.0) + v
        return done

    t0 = time.perf_counter()
    if threads <= 1:
        n = worker(0)
    else:
        with concurrent.futures.ThreadPoolExecutor(threads) as ex:
            n = sum(ex.map(worker, range(threads)))
    el = time.perf_counter() - t0
    return n / el, n, el, phases


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


def stream_setup(mh, synth, ctx, n_jobs, rate):
    """The Poisson stream's libraries and canvases made resident, each job's
    target a contiguous device crop of its canvas, and the libraries'
    caches (resized tiles, term matrices) warmed by one plain job each."""
    st = synth.make_stream(n_jobs, rate)
    jobs = st["jobs"]
    W = st["canvas_size"][0]
    libs = [ctx.library_from_pixels(m) for m in st["libs"]]
    d_can = [ctx.to_device(c) for c in st["canvases"]]
    d_tg = []
    for j in jobs:
        ptr = ctx.device_alloc(j["width"] * j["height"] * 3)
        ctx.copy_2d_on_device(ptr, j["width"] * 3, d_can[j["canvas"]] + (j["oy"] * W + j["ox"]) * 3,
                              W * 3, j["width"] * 3, j["height"])
        d_tg.append(ptr)
    ctx.synchronize()
    prm = [mh.JobParams(j["grid_w"], j["grid_h"], j["max_usage"], j["weight"], j["color"], 1,
                        j["iters"], 100.0, 0.99995, j["seed"]) for j in jobs]
    max_img = max(j["width"] * j["height"] * 3 for j in jobs)
    max_cells = max(j["grid_w"] * j["grid_h"] for j in jobs)
    d_img = ctx.device_alloc(max_img)
    d_grid = ctx.device_alloc(max_cells * 4)
    for li in range(len(libs)):
        k = next(i for i, j in enumerate(jobs) if j["lib"] == li)
        ctx.generate_device(libs[li], d_tg[k], jobs[k]["width"], jobs[k]["height"], prm[k], d_grid,
                            d_img)
    ctx.synchronize()
    return dict(st=st, jobs=jobs, libs=libs, d_can=d_can, d_tg=d_tg, prm=prm, d_img=d_img,
                d_grid=d_grid)


def stream_free(ctx, S):
    for ptr in S["d_tg"] + S["d_can"] + [S["d_img"], S["d_grid"]]:
        ctx.device_free(ptr)
    for lib in S["libs"]:
        lib.close()


def run_served(mh, ctx, S, steps, warmup, budget, streams, max_batch, capacity_only=False):
    """The headline: K timed passes of the whole stream through one server
    (kept across the passes, as a service keeps it) with every job queued at
    once (capacity), after W untimed ones; then the
    stream at its Poisson arrival rate (latency); then the plain path on the
    same jobs one at a time (a job starts when it has arrived and the previous
    one is done).  Latency = completion - arrival.  The placement launches of
    the timed passes are bracketed by HIP events on the server's streams."""
    jobs, libs, prm, d_tg = S["jobs"], S["libs"], S["prm"], S["d_tg"]
    n_jobs = len(jobs)
    sample = list(range(0, n_jobs, max(1, n_jobs // 100)))
    d_gs = {k: ctx.device_alloc(jobs[k]["grid_w"] * jobs[k]["grid_h"] * 4) for k in sample}
    mj = [ctx.job(libs[j["lib"]], d_tg[k], j["width"], j["height"], prm[k], d_gs.get(k), None)
          for k, j in enumerate(jobs)]

    def serve(srv, arrivals, timed=False):
        """Submits job k at t0 + arrivals[k], or every job at once by one call
        when arrivals is None, and drains; returns (times, batches, timing)."""
        b0 = srv.stats()["batches"]
        srv.timing()  # clears what earlier passes left
        srv.set_timing_phases(["place"] if timed else [])
        t0 = time.perf_counter()
        if arrivals is None:  # every job queued at once, by one call
            tickets = srv.submit_many(mj)
        else:
            tickets = []
            for k in range(n_jobs):
                due = t0 + arrivals[k]
                while time.perf_counter() < due:
                    pass
                tickets.append(srv.submit(mj[k]))
        srv.drain()
        times = np.array([srv.times(t) for t in tickets])
        return times, srv.stats()["batches"] - b0, srv.timing()

    # One server for the warm-up and timed capacity passes (a service keeps
    # its server: streams, queues and working memory are made once).
    cap = ctx.server(memory_budget=budget, max_batch=max_batch, streams=streams)
    zero = None
    for _ in range(warmup):
        serve(cap, zero)
    spans, clat, place_ms, place_n, place_cells, cap_batches = [], [], 0.0, 0, 0, None
    ctx.synchronize()
    for _ in range(steps):
        ctimes, cap_batches, tm = serve(cap, zero, timed=True)
        spans.append((ctimes[:, 2].max() - ctimes[:, 0].min()) / 1000.0)
        clat.extend((ctimes[:, 2] - ctimes[:, 0]).tolist())
        place_ms += tm["place"][0]
        place_n += tm["place"][1]
        place_cells += tm["place"][2]
    cstats = cap.stats()
    cap.close()
    ctx.synchronize()
    served_grids = {k: ctx.to_host(d_gs[k], jobs[k]["grid_w"] * jobs[k]["grid_h"], np.int32)
                    for k in sample}
    cells = sum(j["grid_w"] * j["grid_h"] for j in jobs)
    iters = sum(j["iters"] for j in jobs)
    cap_s = float(np.sum(spans))
    capacity_jps = n_jobs * steps / cap_s
    if capacity_only:
        for ptr in d_gs.values():
            ctx.device_free(ptr)
        return {"jobs_per_s": capacity_jps, "step_s": spans,
                "capacity_latency_ms": {"p50": _pct(clat, 50), "p99": _pct(clat, 99)},
                "tiles_placed_per_s": round(cells * steps / cap_s, 1),
                "sa_proposals_per_s": round(iters * steps / cap_s, 1),
                "capacity_batches": cap_batches,
                "streams": streams,
                "place_timing": {"ms": place_ms, "launches": place_n, "cells": place_cells},
                "cells": cells}
    rsrv = ctx.server(memory_budget=budget, max_batch=max_batch, streams=streams)
    times, rate_batches, _ = serve(rsrv, [j["arrival"] for j in jobs])
    stats = rsrv.stats()
    rsrv.close()
    lat = times[:, 2] - times[:, 0]
    span = (times[:, 2].max() - times[:, 0].min()) / 1000.0

    # plain, same arrivals
    plain_lat = []
    same = True
    t0 = time.perf_counter()
    last = 0.0
    for k, j in enumerate(jobs):
        due = t0 + j["arrival"]
        while time.perf_counter() < due:
            pass
        ctx.generate_device(libs[j["lib"]], d_tg[k], j["width"], j["height"], prm[k], S["d_grid"],
                            S["d_img"])
        ctx.synchronize()
        last = time.perf_counter()
        plain_lat.append(1000.0 * (last - due))
        if k in d_gs:
            same &= np.array_equal(ctx.to_host(S["d_grid"], j["grid_w"] * j["grid_h"], np.int32),
                                   served_grids[k])
    plain_span = last - (t0 + jobs[0]["arrival"])
    for ptr in d_gs.values():
        ctx.device_free(ptr)
    plain_jps = n_jobs / plain_span
    return {
        "jobs_per_s": capacity_jps,
        "step_s": spans,
        "capacity_latency_ms": {"p50": _pct(clat, 50), "p99": _pct(clat, 99)},
        "rate": {"arrivals_per_s": S["rate"], "jobs_per_s": round(n_jobs / span, 2),
                 "latency_ms": {"p50": _pct(lat, 50), "p99": _pct(lat, 99)},
                 "batches": rate_batches,
                 "mean_jobs_per_batch": round(n_jobs / max(rate_batches, 1), 2)},
        "plain": {"jobs_per_s": round(plain_jps, 2),
                  "latency_ms": {"p50": _pct(plain_lat, 50), "p99": _pct(plain_lat, 99)}},
        "speedup_vs_plain": round(capacity_jps / plain_jps, 2),
        "speedup_at_rate": round((n_jobs / span) / plain_jps, 2),
        "tiles_placed_per_s": round(cells * steps / cap_s, 1),
        "sa_proposals_per_s": round(iters * steps / cap_s, 1),
        "placements_identical_to_plain": bool(same),
        "checked_jobs": len(sample),
        "capacity_batches": cap_batches,
        "max_jobs_per_batch": cstats["max_batch_jobs"],
        "memory_peak_gib": round(max(stats["memory_peak"], cstats["memory_peak"]) / 2**30, 3),
        "memory_budget_gib": round(stats["memory_budget"] / 2**30, 3),
        "streams": streams,
        "place_timing": {"ms": place_ms, "launches": place_n, "cells": place_cells},
        "cells": cells,
    }


def run_plain(mh, synth, ctx, name, steps, warmup):
    """One plain job at a time (mh_generate_device + wait) on a resident
    library and target.  Only the placement chain is bracketed by events in
    the timed region (events around every phase would add about 46 us to each
    job); the other phases' times come from a second pass with every phase
    timed."""
    p, target, mats = synth.make(name)
    n = p["n"]
    cells = p["grid_w"] * p["grid_h"]
    tw, th = p["width"] // p["grid_w"], p["height"] // p["grid_h"]
    lib = ctx.library_from_pixels(mats)
    d_target = ctx.to_device(target)
    d_grid = ctx.device_alloc(cells * 4)
    d_image = ctx.device_alloc(p["grid_w"] * tw * p["grid_h"] * th * 3)
    prm = mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"], 1,
                       p["iters"], 100.0, 0.99995, 0)
    for _ in range(max(1, warmup)):
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()
    ctx.set_timing_phases(["place"])
    ctx.timing()
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    elapsed = time.perf_counter() - t_start
    place_phase = ctx.timing()["place"]
    ctx.set_timing(False)
    res = ctx.generate_device_result()
    ctx.set_timing(True)
    ctx.timing()
    for _ in range(max(3, steps // 2)):
        ctx.generate_device(lib, d_target, p["width"], p["height"], prm, d_grid, d_image)
        ctx.synchronize()
    phases = ctx.timing()
    ctx.set_timing(False)
    phases["place"] = place_phase
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in phases.items()}
    kernels = {}
    for kname, ms in avg.items():
        if kname in ("sim_matrix", "resize") or ms <= 0:
            continue
        amount, unit, bound = algorithmic_work(kname, p, n)
        rate = amount / (ms * 1e-3)
        peak = HBM_PEAK_GBS * 1e9 if unit == "B" else FP32_PEAK_TFLOPS * 1e12
        kernels[kname] = {"avg_ms": round(ms, 5), "algorithmic": amount, "unit": unit,
                          "rate": rate, "bound": bound, "roofline_frac": round(rate / peak, 6)}
    out = {
        "workload": f"{p['width']}x{p['height']} target, {n} materials {p['msize']}x{p['msize']}, "
                    f"grid {p['grid_w']}x{p['grid_h']}, usage {p['max_usage']}, weight "
                    f"{p['weight']}, colour {p['color']}, {p['iters']} SA proposals",
        "jobs_per_s": round(steps / elapsed, 3),
        "ms_per_job": round(1000.0 * elapsed / steps, 4),
        "latency_ms": {"p50": round(1000 * float(np.percentile(lat, 50)), 4),
                       "p99": round(1000 * float(np.percentile(lat, 99)), 4)},
        "tiles_placed_per_s": round(cells * steps / elapsed, 1),
        "place_chain_tiles_per_s": round(cells / (avg["place"] * 1e-3), 1) if avg["place"] else None,
        "sa_proposals_per_s": round(p["iters"] / (avg["anneal"] * 1e-3), 1) if avg["anneal"] else None,
        "kernels": kernels,
        "job_stats": {"place": res["place"], "anneal_final_cost": res["anneal"]["final_cost"],
                      "anneal_rounds": res["anneal"]["rounds"]},
    }
    return out, (p, target, mats, lib, d_target, d_grid, d_image)


def run_similarity(ctx, lib, n, reps):
    """build_similarities of a resident library: the full n x n f32 matrix
    written to HBM (4 B per pair)."""
    lib.build_similarities()  # allocates the matrix
    ctx.set_timing_phases(["sim_matrix"])
    ctx.timing()
    for _ in range(reps):
        lib.build_similarities()
    ph = ctx.timing()["sim_matrix"]
    ctx.set_timing(False)
    ms = ph[0] / max(ph[1], 1)
    gbs = 4.0 * n * n / (ms * 1e-3) / 1e9
    return {"materials": n, "matrix_mb": round(4.0 * n * n / 1e6, 1), "avg_ms": round(ms, 4),
            "pairs_per_s": round(n * n / (ms * 1e-3), 1), "hbm_gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5,
                    help="timed passes of the served stream (the headline)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--legs", default="served,plain,large,batched,ingest,cpu",
                    help="comma list of served, plain, large, batched, ingest, cpu")
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="also run a device-to-device copy of BYTES (counter calibration)")
    ap.add_argument("--serve-jobs", type=int, default=5000)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-streams", type=int, default=8)
    ap.add_argument("--serve-max-batch", type=int, default=128)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--serve-capacity-only", action="store_true",
                    help="only the capacity passes of the served stream (counter runs)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs one mosaic job stream per GPU and has "
              "no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mh = _load("mosaic_hip", os.path.join(PKG, "mosaic_hip.py"))
    synth = _load("mosaic_synth", os.path.join(PKG, "synth.py"))
    orc = None
    if "cpu" in legs or "ingest" in legs:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        orc.load()
    cpu_threads = max(1, min(16, len(os.sched_getaffinity(0))))
    cpu_name = platform.processor() or platform.machine()

    ctx = mh.Context(0)
    out = {"metric": "jobs/s + p50/p99 latency served vs plain GPU; tiles placed/s, SA proposals/s",
           "value": None, "unit": "jobs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}

    S = None
    if "served" in legs:
        S = stream_setup(mh, synth, ctx, args.serve_jobs, args.serve_rate)
        S["rate"] = args.serve_rate
        sv = run_served(mh, ctx, S, args.steps, args.warmup, int(args.serve_budget_gib * 2**30),
                        args.serve_streams, args.serve_max_batch, args.serve_capacity_only)
        pt = sv.pop("place_timing")
        cells = sv.pop("cells")
        out["value"] = round(sv["jobs_per_s"], 2)
        out["ms_per_step"] = round(1000.0 * float(np.mean(sv["step_s"])), 3)
        sv["jobs_per_s"] = out["value"]
        sv["step_s"] = [round(x, 4) for x in sv["step_s"]]
        pairs = sum(j["grid_w"] * j["grid_h"] * len(S["st"]["libs"][j["lib"]]) for j in S["jobs"])
        out["config"] = {
            "workload": f"served Poisson stream: {args.serve_jobs} jobs, cells log-uniform "
                        f"300-24,000 ({cells} total), 4:3 / 16:9 at 16 px, 4 resident libraries "
                        f"of 1k/2.5k/5k/10k materials 64x64 (mixture colours), usage 3-5, weight "
                        f"0.25, colour 0.4, 1,000 SA proposals each; one step = the whole stream "
                        f"through one long-lived server with every job queued at once (one "
                        f"submission call)",
            "jobs": args.serve_jobs, "cells": cells, "knn_pairs": pairs,
            "libraries": [1000, 2500, 5000, 10000],
            "parallelism": f"1 GPU, continuous batching on {args.serve_streams} stream slots, "
                           f"<= {args.serve_max_batch} jobs per launch"}
        out["latency_ms"] = sv["rate"]["latency_ms"] if "rate" in sv else None
        out["tiles_placed_per_s"] = sv["tiles_placed_per_s"]
        out["sa_proposals_per_s"] = sv["sa_proposals_per_s"]
        # roofline of the dominant kernel: the placement chains of each
        # served batch (one wave per job), 16 B + 12 K B of algorithmic
        # traffic per tile placed (K = 100 for every library of the stream)
        if pt["launches"]:
            per_launch = pt["cells"] / pt["launches"] * (16 + 12 * 100)
            avg_ms = pt["ms"] / pt["launches"]
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            out["roofline"] = {"kernel": "place_raster_kernel (served batches)", "bound": "hbm",
                               "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                               "traffic": measured_traffic("place", "served"),
                               "algorithmic_bytes": int(round(per_launch)),
                               "avg_ms": round(avg_ms, 4), "launches": pt["launches"]}
        out["served"] = sv

    if "plain" in legs:
        pm, keep = run_plain(mh, synth, ctx, "medium", args.plain_steps, 3)
        p, target, mats, lib, d_target, d_grid, d_image = keep
        n = p["n"]
        pk = pm["kernels"]["place"]
        amount, _, _ = algorithmic_work("place", p, n)
        achieved = amount / (pk["avg_ms"] * 1e-3) / 1e9
        pm["roofline"] = {"kernel": "place_raster_kernel (one job)", "bound": "hbm",
                          "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBS, 6),
                          "traffic": measured_traffic("place", "medium"),
                          "algorithmic_bytes": amount, "avg_ms": pk["avg_ms"]}
        pm["similarity"] = run_similarity(ctx, lib, n, 5)
        if "cpu" in legs:
            # the same job on the host: library prepared once (Lab means,
            # tiles resized to the cell size), then the per-job work
            tw, th = p["width"] // p["grid_w"], p["height"] // p["grid_h"]
            prep = orc.prepare_library(mats, [(tw, th)], threads=cpu_threads)
            rate, jobs, el, ph = cpu_jobs(orc, [(target, prep, dict(p, seed=0))], args.cpu_seconds)
            pm["cpu_per_job"] = {"value": round(rate, 4), "unit": "jobs/s", "cores": 1,
                                 "kind": "port", "phase_s_per_job": {k: round(v / jobs, 5)
                                                                     for k, v in ph.items()},
                                 "sample": f"{jobs} medium jobs ({el:.1f} s) on one core of "
                                           f"{cpu_name}, library prepared once"}
            # whole job including material ingest, beside the GPU doing the same
            t0 = time.perf_counter()
            k = 0
            while time.perf_counter() - t0 < args.cpu_seconds / 2 or k < 2:
                orc.generate(target, mats, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                             p["color"], True, p["iters"], seed=k)
                k += 1
            cpu_whole = k / (time.perf_counter() - t0)
            prm = mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"],
                               1, p["iters"], 100.0, 0.99995, 0)
            t0 = time.perf_counter()
            for _ in range(5):
                lib2 = ctx.library_from_pixels(mats)
                ctx.generate_device(lib2, d_target, p["width"], p["height"], prm, d_grid, d_image)
                ctx.synchronize()
                lib2.close()
            gpu_whole = 5 / (time.perf_counter() - t0)
            pm["whole_job_with_ingest"] = {
                "cpu_jobs_per_s": round(cpu_whole, 4), "cpu_cores": 1,
                "gpu_jobs_per_s": round(gpu_whole, 3),
                "note": "material library (1,500 x 128x128 Lab means, resized tiles) rebuilt "
                        "for every job; the GPU figure includes the host-to-device copy of the "
                        "materials"}
        ctx.device_free(d_target)
        ctx.device_free(d_grid)
        ctx.device_free(d_image)
        lib.close()
        out["plain_medium"] = pm

    if "large" in legs:
        pl, keep = run_plain(mh, synth, ctx, "large", 3, 1)
        p, target, mats, lib, d_target, d_grid, d_image = keep
        pl["similarity"] = run_similarity(ctx, lib, p["n"], 5)
        out["large"] = pl
        ctx.device_free(d_target)
        ctx.device_free(d_grid)
        ctx.device_free(d_image)
        lib.close()

    if args.calibrate_copy > 0:
        src = ctx.device_alloc(args.calibrate_copy)
        dst = ctx.device_alloc(args.calibrate_copy)
        for _ in range(3):
            ctx.copy_on_device(dst, src, args.calibrate_copy)
        ctx.synchronize()
        ctx.device_free(src)
        ctx.device_free(dst)

    if "batched" in legs:
        out["batched"] = run_batched(mh, synth, ctx, max(3, args.steps), 1)
    if "ingest" in legs:
        out["ingest"] = run_ingest(mh, synth, ctx, 3, orc if "cpu" in legs else None)

    if "cpu" in legs and S is not None:
        # the headline workload on the host: the stream's jobs in arrival
        # order, on its four libraries prepared once (Lab means, 16x16 tiles)
        st = S["st"]
        preps = [orc.prepare_library(m, [(16, 16)], threads=cpu_threads) for m in st["libs"]]
        cjobs = []
        for j in S["jobs"][:400]:
            c = st["canvases"][j["canvas"]][j["oy"]:j["oy"] + j["height"],
                                            j["ox"]:j["ox"] + j["width"]]
            cjobs.append((np.ascontiguousarray(c), preps[j["lib"]], j))
        rate, jobs, el, ph = cpu_jobs(orc, cjobs, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": round(rate, 4), "unit": "jobs/s", "cores": 1, "kind": "port",
            "sample": f"the first {jobs} jobs of the served stream in arrival order ({el:.1f} s) "
                      f"on one core of {cpu_name}: the C restatement of the reference's per-job "
                      f"path (cell Lab, k-d tree K nearest + placement chain, SA, composite) on "
                      f"libraries prepared once",
            "phase_s_per_job": {k: round(v / jobs, 5) for k, v in ph.items()}}
        rate_all, jobs_all, el_all, _ = cpu_jobs(orc, cjobs, args.cpu_seconds, cpu_threads)
        out["cpu_baseline_all_cores"] = {
            "value": round(rate_all, 4), "unit": "jobs/s", "cores": cpu_threads, "kind": "port",
            "sample": f"{jobs_all} jobs of the same stream ({el_all:.1f} s), one job per thread "
                      f"on {cpu_threads} host threads"}
    elif "cpu_baseline" not in out:
        out["cpu_baseline"] = None

    if S is not None:
        stream_free(ctx, S)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
