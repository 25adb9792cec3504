"""The three large workloads against the oracle, at their full sizes:

  * the large single job: 4000x2400 target, 10,000 mixture-colour materials of
    64x64 (dense clusters and near-ties), grid 200x120, usage 5, 20,000 SA
    proposals, and the library's 10,000 x 10,000 similarity matrix (400 MB)
    in full and packed form;
  * the batched form: the 64-job batch (own libraries of 256-4,096 materials,
    grids 20-160 x 15-120 at 16 px), tuned and replayed from a hipGraph;
  * the served form: a sample of the 5,000-job Poisson stream covering all
    four resident libraries (1k / 2.5k / 5k / 10k materials) and its largest
    jobs, served in a mix with other stream jobs.

Each job's grid, placement counters and annealing costs are compared with
the oracle's generate_mosaic (src/main.rs:477-685) bit for bit, and its
composited image pixel for pixel (the contract allows +-1 LSB; the path is
exact).  The oracle runs on a library prepared once (material Lab means and
resized tiles), on the host's threads.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _check_job(got_grid, got_img, got_res, want, label):
    np.testing.assert_array_equal(got_grid, want["grid"], err_msg=label)
    assert got_res["place"] == want["place"]["stats"], label
    if want["anneal"] is not None:
        a, b = got_res["anneal"], want["anneal"]["result"]
        for key in ("initial_cost", "final_cost", "best_cost"):
            assert _bits(np.float32(a[key])) == _bits(np.float32(b[key])), (label, key)
        for key in ("improved_count", "accepted_count", "iterations", "evaluated"):
            assert a[key] == b[key], (label, key)
    if got_img is not None:
        d = np.abs(got_img.astype(np.int16) - want["image"].astype(np.int16))
        assert d.max() <= 1, label
        np.testing.assert_array_equal(got_img, want["image"], err_msg=label)


def _pool(fn, items):
    import concurrent.futures
    with concurrent.futures.ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(fn, items))


# ------------------------------------------------------------ large job
def test_large_job_matches_oracle(ctx, synth, orc):
    p, target, mats = synth.make("large")
    gw, gh = p["grid_w"], p["grid_h"]
    tw, th = p["width"] // gw, p["height"] // gh
    print(f"large: {p['n']} materials, grid {gw}x{gh}", flush=True)
    prep = orc.prepare_library(mats, [(tw, th)], threads=THREADS)
    lib = ctx.library_from_pixels(mats)
    np.testing.assert_array_equal(_bits(lib.lab()), _bits(prep["lab"]))
    got = ctx.generate(lib, target, gw, gh, p["max_usage"], p["weight"], p["color"], True,
                       p["iters"], seed=0)
    print(f"large: GPU job done ({got['place']})", flush=True)
    want = orc.generate_prepared(target, prep, gw, gh, p["max_usage"], p["weight"], p["color"],
                                 True, p["iters"], seed=0, threads=THREADS)
    print("large: oracle job done", flush=True)
    _check_job(got["grid"], got["image"], got, want, "large")
    # the near-ties of the mixture library send cells down every path
    st = want["place"]["stats"]
    assert st["placed"] == gw * gh and st["main_path"] > 0
    lib.close()


def test_large_similarity_matrix_and_terms(ctx, synth, orc):
    p, target, mats = synth.make("large")
    n = p["n"]
    lab = orc.stack_lab_means_par(mats, THREADS)
    lib = ctx.library_from_lab(lab)
    lib.build_similarities()
    full = lib.similarities()  # 10,000 x 10,000 f32: 400 MB
    assert full.shape == (n, n)
    rng = np.random.default_rng(5)
    rows = np.unique(np.concatenate([[0, 1, n // 2, n - 2, n - 1], rng.integers(0, n, 59)]))
    rows = rows.astype(np.int32)
    want = orc.sim_rows(lab, rows)
    np.testing.assert_array_equal(_bits(full[rows]), _bits(want))
    np.testing.assert_array_equal(_bits(full[:, rows].T), _bits(want))  # bitwise symmetric
    assert np.all(full[rows, rows] == 0.0)
    print("large: full similarity rows match", flush=True)
    del full
    packed = lib.similarities(packed=True)  # n(n-1)/2 entries, src/similarity.rs:100
    assert packed.size == n * (n - 1) // 2
    for r, i in enumerate(rows.tolist()):
        j = np.arange(i + 1, n)
        pos = i * n - i * (i + 1) // 2 + j - i - 1
        np.testing.assert_array_equal(_bits(packed[pos]), _bits(want[r, i + 1:]))
    del packed
    print("large: packed similarity rows match", flush=True)
    # the adjacency terms 1/(sim+1) from the resident term matrix: total cost
    # and swap deltas on a placed 200 x 120 grid (src/adjacency.rs:91-214)
    gw, gh = p["grid_w"], p["grid_h"]
    cl = orc.cell_lab_means_par(target, gw, gh, THREADS)
    grid = orc.place(lab, cl, gw, gh, p["max_usage"], p["weight"])["grid"]
    assert _bits(np.float32(ctx.total_cost(lib, grid, gw, gh))) == \
        _bits(orc.total_cost(lab, grid, gw, gh))
    pairs = np.stack([rng.integers(0, gw, 3000), rng.integers(0, gh, 3000),
                      rng.integers(0, gw, 3000), rng.integers(0, gh, 3000)], 1).astype(np.int32)
    pairs[:500, 2] = np.minimum(pairs[:500, 0] + 1, gw - 1)  # adjacent pairs
    pairs[:500, 3] = pairs[:500, 1]
    got = ctx.swap_deltas(lib, grid, gw, gh, pairs)
    want_d = np.array([orc.swap_delta(lab, grid, gw, gh, *map(int, q)) for q in pairs], np.float32)
    np.testing.assert_array_equal(_bits(got), _bits(want_d))
    lib.close()


# ------------------------------------------------------------ batched form
def test_batched_64_jobs_match_oracle(ctx, mh, synth, orc):
    jobs = synth.make_batch(64)
    print(f"batch: {len(jobs)} jobs, {sum(p['grid_w'] * p['grid_h'] for p, _, _ in jobs)} cells",
          flush=True)
    libs, bufs, mj = [], [], []
    for k, (p, target, mats) in enumerate(jobs):
        lib = ctx.library_from_pixels(mats)
        libs.append(lib)
        d_t = ctx.to_device(target)
        d_g = ctx.device_alloc(p["grid_w"] * p["grid_h"] * 4)
        d_i = ctx.device_alloc(p["width"] * p["height"] * 3)
        bufs += [d_t, d_g, d_i]
        prm = mh.JobParams(p["grid_w"], p["grid_h"], p["max_usage"], p["weight"], p["color"], 1,
                           p["iters"], 100.0, 0.99995, k)
        mj.append(ctx.job(lib, d_t, p["width"], p["height"], prm, d_g, d_i))
    b = ctx.batch(mj)
    b.set_graph(True)
    b.tune(repeats=2)
    b.run()  # captures the tuned launch sequence
    b.run()  # replays it
    res = b.results()
    print("batch: GPU done", flush=True)

    def oracle_job(k):
        p, target, mats = jobs[k]
        prep = orc.prepare_library(mats, [(16, 16)], threads=1)
        return orc.generate_prepared(target, prep, p["grid_w"], p["grid_h"], p["max_usage"],
                                     p["weight"], p["color"], True, p["iters"], seed=k)

    want = _pool(oracle_job, list(range(len(jobs))))
    print("batch: oracle done", flush=True)
    for k, (p, _, _) in enumerate(jobs):
        d_g, d_i = b.outputs(k)
        grid = ctx.to_host(d_g, p["grid_w"] * p["grid_h"], np.int32)
        img = ctx.to_host(d_i, (p["height"], p["width"], 3), np.uint8)
        _check_job(grid, img, res[k], want[k], f"batch job {k}")
    b.close()
    for ptr in bufs:
        ctx.device_free(ptr)
    for lib in libs:
        lib.close()


# ------------------------------------------------------------- served form
def test_served_stream_sample_matches_oracle(ctx, mh, synth, orc):
    st = synth.make_stream(5000, 2000.0)
    jobs = st["jobs"]
    W = st["canvas_size"][0]
    cells = np.array([j["grid_w"] * j["grid_h"] for j in jobs])
    sample = set(np.argsort(cells)[-4:].tolist())  # the largest jobs of the stream
    for li in range(len(st["libs"])):  # per library: its 3 largest jobs and 2 others
        mine = [k for k, j in enumerate(jobs) if j["lib"] == li]
        mine.sort(key=lambda k: cells[k])
        sample.update(mine[-3:])
        sample.update(mine[len(mine) // 3:len(mine) // 3 + 2])
    sample = sorted(sample)
    print(f"served: {len(sample)} sampled jobs, cells {sorted(cells[sample].tolist())}", flush=True)
    libs = [ctx.library_from_pixels(m) for m in st["libs"]]
    d_can = [ctx.to_device(c) for c in st["canvases"]]
    outs = {}
    for k in sample:
        j = jobs[k]
        outs[k] = (ctx.device_alloc(j["grid_w"] * j["grid_h"] * 4),
                   ctx.device_alloc(j["width"] * j["height"] * 3))

    def mjob(k):
        j = jobs[k]
        prm = mh.JobParams(j["grid_w"], j["grid_h"], j["max_usage"], j["weight"], j["color"], 1,
                           j["iters"], 100.0, 0.99995, j["seed"])
        d_g, d_i = outs.get(k, (None, None))
        return ctx.job(libs[j["lib"]], d_can[j["canvas"]] + (j["oy"] * W + j["ox"]) * 3,
                       j["width"], j["height"], prm, d_g, d_i, pitch=W * 3)

    # the sampled jobs in a mix with the stream's first 400 jobs, all queued
    order = sorted(set(range(400)) | set(sample))
    srv = ctx.server(memory_budget=16 << 30, max_batch=64, streams=4)
    tickets = {k: srv.submit(mjob(k)) for k in order}
    got = {k: srv.wait(tickets[k]) for k in sample}
    srv.drain()
    stats = srv.stats()
    srv.close()
    assert stats["completed"] == len(order) and stats["batches"] > 1
    print(f"served: GPU done ({stats['batches']} batches)", flush=True)

    preps = _pool(lambda m: orc.prepare_library(m, [(16, 16)], threads=1), st["libs"])

    def oracle_job(k):
        j = jobs[k]
        c = st["canvases"][j["canvas"]][j["oy"]:j["oy"] + j["height"],
                                        j["ox"]:j["ox"] + j["width"]]
        return orc.generate_prepared(np.ascontiguousarray(c), preps[j["lib"]], j["grid_w"],
                                     j["grid_h"], j["max_usage"], j["weight"], j["color"], True,
                                     j["iters"], 100.0, 0.99995, j["seed"])

    want = dict(zip(sample, _pool(oracle_job, sample)))
    print("served: oracle done", flush=True)
    for k in sample:
        j = jobs[k]
        d_g, d_i = outs[k]
        grid = ctx.to_host(d_g, j["grid_w"] * j["grid_h"], np.int32)
        img = ctx.to_host(d_i, (j["height"], j["width"], 3), np.uint8)
        _check_job(grid, img, got[k], want[k], f"served job {k} (lib {j['lib']})")
        ctx.device_free(d_g)
        ctx.device_free(d_i)
    for ptr in d_can:
        ctx.device_free(ptr)
    for lib in libs:
        lib.close()
