"""Batched (mh_batch) and served (mh_server) forms.

Every job of a fused batch or of the request loop must give exactly what
mh_generate_device gives for the same job alone (grid, composited image,
placement counters, annealing result), and what the oracle's generate_mosaic
(src/main.rs:477-685) gives for it.  Jobs mix grid sizes, libraries (including one with fewer than K=100
materials and an empty one), SA on/off, colour strength 0 and weight 0, and
read their targets as crops of shared canvases (row pitch != width*3);
every third job places with the GUI generator's Euclidean score.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stream(synth, n_jobs=12, seed=3):
    return synth.make_stream(n_jobs=n_jobs, rate=1e9, seed=seed, libraries=(90, 300),
                             cells=(20, 1200), canvas=(720, 560), n_canvases=2, iters=300)


class Setup:
    """Device-resident libraries + canvases and the job list of a stream."""

    def __init__(self, ctx, mh, synth, n_jobs=12, seed=3):
        self.ctx, self.mh = ctx, mh
        st = _stream(synth, n_jobs, seed)
        self.st = st
        self.libs = [ctx.library_from_pixels(m) for m in st["libs"]]
        self.libs.append(ctx.library_from_pixels(np.zeros((0, 64, 64, 3), np.uint8)))
        self.W, self.H = st["canvas_size"]
        self.canvases = [c.copy() for c in st["canvases"]]
        self.d_can = [ctx.to_device(c) for c in self.canvases]
        self.specs = []
        for k, j in enumerate(st["jobs"]):
            j = dict(j)
            j["sa"] = 0 if k % 3 == 1 else 1
            if k % 4 == 2:
                j["color"] = 0.0
            if k % 5 == 3:
                j["weight"] = 0.0
            j["score_mode"] = 1 if k % 3 == 2 else 0  # GUI scoring on every third job
            self.specs.append(j)
        # one job on the empty library
        e = dict(self.specs[0])
        e["lib"] = len(self.libs) - 1
        self.specs.append(e)
        self.allocs = []

    def params(self, j):
        return self.mh.JobParams(j["grid_w"], j["grid_h"], j["max_usage"], j["weight"], j["color"],
                                 j["sa"], j["iters"], 100.0, 0.99995, j["seed"],
                                 j.get("score_mode", 0))

    def alloc(self, nbytes):
        p = self.ctx.device_alloc(nbytes)
        self.allocs.append(p)
        return p

    def job(self, k, own_outputs):
        j = self.specs[k]
        ptr = self.d_can[j["canvas"]] + (j["oy"] * self.W + j["ox"]) * 3
        d_g = d_i = None
        if own_outputs:
            d_g = self.alloc(j["grid_w"] * j["grid_h"] * 4)
            d_i = self.alloc(j["width"] * j["height"] * 3)
        return self.ctx.job(self.libs[j["lib"]], ptr, j["width"], j["height"], self.params(j),
                            d_g, d_i, pitch=self.W * 3)

    def alone(self, k):
        """mh_generate_device on a contiguous copy of the job's crop."""
        j = self.specs[k]
        c = self.canvases[j["canvas"]][j["oy"]:j["oy"] + j["height"], j["ox"]:j["ox"] + j["width"]]
        ctx = self.ctx
        d_t = ctx.to_device(np.ascontiguousarray(c))
        cells = j["grid_w"] * j["grid_h"]
        d_g = ctx.device_alloc(cells * 4)
        d_i = ctx.device_alloc(j["width"] * j["height"] * 3)
        ctx.generate_device(self.libs[j["lib"]], d_t, j["width"], j["height"], self.params(j), d_g,
                            d_i)
        ctx.synchronize()
        res = ctx.generate_device_result()
        grid = ctx.to_host(d_g, cells, np.int32)
        img = ctx.to_host(d_i, (j["height"], j["width"], 3), np.uint8)
        for p in (d_t, d_g, d_i):
            ctx.device_free(p)
        return grid, img, res

    def oracle(self, k, orc):
        """The oracle's whole job on the same crop and materials (cached)."""
        if not hasattr(self, "_want"):
            self._want = {}
        if k not in self._want:
            j = self.specs[k]
            c = self.canvases[j["canvas"]][j["oy"]:j["oy"] + j["height"],
                                           j["ox"]:j["ox"] + j["width"]]
            mats = (self.st["libs"][j["lib"]] if j["lib"] < len(self.st["libs"])
                    else np.zeros((0, 64, 64, 3), np.uint8))
            self._want[k] = orc.generate(np.ascontiguousarray(c), mats, j["grid_w"], j["grid_h"],
                                         j["max_usage"], j["weight"], j["color"], bool(j["sa"]),
                                         j["iters"], 100.0, 0.99995, j["seed"],
                                         score_mode=j.get("score_mode", 0))
        return self._want[k]

    def check_oracle(self, k, orc, grid, img, res):
        w = self.oracle(k, orc)
        np.testing.assert_array_equal(grid, w["grid"])
        np.testing.assert_array_equal(img, w["image"])
        assert res["place"] == w["place"]["stats"]
        if w["anneal"] is not None:
            a, b = res["anneal"], w["anneal"]["result"]
            for key in ("initial_cost", "final_cost", "best_cost"):
                assert np.float32(a[key]).view(np.uint32) == np.float32(b[key]).view(np.uint32), key
            for key in ("improved_count", "accepted_count", "iterations", "evaluated"):
                assert a[key] == b[key], key

    def fetch(self, k, d_g, d_i):
        j = self.specs[k]
        grid = self.ctx.to_host(d_g, j["grid_w"] * j["grid_h"], np.int32)
        img = self.ctx.to_host(d_i, (j["height"], j["width"], 3), np.uint8)
        return grid, img

    def close(self):
        for p in self.allocs + self.d_can:
            self.ctx.device_free(p)
        for lib in self.libs:
            lib.close()


def _same(want, grid, img, res, sa):
    wg, wi, wr = want
    np.testing.assert_array_equal(grid, wg)
    np.testing.assert_array_equal(img, wi)
    assert res["place"] == wr["place"]
    assert res["out_size"] == wr["out_size"]
    if sa:
        assert res["anneal"] == wr["anneal"]


def test_batch_matches_jobs_alone(ctx, mh, synth, orc):
    s = Setup(ctx, mh, synth)
    jobs = [s.job(k, own_outputs=(k % 2 == 1)) for k in range(len(s.specs))]
    b = ctx.batch(jobs)
    b.run()
    res = b.results()
    assert res[-1]["place"]["empty"] == s.specs[-1]["grid_w"] * s.specs[-1]["grid_h"]
    for k in range(len(s.specs)):
        print(f"batch job {k}: {s.specs[k]['grid_w']}x{s.specs[k]['grid_h']} lib {s.specs[k]['lib']}",
              flush=True)
        g, im = s.fetch(k, *b.outputs(k))
        sa = s.specs[k]["sa"] and s.specs[k]["weight"] > 0
        _same(s.alone(k), g, im, res[k], sa)
        s.check_oracle(k, orc, g, im, res[k])
    b.close()
    s.close()


def test_batch_graph_replays(ctx, mh, synth):
    s = Setup(ctx, mh, synth, n_jobs=6, seed=5)
    jobs = [s.job(k, own_outputs=True) for k in range(len(s.specs))]
    b = ctx.batch(jobs)
    b.set_graph(True)
    want = [s.alone(k) for k in range(len(s.specs))]
    for rep in range(3):
        b.run()
        res = b.results()
        for k in range(len(s.specs)):
            g, im = s.fetch(k, *b.outputs(k))
            _same(want[k], g, im, res[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        # eager work on the same stream between replays
        s.alone(0)
    # replays re-read the inputs: change a canvas in place and compare again
    s.canvases[0] = np.ascontiguousarray(255 - s.canvases[0])
    ctx.copy_to_device(s.d_can[0], s.canvases[0])
    b.run()
    res = b.results()
    for k in range(len(s.specs)):
        g, im = s.fetch(k, *b.outputs(k))
        _same(s.alone(k), g, im, res[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
    b.close()
    s.close()


def test_server_results_equal_alone_under_memory_pressure(ctx, mh, synth, orc):
    s = Setup(ctx, mh, synth, n_jobs=16, seed=7)
    jobs = [s.job(k, own_outputs=True) for k in range(len(s.specs))]
    need = max(ctx.job_footprint(j) for j in jobs)
    budget = int(need * 2.5)
    srv = ctx.server(memory_budget=budget, max_batch=4, streams=2)
    tickets = [srv.submit(j) for j in jobs]
    got = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(jobs) and st["submitted"] == len(jobs)
    assert st["memory_peak"] <= budget and st["memory_in_use"] == 0
    assert st["batches"] >= (len(jobs) + 3) // 4 and st["max_batch_jobs"] <= 4
    for k in range(len(s.specs)):
        print(f"served job {k}", flush=True)
        assert got[k]["latency_ms"] >= 0.0
        g, im = s.fetch(k, jobs[k].d_grid_out, jobs[k].d_image_out)
        _same(s.alone(k), g, im, got[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        s.check_oracle(k, orc, g, im, got[k])
    s.close()


def test_server_submit_many(ctx, mh, synth, orc):
    # every job queued by one call; a batch holding an invalid job queues none
    s = Setup(ctx, mh, synth, n_jobs=10, seed=13)
    jobs = [s.job(k, own_outputs=True) for k in range(len(s.specs))]
    srv = ctx.server(memory_budget=1 << 30, max_batch=4, streams=2)
    lab_only = ctx.library_from_lab(np.zeros((4, 3), np.float32))
    bad = ctx.job(lab_only, jobs[0].d_target, jobs[0].width, jobs[0].height, jobs[0].params,
                  pitch=jobs[0].target_pitch)
    with pytest.raises(mh.MhError) as e:
        srv.submit_many(jobs[:3] + [bad])
    assert e.value.code == mh.MH_ERR_INVALID_ARGUMENT and "job 3" in str(e.value)
    assert srv.stats()["submitted"] == 0
    tickets = srv.submit_many(jobs)
    assert tickets == list(range(len(jobs)))
    got = [srv.wait(t) for t in tickets]
    srv.close()
    lab_only.close()
    for k in range(len(s.specs)):
        g, im = s.fetch(k, jobs[k].d_grid_out, jobs[k].d_image_out)
        _same(s.alone(k), g, im, got[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        s.check_oracle(k, orc, g, im, got[k])
    s.close()


def test_server_groups_by_size(ctx, mh, synth, orc):
    # more queued than one batch holds: batches take the oldest queued job and
    # the queued jobs nearest to it in size, so jobs run out of queue order;
    # every result is still the job's own, and the oldest job goes first
    s = Setup(ctx, mh, synth, n_jobs=18, seed=21)
    order = sorted(range(len(s.specs)), key=lambda k: (k % 2, k))  # sizes interleave
    jobs = [s.job(k, own_outputs=True) for k in order]
    srv = ctx.server(memory_budget=1 << 30, max_batch=2, streams=1)
    tickets = srv.submit_many(jobs)
    got = [srv.wait(t) for t in tickets]
    launched = [srv.times(t)[1] for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(jobs) and st["max_batch_jobs"] <= 2
    assert launched[0] == min(launched)
    for i, k in enumerate(order):
        print(f"grouped job {k}", flush=True)
        g, im = s.fetch(k, jobs[i].d_grid_out, jobs[i].d_image_out)
        _same(s.alone(k), g, im, got[i], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        s.check_oracle(k, orc, g, im, got[i])
    s.close()


def test_server_rejects_bad_jobs(ctx, mh, synth):
    s = Setup(ctx, mh, synth, n_jobs=2, seed=9)
    j = s.job(0, own_outputs=False)
    need = ctx.job_footprint(j)
    srv = ctx.server(memory_budget=need // 2, max_batch=2, streams=1)
    with pytest.raises(mh.MhError) as e:
        srv.submit(j)
    assert e.value.code == mh.MH_ERR_OUT_OF_MEMORY
    srv.close()
    srv = ctx.server(memory_budget=need * 4, max_batch=2, streams=1)
    lab_only = ctx.library_from_lab(np.zeros((4, 3), np.float32))
    bad = ctx.job(lab_only, j.d_target, j.width, j.height, j.params, pitch=j.target_pitch)
    with pytest.raises(mh.MhError) as e:
        srv.submit(bad)
    assert e.value.code == mh.MH_ERR_INVALID_ARGUMENT
    t = srv.submit(j)
    assert srv.wait(t)["place"]["placed"] == j.params.grid_w * j.params.grid_h
    srv.close()
    lab_only.close()
    s.close()


def test_handles_outlive_their_context(mh, synth):
    # A context of its own, destroyed while a library and a batch made on it
    # are still open: their device memory goes with it, later calls on them
    # fail cleanly, and destroying them afterwards is safe.
    c = mh.Context(0)
    s = Setup(c, mh, synth, n_jobs=2, seed=7)
    b = c.batch([s.job(k, own_outputs=False) for k in range(len(s.specs))])
    b.run()
    b.results()
    for p in s.d_can:
        c.device_free(p)
    s.d_can = []
    c.close()
    with pytest.raises(mh.MhError):
        b.run()
    with pytest.raises(mh.MhError):
        s.libs[0].similarities()
    b.close()
    for lib in s.libs:
        lib.close()


def test_batch_narrow_grids(ctx, mh, synth, orc):
    # grids of one to three columns (the one-cell-at-a-time chain below three
    # columns, the look-ahead chain from three) in one batch, both scores
    s = Setup(ctx, mh, synth, n_jobs=6, seed=11)
    shapes = [(1, 30), (2, 25), (3, 20), (1, 1), (2, 1), (3, 2)]
    for k, (gw, gh) in enumerate(shapes):
        j = s.specs[k]
        j["grid_w"], j["grid_h"] = gw, gh
        j["width"], j["height"] = gw * 16, gh * 16
        j["ox"], j["oy"] = 0, 0
        j["score_mode"] = k % 2
    jobs = [s.job(k, own_outputs=True) for k in range(len(s.specs))]
    b = ctx.batch(jobs)
    b.run()
    res = b.results()
    for k in range(len(s.specs)):
        print(f"narrow batch job {k}: {s.specs[k]['grid_w']}x{s.specs[k]['grid_h']}", flush=True)
        g, im = s.fetch(k, *b.outputs(k))
        _same(s.alone(k), g, im, res[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        s.check_oracle(k, orc, g, im, res[k])
    b.close()
    s.close()


def test_every_tuned_shape_gives_the_same_outputs(ctx, mh, synth):
    """mh_batch_tune / mh_batch_set_tuning change launch shapes only: every
    candidate shape (top-K cells per wave, composite threads) and the tuner's
    pick, with graph capture after each change, reproduce the jobs alone."""
    s = Setup(ctx, mh, synth, n_jobs=8, seed=11)
    jobs = [s.job(k, own_outputs=True) for k in range(len(s.specs))]
    want = [s.alone(k) for k in range(len(s.specs))]
    b = ctx.batch(jobs)
    b.set_graph(True)
    base = b.tuning()
    assert all(1 <= r <= 8 for r in base["topk_cells_per_wave"]), base
    assert all(t == 64 for t in base["composite_threads"]), base

    def check(label):
        res = b.results()
        for k in range(len(s.specs)):
            g, im = s.fetch(k, *b.outputs(k))
            _same(want[k], g, im, res[k], s.specs[k]["sa"] and s.specs[k]["weight"] > 0)
        print(f"tuned shape {label}: ok", flush=True)

    b.run()
    check("heuristic")
    for run, thr in ((1, 64), (2, 128), (4, 256), (8, 64)):
        b.set_tuning((run, run), (thr, thr))
        got = b.tuning()
        assert got["topk_cells_per_wave"] == [run] * got["groups"], got
        assert got["composite_threads"] == [thr] * got["groups"], got
        b.run()
        check(f"run {run}, {thr} threads")
    t = b.tune(repeats=2)
    assert t == {**b.tuning(), "topk_ms": t["topk_ms"], "composite_ms": t["composite_ms"]}
    assert all(r in (1, 2, 4, 8) for r in t["topk_cells_per_wave"]), t
    assert all(x in (64, 128, 256) for x in t["composite_threads"]), t
    assert all(ms > 0 for ms in t["composite_ms"]), t
    check("after tune")  # the tuner's own run
    b.run()
    check("tuned, captured")
    with pytest.raises(mh.MhError):
        b.set_tuning((9, 0), (0, 0))
    with pytest.raises(mh.MhError):
        b.set_tuning((0, 0), (96, 0))
    with pytest.raises(mh.MhError):
        b.tune(repeats=0)
    b.close()
    s.close()
