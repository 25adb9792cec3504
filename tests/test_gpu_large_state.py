"""Jobs whose working state does not fit LDS, against the oracle: a 300 x 200
grid (60,000 cells: annealed with the grid in HBM and hashed LDS stamps) and
a 50,000-material library (placement with its usage counts and grid in HBM).
The reference takes any grid and any library size (src/lib.rs:57-89,
src/optimizer.rs:41-129); results must be the serial chain's bit for bit,
alone and inside a batch with ordinary jobs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _check(got, want, label):
    np.testing.assert_array_equal(got["grid"], want["grid"], err_msg=label)
    assert got["place"] == want["place"]["stats"], label
    if want["anneal"] is not None:
        a, b = got["anneal"], want["anneal"]["result"]
        for key in ("initial_cost", "final_cost", "best_cost"):
            assert _bits(np.float32(a[key])) == _bits(np.float32(b[key])), (label, key)
        for key in ("improved_count", "accepted_count", "iterations", "evaluated"):
            assert a[key] == b[key], (label, key)
    if got.get("image") is not None:
        np.testing.assert_array_equal(got["image"], want["image"], err_msg=label)


def _big_grid_job(synth):
    rng = np.random.default_rng(41)
    gw, gh, cell = 300, 200, 8
    target = synth.blob_target(gw * cell, gh * cell, 16, rng, sigma=200.0)
    mats = synth.flat_materials(700, 8, rng)
    return dict(grid_w=gw, grid_h=gh, max_usage=100, weight=0.25, color=0.4, iters=6000), target, mats


def _big_library_job(synth):
    rng = np.random.default_rng(43)
    mats = synth.flat_materials(50000, 4, rng, colours=synth.mixture_colours(50000, rng))
    target = synth.blob_target(640, 480, 8, rng)
    return dict(grid_w=40, grid_h=30, max_usage=2, weight=0.25, color=0.4, iters=800), target, mats


def test_anneal_grid_beyond_lds(ctx, synth, orc):
    p, target, mats = _big_grid_job(synth)
    print(f"grid {p['grid_w']}x{p['grid_h']}, {mats.shape[0]} materials", flush=True)
    lib = ctx.library_from_pixels(mats)
    got = ctx.generate(lib, target, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                       p["color"], True, p["iters"], seed=5)
    prep = orc.prepare_library(mats, [(8, 8)], threads=THREADS)
    want = orc.generate_prepared(target, prep, p["grid_w"], p["grid_h"], p["max_usage"],
                                 p["weight"], p["color"], True, p["iters"], seed=5,
                                 threads=THREADS)
    _check(got, want, "300x200")
    assert got["anneal"]["accepted_count"] > 0
    # the same annealing through mh_anneal on the placed grid, with traces
    grid0 = want["place"]["grid"]
    g = ctx.anneal(lib, grid0, p["grid_w"], p["grid_h"], 3000, seed=9, trace=True)
    w = orc.anneal(want["lab"], grid0, p["grid_w"], p["grid_h"], 3000, seed=9, trace=True)
    np.testing.assert_array_equal(g["grid"], w["grid"])
    np.testing.assert_array_equal(_bits(g["trace_delta"]), _bits(w["trace_delta"]))
    np.testing.assert_array_equal(g["trace_accept"], w["trace_accept"])
    lib.close()


def test_place_library_beyond_lds(ctx, synth, orc):
    p, target, mats = _big_library_job(synth)
    print(f"library of {mats.shape[0]} materials", flush=True)
    lib = ctx.library_from_pixels(mats)
    got = ctx.generate(lib, target, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                       p["color"], True, p["iters"], seed=3)
    prep = orc.prepare_library(mats, [(16, 16)], threads=THREADS)
    want = orc.generate_prepared(target, prep, p["grid_w"], p["grid_h"], p["max_usage"],
                                 p["weight"], p["color"], True, p["iters"], seed=3)
    _check(got, want, "50k library")
    # mh_place directly, with its per-cell outputs and usage counts
    cl = orc.cell_lab_means(target, p["grid_w"], p["grid_h"])
    gp = ctx.place(lib, cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"])
    wp = orc.place(prep["lab"], cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"])
    np.testing.assert_array_equal(gp["grid"], wp["grid"])
    np.testing.assert_array_equal(gp["usage"], wp["usage"])
    np.testing.assert_array_equal(gp["stage"], wp["stage"])
    np.testing.assert_array_equal(_bits(gp["dist"]), _bits(wp["dist"]))
    assert gp["stats"] == wp["stats"]
    lib.close()


def _mid_library_job(synth):
    # 32,000 materials on a 40-wide grid: the usage counts fit the raster
    # chain's 160 KB of LDS (4 B per material plus the lane scratch and the
    # column row), the 52 KB of staged candidate rows beside them do not
    # (they fit up to about 27,500 materials), so the chain reads its
    # candidate rows from global memory
    rng = np.random.default_rng(47)
    mats = synth.flat_materials(32000, 4, rng, colours=synth.mixture_colours(32000, rng))
    target = synth.blob_target(640, 480, 8, rng)
    return dict(grid_w=40, grid_h=30, max_usage=2, weight=0.25, color=0.4, iters=800), target, mats


def test_place_library_unstaged(ctx, synth, orc):
    p, target, mats = _mid_library_job(synth)
    print(f"library of {mats.shape[0]} materials", flush=True)
    lib = ctx.library_from_pixels(mats)
    got = ctx.generate(lib, target, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                       p["color"], True, p["iters"], seed=4)
    prep = orc.prepare_library(mats, [(16, 16)], threads=THREADS)
    want = orc.generate_prepared(target, prep, p["grid_w"], p["grid_h"], p["max_usage"],
                                 p["weight"], p["color"], True, p["iters"], seed=4)
    _check(got, want, "32k library")
    cl = orc.cell_lab_means(target, p["grid_w"], p["grid_h"])
    gp = ctx.place(lib, cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"])
    wp = orc.place(prep["lab"], cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"])
    np.testing.assert_array_equal(gp["grid"], wp["grid"])
    np.testing.assert_array_equal(gp["usage"], wp["usage"])
    np.testing.assert_array_equal(_bits(gp["dist"]), _bits(wp["dist"]))
    assert gp["stats"] == wp["stats"]
    lib.close()


def _run_batch(ctx, mh, orc, jobs):
    libs, bufs, mj, want = [], [], [], []
    for k, (p, target, mats) in enumerate(jobs):
        gw, gh = p["grid_w"], p["grid_h"]
        h, w = target.shape[:2]
        tw, th = w // gw, h // gh
        lib = ctx.library_from_pixels(mats)
        libs.append(lib)
        d_t = ctx.to_device(target)
        d_g = ctx.device_alloc(gw * gh * 4)
        d_i = ctx.device_alloc(gw * tw * gh * th * 3)
        bufs += [d_t, d_g, d_i]
        prm = mh.JobParams(gw, gh, p["max_usage"], p["weight"], p["color"], 1, p["iters"], 100.0,
                           0.99995, 20 + k)
        mj.append(ctx.job(lib, d_t, w, h, prm, d_g, d_i))
        prep = orc.prepare_library(mats, [(tw, th)], threads=THREADS)
        want.append(orc.generate_prepared(target, prep, gw, gh, p["max_usage"], p["weight"],
                                          p["color"], True, p["iters"], seed=20 + k))
    b = ctx.batch(mj)
    b.run()
    res = b.results()
    for k, (p, target, _) in enumerate(jobs):
        gw, gh = p["grid_w"], p["grid_h"]
        h, w = target.shape[:2]
        d_g, d_i = b.outputs(k)
        got = dict(res[k])
        got["grid"] = ctx.to_host(d_g, gw * gh, np.int32)
        got["image"] = ctx.to_host(d_i, (gh * (h // gh), gw * (w // gw), 3), np.uint8)
        _check(got, want[k], f"batch job {k}")
    b.close()
    for ptr in bufs:
        ctx.device_free(ptr)
    for lib in libs:
        lib.close()


def test_batch_mixes_hbm_and_lds_jobs(ctx, mh, synth, orc):
    _run_batch(ctx, mh, orc, [_big_grid_job(synth), _big_library_job(synth), synth.make("small")])


def test_batch_unstaged_chains(ctx, mh, synth, orc):
    # one job too large for staged rows: every LDS chain of the launch runs
    # unstaged, the small jobs' too
    _run_batch(ctx, mh, orc, [synth.make("small"), _mid_library_job(synth),
                              synth.make("small", seed=1)])
