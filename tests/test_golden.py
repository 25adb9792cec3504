"""Golden fixtures (tests/golden/golden.json, written by make_golden.py from
the oracle on seeded inputs and the reference tests' pixels): the oracle must
keep reproducing them (CPU), and the HIP path must match them bit for bit
(GPU) without consulting the oracle at all."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def f(a):
    return np.asarray(a, np.float32)


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32).tolist()


# ----------------------------------------------------------------- CPU: oracle
def test_oracle_reproduces_place_fixture(orc):
    g = G["place_6x5"]
    r = orc.place(f(g["lab"]), f(g["cell_lab"]), g["grid_w"], g["grid_h"], g["max_usage"], g["weight"])
    assert r["grid"].tolist() == g["grid"] and r["stage"].tolist() == g["stage"]
    assert r["usage"].tolist() == g["usage"] and r["stats"] == g["stats"]
    assert bits(r["score"]) == g["score_bits"]
    assert g["stats"]["resets"] > 0


def test_oracle_reproduces_swap_and_cost_fixture(orc):
    g = G["swap_4x4"]
    lab, grid = f(g["lab"]), np.array(g["grid"], np.int32)
    assert [bits(orc.swap_delta(lab, grid, 4, 4, *p)) for p in g["pairs"]] == g["delta_bits"]
    assert bits(orc.total_cost(lab, grid, 4, 4)) == g["total_cost_bits"]


def test_oracle_reproduces_anneal_fixture(orc):
    g = G["anneal_5x4"]
    a = orc.anneal(f(g["lab"]), np.array(g["grid"], np.int32), 5, 4, g["iterations"], g["t0"],
                   g["decay"], seed=g["seed"], trace=True)
    assert a["grid"].tolist() == g["final_grid"] and a["trace_accept"].tolist() == g["accept"]
    assert bits(a["trace_delta"]) == g["delta_bits"]


def test_oracle_reproduces_topk_and_colour_fixtures(orc):
    g = G["topk_200"]
    idx, d = orc.nearest_n(f(g["lab"]), f(g["query"]), 100)
    assert idx.tolist() == g["idx"] and bits(d) == g["sqdist_bits"]
    lp = G["lab_pixels"]
    assert [bits(orc.rgb8_to_lab(*p)) for p in lp["rgb"]] == lp["lab_bits"]
    ap = G["adjust_pixels"]
    for a, row in zip(ap["adj"], ap["out"]):
        assert [orc.adjust_pixel(f(a), p).tolist() for p in ap["rgb"]] == row
    oa = G["optimal_adjustment"]
    assert [bits(orc.optimal_adjustment(t, u, 1.0)) for t, u in oa["pairs"]] == oa["adj_bits"]


# ------------------------------------------------------------ GPU: HIP path
@pytest.mark.gpu
def test_hip_matches_fixtures(ctx):
    g = G["place_6x5"]
    lib = ctx.library_from_lab(f(g["lab"]))
    r = ctx.place(lib, f(g["cell_lab"]), g["grid_w"], g["grid_h"], g["max_usage"], g["weight"])
    assert r["grid"].tolist() == g["grid"] and r["stage"].tolist() == g["stage"]
    assert r["usage"].tolist() == g["usage"] and r["stats"] == g["stats"]
    assert bits(r["score"]) == g["score_bits"]
    lib.close()

    g = G["swap_4x4"]
    lib = ctx.library_from_lab(f(g["lab"]))
    grid = np.array(g["grid"], np.int32)
    assert bits(ctx.swap_deltas(lib, grid, 4, 4, g["pairs"])) == g["delta_bits"]
    assert bits(ctx.total_cost(lib, grid, 4, 4)) == g["total_cost_bits"]
    lib.close()

    g = G["anneal_5x4"]
    lib = ctx.library_from_lab(f(g["lab"]))
    a = ctx.anneal(lib, np.array(g["grid"], np.int32), 5, 4, g["iterations"], g["t0"], g["decay"],
                   seed=g["seed"], trace=True)
    assert a["grid"].tolist() == g["final_grid"] and a["trace_accept"].tolist() == g["accept"]
    assert bits(a["trace_delta"]) == g["delta_bits"]
    assert bits(a["result"]["final_cost"]) == g["final_cost_bits"]
    lib.close()

    g = G["topk_200"]
    lib = ctx.library_from_lab(f(g["lab"]))
    idx, d = ctx.nearest_n(lib, f(g["query"]), 100)
    assert idx.tolist() == g["idx"] and bits(d) == g["sqdist_bits"]
    lib.close()

    lp = G["lab_pixels"]
    img = np.array(lp["rgb"], np.uint8)[None]
    assert bits(ctx.cell_lab_means(img, len(lp["rgb"]), 1).ravel()) == [x for row in lp["lab_bits"] for x in row]
    ap = G["adjust_pixels"]
    for a_, row in zip(ap["adj"], ap["out"]):
        assert ctx.adjust_pixels(f(a_), np.array(ap["rgb"], np.uint8)).tolist() == row
    oa = G["optimal_adjustment"]
    t = np.array([p[0] for p in oa["pairs"]], np.uint8)
    u = np.array([p[1] for p in oa["pairs"]], np.uint8)
    assert [bits(r_) for r_ in ctx.optimal_adjustments(t, u, 1.0)] == oa["adj_bits"]
