"""Pins the C oracle (the checker) to the reference and to mathematics.

Three kinds of checks, chosen so that a plausible slip in the restatement
(a dropped term, a swapped operand, a wrong index or tie rule) fails:
  * known answers from the reference's own unit tests (file:line cited);
  * closed forms and independent implementations (numpy float64/float32,
    colorsys) of the same published formulas;
  * invariants (swap delta == cost difference, SA bookkeeping, grid
    permutation, no cooling on skipped proposals).
Runs on the CPU only.
"""
import colorsys
import math

import numpy as np
import pytest

f32 = np.float32


def _rng(s):
    return np.random.default_rng(s)


# ------------------------------------------------------------ building blocks
def test_srgb_lut_matches_the_transfer_function(orc):
    lut = orc.srgb_lut()
    for v in range(256):
        c = f32(v) / f32(255.0)
        want = c / f32(12.92) if c <= f32(0.04045) else ((float(c) + 0.055) / 1.055) ** 2.4
        assert abs(float(lut[v]) - float(want)) <= 1e-7 * max(1.0, float(want))
    assert np.all(np.diff(lut) > 0) and lut[0] == 0.0 and lut[255] == 1.0


def test_cbrt_is_correctly_rounded(orc):
    xs = np.concatenate([_rng(0).uniform(0.0088, 1.2, 20000), [1.0, 0.125, 0.008856452, 1.0889]])
    xs = xs.astype(np.float32)
    got = np.array([orc.cbrtf(x) for x in xs], np.float32)
    want = np.cbrt(xs.astype(np.float64)).astype(np.float32)
    ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32))
    assert ulp.max() <= 1 and np.mean(ulp == 0) > 0.999


def test_exp_is_correctly_rounded_and_handles_specials(orc):
    xs = np.concatenate([_rng(1).uniform(-110, 0, 20000), _rng(2).uniform(-1e-6, 0, 1000)])
    xs = xs.astype(np.float32)
    got = np.array([orc.expf(x) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32))
    assert ulp.max() <= 1
    assert orc.expf(0.0) == 1.0 and orc.expf(float("-inf")) == 0.0
    assert math.isnan(orc.expf(float("nan"))) and orc.expf(100.0) == float("inf")


# ----------------------------------------------------------------- a1 Lab
def test_lab_known_answers(orc):
    # src/lib.rs:133-144 (red) and :147-158 (grey 128)
    red = orc.rgb8_to_lab(255, 0, 0)
    assert abs(red[0] - 53.24) < 1.0 and abs(red[1] - 80.09) < 1.0 and abs(red[2] - 67.20) < 1.0
    grey = orc.rgb8_to_lab(128, 128, 128)
    assert abs(grey[0] - 53.59) < 1.0 and abs(grey[1]) < 1.0 and abs(grey[2]) < 1.0
    # the Lab points of the reference's k-d tree test (src/lib.rs:213-220),
    # quoted there to two decimals
    assert np.allclose(orc.rgb8_to_lab(0, 255, 0), [87.74, -86.18, 83.18], atol=0.02)
    assert np.allclose(orc.rgb8_to_lab(0, 0, 255), [32.30, 79.20, -107.86], atol=0.02)
    assert np.array_equal(orc.rgb8_to_lab(0, 0, 0), [0, 0, 0])
    assert orc.rgb8_to_lab(255, 255, 255)[0] == 100.0  # FromColor clamps L to [0, 100]


def _lab_f64(rgb):
    """Independent float64 CIE 1976 L*a*b* (sRGB, D65) for comparison."""
    c = np.asarray(rgb, np.float64) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    m = np.array([[0.4124564, 0.3575761, 0.1804375], [0.2126729, 0.7151522, 0.0721750],
                  [0.0193339, 0.1191920, 0.9503041]])
    xyz = m @ lin / np.array([0.95047, 1.0, 1.08883])
    eps, kap = 216.0 / 24389.0, 24389.0 / 27.0
    f = np.where(xyz > eps, np.cbrt(xyz), (kap * xyz + 16.0) / 116.0)
    return np.array([116 * f[1] - 16, 500 * (f[0] - f[1]), 200 * (f[1] - f[2])])


def test_lab_closed_form_on_random_colours(orc):
    for rgb in _rng(3).integers(0, 256, (3000, 3)):
        got = orc.rgb8_to_lab(*rgb)
        want = np.clip(_lab_f64(rgb), [0, -128, -128], [100, 127, 127])
        assert np.max(np.abs(got - want)) < 2e-3, (rgb, got, want)


def test_average_lab_is_a_serial_f32_fold(orc):
    img = _rng(4).integers(0, 256, (7, 9, 3)).astype(np.uint8)
    got = orc.average_lab(img)
    s = np.zeros(3, np.float32)
    for y in range(7):
        for x in range(9):
            s = (s + orc.rgb8_to_lab(*img[y, x])).astype(np.float32)  # row-major order
    assert np.array_equal(got, (s / f32(63)).astype(np.float32))


def test_cell_lab_means_crop_geometry(orc):
    # src/main.rs:495-496, :536-541: cells of W/gw x H/gh at (x*tw, y*th);
    # the right/bottom remainder is ignored
    img = _rng(5).integers(0, 256, (23, 31, 3)).astype(np.uint8)
    got = orc.cell_lab_means(img, 4, 3)
    tw, th = 31 // 4, 23 // 3
    for c in range(12):
        x, y = c % 4, c // 4
        assert np.array_equal(got[c], orc.average_lab(img[y * th:(y + 1) * th, x * tw:(x + 1) * tw]))


# ----------------------------------------------------------- a2 similarity
def test_similarity_known_answers(orc):
    # src/similarity.rs:186-218, :320-362
    lab = np.array([[50, 0, 0], [60, 10, 10], [40, -10, -10]], np.float32)
    assert orc.lab_distance(lab[0], lab[1]) == f32(np.sqrt(f32(300.0)))
    assert orc.lab_distance(lab[0], lab[0]) == 0.0
    full = orc.sim_full(lab)
    assert np.array_equal(full, full.T) and np.all(np.diag(full) == 0)
    assert np.all(full[~np.eye(3, dtype=bool)] > 0)


def test_packed_layout_matches_the_reference_index_formula(orc):
    lab = _rng(6).uniform(-50, 100, (17, 3)).astype(np.float32)
    packed, full = orc.sim_packed(lab), orc.sim_full(lab)
    n = 17
    for i in range(n):
        for j in range(i + 1, n):
            assert packed[i * n - i * (i + 1) // 2 + j - i - 1] == full[i, j]  # similarity.rs:100


def test_distance_is_bitwise_symmetric(orc):
    a = _rng(7).uniform(-100, 100, (5000, 3)).astype(np.float32)
    b = _rng(8).uniform(-100, 100, (5000, 3)).astype(np.float32)
    for i in range(5000):
        assert orc.lab_distance(a[i], b[i]) == orc.lab_distance(b[i], a[i])


# -------------------------------------------------------------- a3 top-K
def test_nearest_n_is_brute_force_sorted_by_distance_then_index(orc):
    rng = _rng(9)
    lab = np.round(rng.uniform(-20, 20, (500, 3))).astype(np.float32)  # many exact ties
    q = np.round(rng.uniform(-20, 20, (40, 3))).astype(np.float32)
    idx, d = orc.nearest_n(lab, q, 100)
    for r in range(40):
        diff = q[r] - lab
        sq = ((diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1]) + diff[:, 2] * diff[:, 2]).astype(np.float32)
        order = np.lexsort((np.arange(500), sq))[:100]
        assert np.array_equal(idx[r], order) and np.array_equal(d[r], sq[order])


def test_nearest_one_known_answer(orc):
    # src/lib.rs:213-227
    lab = np.array([[50, 0, 0], [53.24, 80.09, 67.20], [87.74, -86.18, 83.18],
                    [32.30, 79.20, -107.86]], np.float32)
    idx, _ = orc.nearest_n(lab, np.array([53.0, 80.0, 67.0], np.float32), 1)
    assert idx[0, 0] == 1
    kidx, _ = orc.nearest_n_kd(lab, np.array([53.0, 80.0, 67.0], np.float32), 1)
    assert kidx[0, 0] == 1


@pytest.mark.parametrize("n", [0, 1, 7, 100, 101, 256, 257, 1500, 10000])
def test_kd_tree_nearest_equals_brute_force(orc, n):
    # The placement queries a k-d tree (the reference's kiddo structure,
    # src/main.rs:145-151); its K nearest must be the brute-force lexsort
    # answer exactly, including exact ties (integer coordinates, duplicated
    # points, queries on split planes) and the order among them.
    rng = _rng(n + 17)
    lab = np.round(rng.uniform(-30, 30, (n, 3))).astype(np.float32)
    if n > 10:
        lab[n // 2:n // 2 + 5] = lab[3]  # duplicates
    q = np.concatenate([np.round(rng.uniform(-35, 35, (150, 3))),
                        rng.uniform(-35, 35, (150, 3)),
                        lab[:min(n, 20)]]).astype(np.float32)
    for k in (1, 7, 100):
        bi, bd = orc.nearest_n(lab, q, k) if n else (np.zeros((q.shape[0], 0), np.int32),) * 2
        ki, kd = orc.nearest_n_kd(lab, q, k)
        np.testing.assert_array_equal(ki, bi)
        if n:
            np.testing.assert_array_equal(kd.view(np.uint32), bd.view(np.uint32))


def test_kd_tree_clustered_library(orc, synth):
    # dense mixture clusters with near-ties, as the large library has
    rng = _rng(31)
    cols = synth.mixture_colours(4000, rng)
    lab = np.stack([orc.rgb8_to_lab(*np.clip(np.rint(c), 0, 255).astype(int)) for c in cols])
    q = lab[rng.integers(0, 4000, 300)] + rng.normal(0, 0.5, (300, 3)).astype(np.float32)
    bi, bd = orc.nearest_n(lab, q, 100)
    ki, kd = orc.nearest_n_kd(lab, q, 100)
    np.testing.assert_array_equal(ki, bi)
    np.testing.assert_array_equal(kd.view(np.uint32), bd.view(np.uint32))


def test_sim_rows_are_rows_of_the_full_matrix(orc):
    lab = _rng(33).uniform(-50, 100, (57, 3)).astype(np.float32)
    rows = np.array([0, 5, 56, 5], np.int32)
    np.testing.assert_array_equal(orc.sim_rows(lab, rows), orc.sim_full(lab)[rows])


def test_prepared_job_equals_the_whole_job(orc, synth):
    # a library prepared once (Lab means, tiles resized per cell size) gives
    # the whole job's result: the per-job work the CPU baseline times
    p, target, mats = synth.make("small")
    prep = orc.prepare_library(mats, [(16, 16)], threads=4)
    np.testing.assert_array_equal(prep["lab"], orc.stack_lab_means(mats))
    np.testing.assert_array_equal(orc.cell_lab_means_par(target, 40, 30, threads=3),
                                  orc.cell_lab_means(target, 40, 30))
    times = {}
    a = orc.generate_prepared(target, prep, 40, 30, 3, 0.25, 0.4, True, 500, seed=2, times=times)
    b = orc.generate(target, mats, 40, 30, 3, 0.25, 0.4, True, 500, seed=2)
    np.testing.assert_array_equal(a["grid"], b["grid"])
    np.testing.assert_array_equal(a["image"], b["image"])
    assert a["place"]["stats"] == b["place"]["stats"]
    assert a["anneal"]["result"] == b["anneal"]["result"]
    assert set(times) == {"cell_lab", "place", "anneal", "composite"}


# ------------------------------------------------------------ a4-a9 place
def _lab_line(n):
    # materials on a line in L so distances are easy to reason about
    return np.stack([np.arange(n, dtype=np.float32) * 10, np.zeros(n), np.zeros(n)], 1).astype(np.float32)


def test_place_picks_the_nearest_when_unconstrained(orc):
    lab = _lab_line(5)
    cells = np.array([[0, 0, 0], [20, 0, 0], [40, 0, 0]], np.float32)
    r = orc.place(lab, cells, 3, 1, 3, 0.0)
    assert list(r["grid"]) == [0, 2, 4] and list(r["stage"]) == [0, 0, 0]


def test_place_bans_the_same_material_next_to_itself(orc):
    # src/main.rs:951-1024: same tile never adjacent; with weight 0 the next
    # nearest is taken
    lab = _lab_line(3)
    cells = np.zeros((4, 3), np.float32)  # all want material 0
    r = orc.place(lab, cells, 2, 2, 10, 0.0)
    g = r["grid"].reshape(2, 2)
    assert g[0, 0] == 0 and g[0, 1] == 1 and g[1, 0] == 1 and g[1, 1] == 0


def test_place_usage_limit_then_reset(orc):
    # UsageTracker (src/lib.rs:252-306): count < max; fallback resets all
    lab = _lab_line(2)
    cells = np.zeros((5, 3), np.float32)
    r = orc.place(lab, cells, 5, 1, 1, 0.0)
    # cell0 -> 0 (nearest), cell1: 0 used up and banned -> 1; cell2: both used
    # up -> reset, first placeable (not 1) -> 0; cell3 -> 1; cell4 -> reset -> 0
    assert list(r["grid"]) == [0, 1, 0, 1, 0]
    assert list(r["stage"]) == [0, 0, 1, 0, 1]
    assert r["stats"]["resets"] == 2 and list(r["usage"]) == [1, 0]


def test_place_max_usage_zero_falls_back_every_cell(orc):
    lab = _lab_line(3)
    r = orc.place(lab, np.zeros((6, 3), np.float32), 3, 2, 0, 0.3)
    assert r["stats"]["resets"] == 6 and r["stats"]["main_path"] == 0


def test_place_single_material_uses_nearest_one(orc):
    lab = _lab_line(1)
    r = orc.place(lab, np.zeros((3, 3), np.float32), 3, 1, 5, 0.3)
    # cell1 is banned next to cell0 with the only material: reset + nearest_one
    assert list(r["grid"]) == [0, 0, 0] and list(r["stage"]) == [0, 2, 2]


def test_place_penalty_prefers_dissimilar_neighbours(orc):
    # distance ties broken by the adjacency penalty: the candidate farther
    # (in Lab) from the left neighbour has the lower 1/(d+1) term
    lab = np.array([[0, 0, 0], [10, 0, 0], [-10, 0, 0], [30, 0, 0]], np.float32)
    cells = np.array([[30, 0, 0], [0, 0, 0]], np.float32)
    r = orc.place(lab, cells, 2, 1, 5, 0.25)
    assert r["grid"][0] == 3
    # cell1: 0 (d=0) beats 1 and 2 (d=100) regardless of the small penalty
    assert r["grid"][1] == 0
    cells2 = np.array([[30, 0, 0], [0.5, 0, 0]], np.float32)
    lab2 = np.array([[0, 0, 0], [1, 0, 0], [-1000, 0, 0], [30, 0, 0]], np.float32)
    r2 = orc.place(lab2, cells2, 2, 1, 5, 100.0)
    # 0 and 1 are equally close (0.25); material 0 is farther from 3 (30 vs 29)
    assert r2["grid"][1] == 0
    score = f32(0.25) + f32(f32(1.0) / f32(f32(30.0) + f32(1.0))) * f32(100.0)
    assert r2["score"][1] == score


def _place_bruteforce(lab, cell_lab, gw, gh, max_usage, weight, score_mode=0):
    """Independent restatement for tiny grids: full candidate scoring with
    numpy, ranks by lexsort, reset on failure (src/main.rs:335-475).
    score_mode 1 scores the Euclidean distance as the GUI generator does
    (src/gui/app_full.rs:1436-1448)."""
    n = lab.shape[0]
    K = min(n, 100)
    grid = np.full(gw * gh, -1)
    usage = np.zeros(n, int)
    for c in range(gw * gh):
        x, y = c % gw, c // gw
        diff = cell_lab[c] - lab
        sq = ((diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1]) + diff[:, 2] * diff[:, 2]).astype(np.float32)
        cand = np.lexsort((np.arange(n), sq))[:K]
        nbs = [(x, y - 1), (x, y + 1), (x - 1, y), (x + 1, y)]
        filled = [grid[ny * gw + nx] for nx, ny in nbs if 0 <= nx < gw and 0 <= ny < gh and grid[ny * gw + nx] >= 0]
        best = None
        for r, m in enumerate(cand):
            if usage[m] >= max_usage or m in filled:
                continue
            p = f32(0.0)
            if weight > 0:
                for o in filled:
                    d = lab[m] - lab[o]
                    s = f32(np.sqrt(f32((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))
                    p = f32(p + f32(f32(1.0) / f32(s + f32(1.0))))
                p = f32(p * f32(weight))
            d_m = f32(np.sqrt(sq[m])) if score_mode == 1 else sq[m]
            score = f32(d_m + p)
            if best is None or score < best[0]:
                best = (score, m)
        if best is None:
            usage[:] = 0
            pick = next((m for m in cand if m not in filled), cand[0])
        else:
            pick = best[1]
        usage[pick] += 1
        grid[c] = pick
    return grid


@pytest.mark.parametrize("score_mode", [0, 1])
@pytest.mark.parametrize("seed", range(6))
def test_place_matches_bruteforce_on_tiny_grids(orc, seed, score_mode):
    rng = _rng(100 + seed)
    n = int(rng.integers(1, 30))
    gw, gh = int(rng.integers(1, 7)), int(rng.integers(1, 6))
    lab = np.round(rng.uniform(-20, 20, (n, 3))).astype(np.float32)
    cl = np.round(rng.uniform(-20, 20, (gw * gh, 3))).astype(np.float32)
    usage, w = int(rng.integers(0, 4)), float(rng.choice([0.0, 0.25, 3.0]))
    got = orc.place(lab, cl, gw, gh, usage, w, score_mode=score_mode)["grid"]
    assert np.array_equal(got, _place_bruteforce(lab, cl, gw, gh, usage, w, score_mode))


def test_gui_and_cli_scores_pick_different_winners(orc):
    # material 1 is nearer (sqdist 1 vs 4) but identical to the left neighbour
    # 0 (penalty 4*1 vs 4*1/2): squared 1+4 < 4+2, Euclidean 1+4 > 2+2
    lab = np.array([[1, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    cl = np.array([[9, 9, 9], [0, 0, 0]], np.float32)
    for mode, pick, dist, score in [(0, 1, 1.0, 5.0), (1, 2, 2.0, 4.0)]:
        r = orc.place(lab, cl, 2, 1, 3, 4.0, order=np.array([1], np.int32),
                      grid=np.array([0, -1], np.int32), score_mode=mode)
        assert r["grid"][1] == pick and r["stage"][1] == 0
        assert r["dist"][1] == dist and r["score"][1] == score


def _gui_tiles():
    # the two tiles of the GUI generator tests (src/gui/app_full.rs:2813-2824)
    return np.array([[50, 0, 0], [75, 10, 5]], np.float32)


def _gui_place(orc, lab, gw, gh, max_usage, weight, cells, target=(50.0, 0.0, 0.0)):
    """find_best_tile_for_position at `cells` in turn with one target Lab
    (src/gui/app_full.rs:1347-1394); the reference tests use an empty
    similarity database, i.e. zero penalties, so weight 0 matches them."""
    cl = np.tile(np.array(target, np.float32), (gw * gh, 1))
    grid = np.full(gw * gh, -1, np.int32)
    usage = np.zeros(max(len(lab), 1), np.uint32)
    stages = []
    for c in cells:
        r = orc.place(lab, cl, gw, gh, max_usage, weight, order=np.array([c], np.int32),
                      grid=grid, usage=usage, score_mode=1)
        grid, usage = r["grid"], np.resize(r["usage"], max(len(lab), 1)).astype(np.uint32)
        stages.append(int(r["stage"][c]))
    return grid, stages


@pytest.mark.parametrize("weight", [0.0, 0.8])
def test_gui_comprehensive_fallback_scenario(orc, weight):
    # src/gui/app_full.rs:2804-2856: 2x2, usage limit 1, every cell filled;
    # (0,1) needs the usage reset (both tiles used, tile 0 is above it)
    grid, stages = _gui_place(orc, _gui_tiles(), 2, 2, 1, weight, [0, 1, 2, 3])
    assert grid.tolist() == [0, 1, 1, 0]
    assert stages == [0, 0, 1, 0]


def test_gui_usage_tracker_reset_in_fallback(orc):
    # src/gui/app_full.rs:3132-3165: one tile, limit 1, (0,0) then (1,1)
    grid, stages = _gui_place(orc, _gui_tiles()[:1], 2, 2, 1, 0.3, [0, 3])
    assert grid.tolist() == [0, -1, -1, 0] and stages == [0, 1]


def test_gui_fallback_with_adjacency_constraints(orc):
    # src/gui/app_full.rs:2682-2723: limit 1, (0,0) then (1,0)
    grid, stages = _gui_place(orc, _gui_tiles(), 2, 2, 1, 0.3, [0, 1])
    assert grid.tolist() == [0, 1, -1, -1] and stages == [0, 0]


def test_gui_final_fallback_and_empty_library(orc):
    # src/gui/app_full.rs:2725-2753: the only tile is banned by its
    # neighbour, so nearest_one places it anyway; :2640-2654: no tiles -> None
    grid, stages = _gui_place(orc, _gui_tiles()[:1], 2, 2, 3, 0.3, [0, 1])
    assert grid.tolist() == [0, 0, -1, -1] and stages == [0, 2]
    r = orc.place(np.zeros((0, 3), np.float32), np.zeros((4, 3), np.float32), 2, 2, 3, 0.3,
                  score_mode=1)
    assert r["grid"].tolist() == [-1] * 4 and r["stats"]["empty"] == 4


# ------------------------------------------------ a10-a12 cost, delta, SA
def test_total_cost_known_answer(orc):
    # src/adjacency.rs:274-290: identical tiles side by side -> cost 1 > 0.5
    lab = np.array([[50, 0, 0], [50, 0, 0]], np.float32)
    assert orc.total_cost(lab, np.array([0, 1, -1, -1], np.int32), 2, 2) == 1.0
    lab3 = np.array([[50, 0, 0], [60, 10, 10], [40, -10, -10]], np.float32)
    g = np.array([0, 1, 2, -1], np.int32)  # right(0,1) + down(0,2)
    want = f32(f32(1) / f32(np.sqrt(f32(300)) + f32(1))) + f32(f32(1) / f32(np.sqrt(f32(300)) + f32(1)))
    assert orc.total_cost(lab3, g, 2, 2) == want


def test_swap_delta_equals_cost_difference(orc):
    rng = _rng(11)
    for _ in range(300):
        n, gw, gh = int(rng.integers(2, 12)), int(rng.integers(1, 6)), int(rng.integers(1, 6))
        lab = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
        g = rng.integers(-1, n, gw * gh).astype(np.int32)
        x1, y1, x2, y2 = (int(rng.integers(0, gw)), int(rng.integers(0, gh)),
                          int(rng.integers(0, gw)), int(rng.integers(0, gh)))
        d = orc.swap_delta(lab, g, gw, gh, x1, y1, x2, y2)
        g2 = g.copy()
        g2[y1 * gw + x1], g2[y2 * gw + x2] = g[y2 * gw + x2], g[y1 * gw + x1]
        if g[y1 * gw + x1] < 0 or g[y2 * gw + x2] < 0:
            assert d == 0.0
            continue
        diff = float(orc.total_cost(lab, g2, gw, gh)) - float(orc.total_cost(lab, g, gw, gh))
        assert abs(float(d) - diff) <= 1e-4 * max(1.0, gw * gh), (d, diff)


def test_anneal_bookkeeping_invariants(orc):
    rng = _rng(12)
    n, gw, gh = 40, 9, 7
    lab = rng.uniform(-40, 40, (n, 3)).astype(np.float32)
    g = rng.integers(0, n, gw * gh).astype(np.int32)
    g[rng.random(gw * gh) < 0.1] = -1
    r = orc.anneal(lab, g, gw, gh, 3000, 2.0, 0.999, seed=5, trace=True)
    res, td, ta = r["result"], r["trace_delta"], r["trace_accept"]
    assert res["iterations"] == 3000
    assert res["evaluated"] == int(np.sum(~np.isnan(td)))
    assert res["accepted_count"] == int(ta.sum())
    # running cost: initial + accepted deltas in proposal order (f32)
    cost = f32(res["initial_cost"])
    best, improved = cost, 0
    for i in np.nonzero(ta)[0]:
        cost = f32(cost + td[i])
        if cost < best:
            best, improved = cost, improved + 1
    assert cost == f32(res["final_cost"]) and best == f32(res["best_cost"])
    assert improved == res["improved_count"]
    assert res["initial_cost"] == orc.total_cost(lab, g, gw, gh)
    assert sorted(r["grid"].tolist()) == sorted(g.tolist())  # swaps permute
    assert abs(orc.total_cost(lab, r["grid"], gw, gh) - res["final_cost"]) < 1e-2


def test_anneal_skips_do_not_cool(orc):
    # src/optimizer.rs:66-73 vs :102: recompute the accept decisions from the
    # trace with T advancing only on evaluated proposals
    rng = _rng(13)
    n, gw, gh = 10, 3, 3
    lab = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    g = rng.integers(0, n, 9).astype(np.int32)
    g[4] = -1
    t0, decay, seed = 0.05, 0.97, 9
    r = orc.anneal(lab, g, gw, gh, 400, t0, decay, seed=seed, trace=True)
    T = f32(t0)
    for it in range(400):
        d = r["trace_delta"][it]
        if np.isnan(d):
            continue
        u = f32(orc.draw(seed, it, 4) >> 8) * f32(1.0 / 16777216.0)
        want = bool(d < 0 or u < orc.expf(f32(-d) / T))
        assert bool(r["trace_accept"][it]) == want, it
        T = f32(T * f32(decay))


def test_anneal_edge_cases(orc):
    lab = np.array([[50, 0, 0], [60, 10, 10], [40, -10, -10], [55, 5, 5]], np.float32)
    g = np.array([0, 1, 2, 3], np.int32)
    r = orc.anneal(lab, g, 2, 2, 0)  # src/optimizer.rs:440-457
    assert r["result"]["initial_cost"] == r["result"]["final_cost"]
    assert r["result"]["accepted_count"] == 0 and r["result"]["improved_count"] == 0
    one = orc.anneal(lab[:1], np.array([0], np.int32), 1, 1, 100)  # :296-322
    assert one["result"]["initial_cost"] == 0.0 and one["result"]["accepted_count"] == 0
    sparse = np.full(9, -1, np.int32)
    sparse[0], sparse[8] = 0, 1
    assert orc.anneal(lab, sparse, 3, 3, 50)["result"]["initial_cost"] == 0.0  # :336-357
    gr = orc.anneal(lab, g, 2, 2, 200, greedy=True)["result"]  # :264-277
    assert gr["final_cost"] <= gr["initial_cost"] and gr["improved_count"] == gr["accepted_count"]
    empty = orc.anneal(lab, np.zeros(0, np.int32), 0, 0, 10)["result"]  # :294-303
    assert empty["initial_cost"] == 0.0 and empty["iterations"] == 0


# ---------------------------------------------- a13-a16 adjust/composite
def test_average_rgb_known_answers(orc):
    img = np.array([[[255, 0, 0], [0, 255, 0]], [[0, 0, 255], [255, 255, 255]]], np.uint8)
    assert list(orc.average_rgb(img)) == [127, 127, 127]  # src/main.rs:877-889
    red = np.zeros((10, 10, 3), np.uint8)
    red[..., 0] = 255
    assert list(orc.average_rgb(red)) == [255, 0, 0]  # :866-875


def test_adjustment_new_clamps(orc):
    # src/color_adjustment.rs:170-176, :445-452, :434-443
    assert list(orc.adjustment_new(2.0, -1.0, 400.0, 3.0)) == [1.0, 0.0, 180.0, 2.0]
    assert list(orc.adjustment_new(-2.0, 5.0, -500.0, 10.0)) == [-1.0, 2.0, -180.0, 2.0]
    assert orc.adjustment_new(0.0, 1.0, -100.0, 1.0)[2] == -100.0


def test_hsv_matches_the_hexcone_model(orc):
    for rgb in _rng(14).uniform(0, 1, (3000, 3)).astype(np.float32):
        h, s, v = orc.rgb_to_hsv(*rgb)
        ch, cs, cv = colorsys.rgb_to_hsv(*map(float, rgb))
        assert abs((h % 360.0) - ch * 360.0) < 1e-3 or abs(abs((h % 360.0) - ch * 360.0) - 360) < 1e-3
        assert abs(s - cs) < 1e-6 and abs(v - cv) < 1e-6
        back = orc.hsv_to_rgb(h, s, v)
        assert np.max(np.abs(back - rgb)) < 1e-5
    assert orc.rgb_to_hsv(0.5, 0.5, 0.5)[1] == 0.0


def test_adjust_pixel_reference_cases(orc):
    ident = orc.adjustment_new(0.0, 1.0, 0.0, 1.0)
    px = orc.adjust_pixel(ident, [128, 64, 192])  # :178-192
    assert np.max(np.abs(px.astype(int) - [128, 64, 192])) <= 1
    assert np.all(orc.adjust_pixel(orc.adjustment_new(0.2, 1.0, 0, 1), [100, 100, 100]) > 100)
    c15 = orc.adjustment_new(0.0, 1.5, 0, 1)
    assert orc.adjust_pixel(c15, [50, 50, 50])[0] < 50 and orc.adjust_pixel(c15, [200, 200, 200])[0] > 200
    g = orc.adjust_pixel(orc.adjustment_new(0, 1, 120, 1), [255, 0, 0])  # :272-283
    assert g[1] > g[0] and g[1] > g[2]
    s = orc.adjust_pixel(orc.adjustment_new(0, 1, 0, 2.0), [128, 128, 128])  # :285-305
    assert np.max(np.abs(s.astype(int) - 128)) <= 5
    d = orc.adjust_pixel(orc.adjustment_new(0, 1, 0, 0.5), [255, 100, 50])
    assert int(d.max()) - int(d.min()) < 205
    ext = orc.adjustment_new(0.5, 1.5, 60.0, 1.5)  # :476-498
    assert np.all(orc.adjust_pixel(ext, [255, 255, 255]) >= 250)
    assert np.all(orc.adjust_pixel(ext, [0, 0, 0]) > 0)


def _bc_closed_form(levels, brightness, contrast):
    """adjust_pixel with hue_shift 0 and saturation 1 (no HSV round trip),
    src/color_adjustment.rs:53-96, restated in numpy float32 one rounded
    operation at a time: v = x/255; c = clamp((v - 0.5)*contrast + 0.5, 0, 1);
    c = clamp(c + brightness, 0, 1); then `(c * 255).clamp(0, 255) as u8`,
    Rust's saturating float->int cast, which truncates toward zero.  Returns
    the u8 result and the pre-cast float."""
    f = np.float32
    v = levels.astype(f) / f(255.0)
    c = np.clip((v - f(0.5)) * f(contrast) + f(0.5), f(0.0), f(1.0)).astype(f)
    c = np.clip(c + f(brightness), f(0.0), f(1.0)).astype(f)
    t = np.clip(c * f(255.0), f(0.0), f(255.0)).astype(f)
    return np.trunc(t).astype(np.uint8), t


def test_adjust_pixel_u8_conversion_truncates(orc):
    # All 256 levels on the brightness/contrast-only path against a numpy
    # closed form of src/color_adjustment.rs:82-96.  The cases are chosen so
    # that rounding and truncation disagree on many levels: an oracle that
    # rounded (or took floor(t + 0.5)) fails here.
    levels = np.arange(256, dtype=np.uint8)
    px = np.repeat(levels[:, None], 3, axis=1)
    disagree = 0
    for brightness, contrast in ((0.1, 1.0), (-0.139, 1.0), (0.0, 1.3), (0.05, 0.77),
                                 (0.3, 1.9), (-0.42, 0.5), (0.0, 1.0)):
        adj = orc.adjustment_new(brightness, contrast, 0.0, 1.0)
        want, t = _bc_closed_form(levels, adj[0], adj[1])
        got = orc.apply_adjustment(adj, px)
        np.testing.assert_array_equal(got[:, 0], want, err_msg=f"b={brightness} c={contrast}")
        np.testing.assert_array_equal(got[:, 1], want)
        np.testing.assert_array_equal(got[:, 2], want)
        disagree += int(np.count_nonzero(np.rint(t) != np.trunc(t)))
    assert disagree > 200  # the cases separate truncation from rounding
    # identity: x/255*255 truncates back to x for every level (the reference's
    # identity test allows +-1, src/color_adjustment.rs:178-192)
    ident = orc.apply_adjustment(orc.adjustment_new(0.0, 1.0, 0.0, 1.0), px)
    np.testing.assert_array_equal(ident[:, 0], _bc_closed_form(levels, 0.0, 1.0)[0])


def test_optimal_adjustment_reference_cases(orc):
    # src/color_adjustment.rs:230-240, :335-470
    a = orc.optimal_adjustment([128, 128, 128], [128, 128, 128], 1.0)
    assert abs(a[0]) < 0.01 and abs(a[1] - 1) < 0.01 and abs(a[2]) < 0.01 and abs(a[3] - 1) < 0.01
    b = orc.optimal_adjustment([50, 50, 50], [200, 200, 200], 1.0)
    assert 0 < b[0] <= 1.0
    assert -1.0 <= orc.optimal_adjustment([200, 200, 200], [50, 50, 50], 1.0)[0] < 0
    h = orc.optimal_adjustment([255, 0, 0], [0, 255, 0], 1.0)
    assert 10 < abs(h[2]) <= 180
    assert orc.optimal_adjustment([255, 0, 0], [200, 150, 150], 1.0)[3] < 1.0
    full = orc.optimal_adjustment([100, 100, 100], [200, 200, 200], 1.0)
    half = orc.optimal_adjustment([100, 100, 100], [200, 200, 200], 0.5)
    assert abs(half[0] - full[0] * 0.5) < 0.01
    gr = orc.optimal_adjustment([128, 128, 128], [64, 64, 64], 1.0)
    assert gr[0] < 0 and abs(gr[2]) < 5 and abs(gr[3] - 1) < 0.1
    z = orc.optimal_adjustment([255, 0, 0], [0, 255, 0], 0.0)
    assert abs(z[0]) < 0.01 and abs(z[2]) < 0.01 and abs(z[3] - 1) < 0.01
    assert abs(orc.optimal_adjustment([255, 0, 0], [255, 0, 255], 1.0)[2]) <= 180
    assert abs(orc.optimal_adjustment([100, 95, 105], [105, 95, 100], 1.0)[2]) < 10
    # closed form of the brightness term
    lt = (0.299 * 100 + 0.587 * 100 + 0.114 * 100) / 255
    lg = (0.299 * 200 + 0.587 * 200 + 0.114 * 200) / 255
    assert abs(full[0] - (lg - lt)) < 1e-6


def _lanczos_f64(src, dw, dh):
    """Independent float64 separable Lanczos3 (same support and centring);
    the horizontal pass is rounded and clipped to u8 like any two-pass u8
    resampler, the rest stays in float64."""
    def axis(insz, outsz):
        sc = insz / outsz
        fs = max(sc, 1.0)
        W = np.zeros((outsz, insz))
        for o in range(outsz):
            cen = (o + 0.5) * sc
            xs = np.arange(insz)
            t = (xs - cen + 0.5) / fs
            w = np.where(np.abs(t) < 3, np.sinc(t) * np.sinc(t / 3), 0.0)
            W[o] = w / w.sum()
        return W
    sh, sw = src.shape[:2]
    Wx, Wy = axis(sw, dw), axis(sh, dh)
    tmp = np.clip(np.rint(np.einsum("yxc,px->ypc", src.astype(np.float64), Wx)), 0, 255)
    return np.einsum("oy,ypc->opc", Wy, tmp)


@pytest.mark.parametrize("sw,sh,dw,dh", [(64, 64, 16, 16), (128, 128, 24, 18), (30, 20, 60, 40)])
def test_resize_matches_float_lanczos3(orc, sw, sh, dw, dh):
    src = _rng(sw + dw).integers(0, 256, (sh, sw, 3)).astype(np.uint8)
    got = orc.resize(src, dw, dh).astype(np.float64)
    want = np.clip(_lanczos_f64(src, dw, dh), 0, 255)
    assert np.max(np.abs(got - want)) <= 2.0  # two u8 roundings + 16-bit taps


@pytest.mark.parametrize("sw,sh,dw,dh", [(100, 100, 50, 50), (128, 128, 24, 18), (30, 20, 61, 43),
                                         (7, 300, 3, 11)])
def test_resize_keeps_every_solid_level(orc, sw, sh, dw, dh):
    # the rounded taps of an output pixel sum to 2^p within a few units, and
    # the 2^(p-1) offset absorbs that: a flat image stays flat at every level
    # (the reference's own test checks red, src/main.rs:892-908)
    lv = np.arange(256, dtype=np.uint8)
    for c in range(3):
        src = np.zeros((sh, sw, 3), np.uint8)
        for v in lv[::5]:
            src[..., c] = v
            out = orc.resize(src, dw, dh)
            assert np.all(out[..., c] == v) and np.all(np.delete(out, c, axis=2) == 0)


def test_resize_identity_and_solid(orc):
    src = _rng(15).integers(0, 256, (9, 13, 3)).astype(np.uint8)
    assert np.array_equal(orc.resize(src, 13, 9), src)
    red = np.zeros((100, 100, 3), np.uint8)
    red[..., 0] = 255
    out = orc.resize(red, 50, 50)  # src/main.rs:892-908
    assert np.all(out == [255, 0, 0])


def test_composite_layout(orc):
    rng = _rng(16)
    mats = rng.integers(0, 256, (3, 8, 8, 3)).astype(np.uint8)
    target = rng.integers(0, 256, (13, 17, 3)).astype(np.uint8)
    grid = np.array([0, 1, 2, -1, 2, 1], np.int32)  # 3 x 2 cells of 5 x 6
    out = orc.composite(target, mats, grid, 3, 2, 0.0)
    assert out.shape == (12, 15, 3)
    assert np.all(out[6:12, 0:5] == 0)  # empty cell stays black (src/main.rs:515-517)
    assert np.array_equal(out[0:6, 5:10], orc.resize(mats[1], 5, 6))
    adj = orc.composite(target, mats, grid, 3, 2, 0.4)
    tile = orc.resize(mats[2], 5, 6)
    a = orc.optimal_adjustment(orc.average_rgb(tile), orc.average_rgb(target[0:6, 10:15]), 0.4)
    assert np.array_equal(adj[0:6, 10:15], orc.apply_adjustment(a, tile))
