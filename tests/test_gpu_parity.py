"""Parity of the HIP path (through the C ABI) against the C oracle.

Bar: bit-exact for Lab means, similarities, candidate lists, placement
(indices, usage, stages, distances, scores, stats), costs, swap deltas and
annealing (the counter-RNG chain is reproduced exactly); composited pixels
within +-1 LSB of u8 (they are expected to be exact).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape
    diff = np.nonzero(_bits(a) != _bits(b))
    assert diff[0].size == 0, f"{diff[0].size} mismatches, first at {diff[0][:5]}: {a[diff][:5]} vs {b[diff][:5]}"


# ------------------------------------------------------------------- a1 Lab
def test_average_lab_known_answers(ctx, orc):
    # src/lib.rs:133-158
    red = np.zeros((10, 10, 3), np.uint8)
    red[..., 0] = 255
    lab = ctx.average_lab(red)
    assert abs(lab[0] - 53.24) < 1.0 and abs(lab[1] - 80.09) < 1.0 and abs(lab[2] - 67.20) < 1.0
    assert_bits_equal(lab, orc.average_lab(red))
    grey = np.full((10, 10, 3), 128, np.uint8)
    lab = ctx.average_lab(grey)
    assert abs(lab[0] - 53.59) < 1.0 and abs(lab[1]) < 1.0 and abs(lab[2]) < 1.0
    assert_bits_equal(lab, orc.average_lab(grey))


def test_every_rgb_value_converts_like_oracle(ctx, orc):
    # all 256 grey levels + 4096 random colours as 1-pixel images in a strip
    rng = _rng(1)
    cols = np.concatenate([np.repeat(np.arange(256, dtype=np.uint8)[:, None], 3, 1),
                           rng.integers(0, 256, (4096, 3)).astype(np.uint8)])
    img = cols[None, :, :]
    got = ctx.cell_lab_means(img, cols.shape[0], 1)
    want = orc.cell_lab_means(img, cols.shape[0], 1)
    assert_bits_equal(got, want)


def test_every_colour_converts_like_oracle(ctx, orc):
    # all 2^24 sRGB colours, once each, as the 1x1 cells of a 4096 x 4096
    # image: the per-pixel conversion is bit-exact everywhere it can be fed
    v = np.arange(1 << 24, dtype=np.uint32)
    img = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], -1).astype(np.uint8)
    img = img.reshape(4096, 4096, 3)
    got = ctx.cell_lab_means(img, 4096, 4096)
    want = orc.cell_lab_means(img, 4096, 4096)
    bad = np.nonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))[0]
    assert bad.size == 0, f"{bad.size} colours differ, first {bad[:8]}"


# the last two have 262,144 and 360,000 cells (the launch shape of the
# served batches: 16 cells per block)
@pytest.mark.parametrize("w,h,gw,gh", [(640, 480, 40, 30), (101, 77, 7, 5), (64, 64, 1, 1),
                                       (1000, 37, 250, 37), (2048, 2048, 512, 512),
                                       (6000, 1500, 1200, 300)])
def test_cell_lab_means(ctx, orc, w, h, gw, gh):
    img = _rng(w * h).integers(0, 256, (h, w, 3)).astype(np.uint8)
    assert_bits_equal(ctx.cell_lab_means(img, gw, gh), orc.cell_lab_means(img, gw, gh))


def test_material_lab_from_pixels(ctx, orc, synth):
    mats = synth.flat_materials(300, 64, _rng(2))
    lib = ctx.library_from_pixels(mats)
    assert_bits_equal(lib.lab(), orc.stack_lab_means(mats))
    lib.close()


# ------------------------------------------------------------ a2 similarity
@pytest.mark.parametrize("n", [1, 3, 200, 201, 1500])
def test_similarity_matrix(ctx, orc, n):
    lab = (_rng(n).random((n, 3)) * [100, 200, 200] - [0, 100, 100]).astype(np.float32)
    lib = ctx.library_from_lab(lab)
    full = lib.similarities()
    assert_bits_equal(full, orc.sim_full(lab))
    assert_bits_equal(full, full.T)
    assert np.all(np.diag(full) == 0.0)
    if n > 1:
        assert_bits_equal(lib.similarities(packed=True), orc.sim_packed(lab))
    lib.close()


def test_similarity_known_answers(ctx):
    # src/similarity.rs:186-205, :320-362
    lab = np.array([[50, 0, 0], [60, 10, 10], [40, -10, -10]], np.float32)
    lib = ctx.library_from_lab(lab)
    s12 = lib.get_similarity(0, 1)
    assert s12 == np.float32(np.sqrt(np.float32(300.0)))
    assert lib.get_similarity(1, 0) == s12
    assert lib.get_similarity(0, 0) == 0.0
    assert lib.get_similarity(0, 3) is None
    lib.close()


# -------------------------------------------------------------- a3 top-K
@pytest.mark.parametrize("n,k", [(200, 100), (1500, 100), (50, 100), (128, 100), (129, 100),
                                 (10000, 100), (3000, 7)])
def test_nearest_n(ctx, orc, n, k):
    rng = _rng(n + k)
    lab = (rng.random((n, 3)) * [100, 200, 200] - [0, 100, 100]).astype(np.float32)
    q = (rng.random((64, 3)) * [100, 200, 200] - [0, 100, 100]).astype(np.float32)
    lib = ctx.library_from_lab(lab)
    gi, gd = ctx.nearest_n(lib, q, k)
    oi, od = orc.nearest_n(lab, q, k)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(gd, od)
    lib.close()


@pytest.mark.parametrize("n,nq,walk", [(1500, 20000, True), (1500, 20000, False),
                                       (300, 70000, True), (5000, 17000, True)])
def test_nearest_n_runs_of_cells(ctx, orc, n, nq, walk):
    # enough queries that each wave answers a run of cells, the later ones
    # bounded by the previous cell's answers; neighbouring cells of an image
    # are close in Lab (a random walk) or, in the worst case, unrelated
    rng = _rng(n + nq)
    lab = (rng.random((n, 3)) * [100, 200, 200] - [0, 100, 100]).astype(np.float32)
    if walk:
        q = np.cumsum(rng.normal(0, 1.5, (nq, 3)), axis=0) % [100, 200, 200] - [0, 100, 100]
    else:
        q = rng.random((nq, 3)) * [100, 200, 200] - [0, 100, 100]
    q = q.astype(np.float32)
    lib = ctx.library_from_lab(lab)
    gi, gd = ctx.nearest_n(lib, q, 100)
    oi, od = orc.nearest_n(lab, q, 100)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(gd, od)
    lib.close()


@pytest.mark.parametrize("n,k,kind", [(255, 100, "mix"), (256, 128, "mix"), (257, 100, "mix"),
                                      (257, 1, "mix"), (320, 65, "dup"), (4097, 100, "mix"),
                                      (4097, 64, "far"), (10000, 128, "dup"), (2500, 100, "far")])
def test_nearest_n_blocks(ctx, orc, n, k, kind):
    # the block layout's edges: libraries just below, at and above the
    # 256-material buffer and the 64-material blocks, clustered colours (most
    # blocks out of reach), queries far from every cluster (many blocks in
    # reach, loose thresholds) and exact duplicates split across blocks
    rng = _rng(n * 7 + k)
    means = rng.random((3, 3)) * [100, 200, 200] - [0, 100, 100]
    lab = means[rng.integers(0, 3, n)] + rng.normal(0, 4.0, (n, 3))
    if kind == "dup":
        lab[n // 2:] = lab[: n - n // 2][rng.permutation(n - n // 2)]
    lab = lab.astype(np.float32)
    if kind == "far":
        q = rng.random((2000, 3)) * [100, 200, 200] - [0, 100, 100]
    else:
        q = np.cumsum(rng.normal(0, 1.0, (2000, 3)), axis=0) % [100, 200, 200] - [0, 100, 100]
    q = q.astype(np.float32)
    lib = ctx.library_from_lab(lab)
    gi, gd = ctx.nearest_n(lib, q, k)
    oi, od = orc.nearest_n(lab, q, k)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(gd, od)
    lib.close()


def test_nearest_n_ties_by_index(ctx, orc):
    # many exact duplicates: equal squared distances must come out by index
    rng = _rng(7)
    base = (rng.random((40, 3)) * 50).astype(np.float32)
    lab = np.concatenate([base] * 25)  # 1000 materials, every colour 25 times
    perm = rng.permutation(lab.shape[0])
    lab = lab[perm]
    q = base[:16] + np.float32(0.25)
    lib = ctx.library_from_lab(lab)
    gi, gd = ctx.nearest_n(lib, q, 100)
    oi, od = orc.nearest_n(lab, q, 100)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(gd, od)
    for r in range(gi.shape[0]):
        keys = list(zip(gd[r].tolist(), gi[r].tolist()))
        assert keys == sorted(keys)
    lib.close()


def test_nearest_one_known_answer(ctx):
    # src/lib.rs:213-227
    lab = np.array([[50, 0, 0], [53.24, 80.09, 67.20], [87.74, -86.18, 83.18],
                    [32.30, 79.20, -107.86]], np.float32)
    lib = ctx.library_from_lab(lab)
    gi, _ = ctx.nearest_n(lib, np.array([[53.0, 80.0, 67.0]], np.float32), 1)
    assert gi[0, 0] == 1
    lib.close()


# ------------------------------------------------------------ a4-a9 place
def _check_place(got, want):
    np.testing.assert_array_equal(got["grid"], want["grid"])
    np.testing.assert_array_equal(got["usage"], want["usage"])
    np.testing.assert_array_equal(got["stage"], want["stage"])
    assert_bits_equal(got["dist"], want["dist"])
    assert_bits_equal(got["score"], want["score"])
    assert got["stats"] == want["stats"]


@pytest.mark.parametrize("score_mode", [0, 1])
@pytest.mark.parametrize("name", ["small", "medium"])
def test_place_workloads(ctx, orc, synth, name, score_mode):
    p, target, mats = synth.make(name)
    lab = orc.stack_lab_means(mats)
    cl = orc.cell_lab_means(target, p["grid_w"], p["grid_h"])
    lib = ctx.library_from_lab(lab)
    got = ctx.place(lib, cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                    score_mode=score_mode)
    want = orc.place(lab, cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                     score_mode=score_mode)
    _check_place(got, want)
    if name == "small":  # 200 x 3 = 600 < 1200 cells: the reset path must run
        assert want["stats"]["resets"] >= 1
    lib.close()


@pytest.mark.parametrize("n,gw,gh,usage,weight", [
    (12, 6, 5, 2, 0.25),   # forces resets (24 < 30)
    (12, 6, 5, 0, 0.25),   # max_usage = 0: fallback on every cell
    (1, 4, 3, 3, 0.3),     # single material: nearest_one after the ban
    (2, 5, 4, 1, 0.5),     # two materials, ban + usage
    (3, 7, 7, 2, 0.0),     # weight 0: no penalty
    (150, 13, 11, 1, -1.0),  # negative weight behaves like 0
    (500, 20, 20, 3, 2.0),
    (300, 1, 40, 2, 4.0),   # one column: the up neighbour is the previous winner
    (300, 2, 30, 2, 4.0),   # two columns
    (300, 3, 25, 2, 4.0),   # three columns: the narrowest grid with the read-ahead chain
])
@pytest.mark.parametrize("score_mode", [0, 1])
def test_place_edge_cases(ctx, orc, n, gw, gh, usage, weight, score_mode):
    rng = _rng(n * 131 + gw)
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    cl = (rng.random((gw * gh, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    lib = ctx.library_from_lab(lab)
    got = ctx.place(lib, cl, gw, gh, usage, weight, score_mode=score_mode)
    want = orc.place(lab, cl, gw, gh, usage, weight, score_mode=score_mode)
    _check_place(got, want)
    lib.close()


def test_place_empty_library(ctx, orc):
    lib = ctx.library_from_lab(np.zeros((0, 3), np.float32))
    cl = np.zeros((12, 3), np.float32)
    got = ctx.place(lib, cl, 4, 3, 3, 0.25)
    assert np.all(got["grid"] == -1)
    assert got["stats"]["empty"] == 12 and got["stats"]["placed"] == 0
    lib.close()


@pytest.mark.parametrize("score_mode", [0, 1])
def test_place_order_and_prefilled_state(ctx, orc, score_mode):
    # find_and_use_best_tile_with_position at arbitrary cells with a
    # pre-filled grid and usage (src/main.rs:1250-1283): all 4 neighbours count
    rng = _rng(11)
    n, gw, gh = 40, 6, 6
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    cl = (rng.random((gw * gh, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    grid = np.full(gw * gh, -1, np.int32)
    grid[rng.choice(gw * gh, 10, replace=False)] = rng.integers(0, n, 10)
    usage = rng.integers(0, 3, n).astype(np.uint32)
    order = rng.permutation(np.nonzero(grid < 0)[0]).astype(np.int32)
    lib = ctx.library_from_lab(lab)
    got = ctx.place(lib, cl, gw, gh, 3, 0.5, order=order, grid=grid, usage=usage,
                    score_mode=score_mode)
    want = orc.place(lab, cl, gw, gh, 3, 0.5, order=order, grid=grid, usage=usage,
                     score_mode=score_mode)
    _check_place(got, want)
    lib.close()


@pytest.mark.parametrize("raster", [True, False])
def test_place_gui_score_changes_the_winner(ctx, raster):
    # material 1 is nearer (sqdist 1 vs 4) but identical to the left
    # neighbour 0, so its penalty is 4*1 against material 2's 4*(1/2):
    # squared scores 1+4 < 4+2 pick 1, Euclidean scores 1+4 > 2+2 pick 2
    # (src/main.rs:385-398 vs src/gui/app_full.rs:1436-1448)
    lab = np.array([[1, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    lib = ctx.library_from_lab(lab)
    for mode, pick, dist, score in [(0, 1, 1.0, 5.0), (1, 2, 2.0, 4.0)]:
        if raster:
            # raster chain: cell 0 (target = material 0's Lab) takes material 0
            cl = np.array([[1, 0, 0], [0, 0, 0]], np.float32)
            got = ctx.place(lib, cl, 2, 1, 1, 4.0, score_mode=mode)
            assert got["grid"][0] == 0
        else:
            cl = np.array([[9, 9, 9], [0, 0, 0]], np.float32)
            got = ctx.place(lib, cl, 2, 1, 3, 4.0, order=np.array([1], np.int32),
                            grid=np.array([0, -1], np.int32), score_mode=mode)
        assert got["grid"][1] == pick and got["stage"][1] == 0
        assert got["dist"][1] == np.float32(dist) and got["score"][1] == np.float32(score)
    lib.close()


# ------------------------------------------------- a10-a12 cost, delta, SA
def _random_grid(rng, n, gw, gh, empty_frac=0.1):
    g = rng.integers(0, n, gw * gh).astype(np.int32)
    g[rng.random(gw * gh) < empty_frac] = -1
    return g


@pytest.mark.parametrize("n,gw,gh", [(5, 4, 4), (200, 40, 30), (1500, 80, 60), (3, 1, 9)])
def test_total_cost_and_penalties(ctx, orc, n, gw, gh):
    rng = _rng(n + gw)
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    g = _random_grid(rng, n, gw, gh)
    lib = ctx.library_from_lab(lab)
    assert _bits(np.float32(ctx.total_cost(lib, g, gw, gh))) == _bits(orc.total_cost(lab, g, gw, gh))
    q = np.stack([rng.integers(0, n, 256), rng.integers(0, gw, 256), rng.integers(0, gh, 256)], 1)
    got = ctx.penalties(lib, g, gw, gh, 0.25, q)
    sim = orc.sim_full(lab)
    want = []
    for m, x, y in q:  # calculate_penalty restated (Up, Down, Left, Right)
        p = np.float32(0.0)
        for (nx, ny) in [(x, y - 1), (x, y + 1), (x - 1, y), (x + 1, y)]:
            if 0 <= nx < gw and 0 <= ny < gh and g[ny * gw + nx] >= 0:
                p = np.float32(p + np.float32(1.0) / np.float32(sim[m, g[ny * gw + nx]] + np.float32(1.0)))
        want.append(np.float32(p * np.float32(0.25)))
    assert_bits_equal(got, np.array(want, np.float32))
    lib.close()


@pytest.mark.parametrize("n,gw,gh", [(4, 2, 2), (25, 5, 5), (200, 40, 30), (6, 3, 1)])
def test_swap_deltas(ctx, orc, n, gw, gh):
    rng = _rng(n * 7 + gh)
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    g = _random_grid(rng, n, gw, gh)
    pairs = np.stack([rng.integers(0, gw, 500), rng.integers(0, gh, 500),
                      rng.integers(0, gw, 500), rng.integers(0, gh, 500)], 1).astype(np.int32)
    lib = ctx.library_from_lab(lab)
    got = ctx.swap_deltas(lib, g, gw, gh, pairs)
    want = np.array([orc.swap_delta(lab, g, gw, gh, *map(int, p)) for p in pairs], np.float32)
    assert_bits_equal(got, want)
    lib.close()


def _check_anneal(got, want):
    np.testing.assert_array_equal(got["grid"], want["grid"])
    for key in ["initial_cost", "final_cost", "best_cost"]:
        assert _bits(np.float32(got["result"][key])) == _bits(np.float32(want["result"][key])), key
    for key in ["improved_count", "accepted_count", "iterations", "evaluated"]:
        assert got["result"][key] == want["result"][key], key
    if "trace_delta" in want:
        assert_bits_equal(got["trace_delta"], want["trace_delta"])
        np.testing.assert_array_equal(got["trace_accept"], want["trace_accept"])


@pytest.mark.parametrize("n,gw,gh,iters,t0,decay,greedy,seed", [
    (200, 40, 30, 500, 100.0, 0.99995, False, 0),
    (1500, 80, 60, 2000, 100.0, 0.99995, False, 1),
    (25, 5, 5, 300, 100.0, 0.99995, False, 2),
    (4, 2, 2, 50, 0.001, 0.99, False, 3),       # low temperature
    (4, 2, 2, 10, 100.0, 0.0, False, 4),        # T -> 0: -d/0 edge cases
    (4, 2, 2, 50, 1000.0, 0.5, False, 5),       # very fast cooling
    (200, 40, 30, 3000, 100.0, 0.99995, True, 6),  # optimize_greedy
    (9, 3, 3, 0, 100.0, 0.99995, False, 7),     # zero iterations
    (1, 1, 1, 100, 100.0, 0.99995, False, 8),   # single tile
    (30, 17, 9, 4000, 5.0, 0.999, False, 9),
])
def test_anneal_matches_sequential_chain(ctx, orc, n, gw, gh, iters, t0, decay, greedy, seed):
    rng = _rng(seed + 100)
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    g = _random_grid(rng, n, gw, gh, 0.05 if gw * gh > 4 else 0.0)
    lib = ctx.library_from_lab(lab)
    got = ctx.anneal(lib, g, gw, gh, iters, t0, decay, seed, greedy, trace=True)
    want = orc.anneal(lab, g, gw, gh, iters, t0, decay, seed, greedy, trace=True)
    _check_anneal(got, want)
    # the recorded cost equals the cost function on the final grid up to f32 drift
    final = orc.total_cost(lab, got["grid"], gw, gh)
    assert abs(final - got["result"]["final_cost"]) <= 1e-3 * max(1.0, abs(final))
    lib.close()


def test_anneal_without_term_matrix(ctx, orc):
    # libraries above the term-matrix cap compute every term from the Labs
    rng = _rng(31)
    n, gw, gh = 17000, 30, 20
    lab = (rng.random((n, 3)) * [100, 60, 60] - [0, 30, 30]).astype(np.float32)
    g = rng.integers(0, n, gw * gh).astype(np.int32)
    lib = ctx.library_from_lab(lab)
    got = ctx.anneal(lib, g, gw, gh, 1500, 100.0, 0.99995, 4, False, trace=True)
    want = orc.anneal(lab, g, gw, gh, 1500, 100.0, 0.99995, 4, False, trace=True)
    _check_anneal(got, want)
    pairs = np.stack([rng.integers(0, gw, 300), rng.integers(0, gh, 300),
                      rng.integers(0, gw, 300), rng.integers(0, gh, 300)], 1).astype(np.int32)
    assert_bits_equal(ctx.swap_deltas(lib, g, gw, gh, pairs),
                      [orc.swap_delta(lab, g, gw, gh, *map(int, p)) for p in pairs])
    assert _bits(np.float32(ctx.total_cost(lib, g, gw, gh))) == _bits(orc.total_cost(lab, g, gw, gh))
    lib.close()


def test_anneal_sparse_grid(ctx, orc):
    # src/optimizer.rs:336-357: non-adjacent tiles have zero cost
    lab = np.array([[50, 0, 0], [60, 10, 10]], np.float32)
    g = np.full(9, -1, np.int32)
    g[0], g[8] = 0, 1
    lib = ctx.library_from_lab(lab)
    got = ctx.anneal(lib, g, 3, 3, 50, seed=3, trace=True)
    want = orc.anneal(lab, g, 3, 3, 50, seed=3, trace=True)
    _check_anneal(got, want)
    assert got["result"]["initial_cost"] == 0.0
    lib.close()


# ------------------------------------------------ a13-a16 adjust/composite
def test_average_rgb_known_answer(ctx):
    # src/main.rs:877-889
    img = np.array([[[255, 0, 0], [0, 255, 0]], [[0, 0, 255], [255, 255, 255]]], np.uint8)
    np.testing.assert_array_equal(ctx.average_rgb(img), [127, 127, 127])


def test_optimal_adjustments(ctx, orc):
    rng = _rng(21)
    t = rng.integers(0, 256, (4000, 3)).astype(np.uint8)
    g = rng.integers(0, 256, (4000, 3)).astype(np.uint8)
    t[:256] = np.arange(256)[:, None]  # greys (saturation 0 branches)
    for k in [0.0, 0.4, 1.0]:
        got = ctx.optimal_adjustments(t, g, k)
        want = np.stack([orc.optimal_adjustment(t[i], g[i], k) for i in range(t.shape[0])])
        assert_bits_equal(got, want)


def test_adjust_pixels(ctx, orc):
    rng = _rng(22)
    levels = np.arange(256, dtype=np.uint8)
    px = np.concatenate([rng.integers(0, 256, (20000, 3)).astype(np.uint8),
                         np.stack([levels, levels, levels], 1),  # every u8 value per channel
                         np.stack([levels, levels[::-1], np.roll(levels, 85)], 1)])
    adjs = [orc.adjustment_new(0.0, 1.0, 0.0, 1.0), orc.adjustment_new(0.2, 1.0, 0.0, 1.0),
            orc.adjustment_new(0.0, 1.5, 0.0, 1.0), orc.adjustment_new(0.0, 1.0, 120.0, 1.0),
            orc.adjustment_new(-0.3, 0.7, -47.5, 0.5), orc.adjustment_new(0.5, 1.5, 60.0, 1.5),
            orc.adjustment_new(0.05, 1.0, 179.9, 2.0)]
    # raw hue shifts beyond ColorAdjustment::new's clamp: hue + shift crosses
    # +-360 and +-720 (the exact and the fmodf branches of the reduction)
    adjs += [np.array([0.0, 1.0, hs, 1.0], np.float32)
             for hs in (-180.0, -360.0, 359.9, 400.0, 700.0, 725.5, -1000.25, 1e6)]
    for a in adjs:
        got = ctx.adjust_pixels(a, px)
        want = orc.apply_adjustment(a, px)
        assert np.max(np.abs(got.astype(int) - want.astype(int))) <= 1
        np.testing.assert_array_equal(got, want)


def test_adjust_pixels_u8_conversion_truncates(ctx):
    # The brightness/contrast-only path (no HSV) over all 256 levels against
    # a numpy float32 closed form of src/color_adjustment.rs:82-96, which ends
    # in Rust's truncating `as u8`; independent of the oracle.
    f = np.float32
    levels = np.arange(256, dtype=np.uint8)
    px = np.repeat(levels[:, None], 3, axis=1)
    for brightness, contrast in ((0.1, 1.0), (-0.139, 1.0), (0.0, 1.3), (0.05, 0.77),
                                 (0.3, 1.9), (-0.42, 0.5)):
        v = levels.astype(f) / f(255.0)
        c = np.clip((v - f(0.5)) * f(contrast) + f(0.5), f(0.0), f(1.0)).astype(f)
        c = np.clip(c + f(brightness), f(0.0), f(1.0)).astype(f)
        t = np.clip(c * f(255.0), f(0.0), f(255.0)).astype(f)
        want = np.trunc(t).astype(np.uint8)
        assert np.count_nonzero(np.rint(t) != np.trunc(t)) > 20  # separates the two
        got = ctx.adjust_pixels(np.array([brightness, contrast, 0.0, 1.0], np.float32), px)
        for ch in range(3):
            np.testing.assert_array_equal(got[:, ch], want, err_msg=f"b={brightness} c={contrast}")


@pytest.mark.parametrize("sw,sh,dw,dh", [(64, 64, 16, 16), (128, 128, 24, 18), (64, 64, 20, 20),
                                         (100, 100, 50, 50), (100, 100, 200, 100), (7, 5, 13, 3),
                                         (256, 256, 30, 20), (61, 47, 19, 23)])
def test_resize(ctx, orc, sw, sh, dw, dh):
    img = _rng(sw * dh).integers(0, 256, (sh, sw, 3)).astype(np.uint8)
    np.testing.assert_array_equal(ctx.resize(img, dw, dh), orc.resize(img, dw, dh))


def test_resize_solid_colour(ctx):
    # src/main.rs:892-908
    red = np.zeros((100, 100, 3), np.uint8)
    red[..., 0] = 255
    out = ctx.resize(red, 50, 50)
    assert out.shape == (50, 50, 3)
    assert out[25, 25, 0] > 200 and out[25, 25, 1] < 50 and out[25, 25, 2] < 50
    assert np.all(out == out[0, 0])


@pytest.mark.parametrize("strength", [0.0, 0.4, 1.0])
def test_composite(ctx, orc, synth, strength):
    p, target, mats = synth.make("small")
    lab = orc.stack_lab_means(mats)
    cl = orc.cell_lab_means(target, p["grid_w"], p["grid_h"])
    grid = orc.place(lab, cl, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"])["grid"]
    grid[5] = -1  # an empty cell stays black
    lib = ctx.library_from_pixels(mats)
    got = ctx.composite(lib, target, grid, p["grid_w"], p["grid_h"], strength)
    want = orc.composite(target, mats, grid, p["grid_w"], p["grid_h"], strength)
    assert np.max(np.abs(got.astype(int) - want.astype(int))) <= 1
    np.testing.assert_array_equal(got, want)
    lib.close()


# ------------------------------------------- material ingest, mixed sizes
_MIXED = [(64, 64), (3, 7), (100, 37), (64, 64), (1, 1), (5, 300), (3, 7), (48, 48), (64, 64),
          (200, 150)]


def _mixed_materials(seed, sizes=_MIXED):
    rng = _rng(seed)
    out = []
    for w, h in sizes:
        base = rng.integers(0, 256, 3)
        out.append(np.clip(base + rng.integers(-40, 41, (h, w, 3)), 0, 255).astype(np.uint8))
    return out


def test_average_lab_many(ctx, orc):
    # load_tiles' per-file calculate_average_lab (src/main.rs:263-301) for
    # images of several sizes in mixed order, from one launch
    mats = _mixed_materials(31)
    got = ctx.average_lab_many(mats)
    assert_bits_equal(got, np.stack([orc.average_lab(m) for m in mats]))
    assert ctx.average_lab_many([]).shape == (0, 3)


def test_library_from_images(ctx, orc):
    mats = _mixed_materials(32)
    lib = ctx.library_from_images(mats)
    assert lib.n == len(mats)
    assert_bits_equal(lib.lab(), np.stack([orc.average_lab(m) for m in mats]))
    lib.close()
    # one size: the same library as mh_library_from_pixels
    uni = _mixed_materials(33, [(30, 20)] * 6)
    a, b = ctx.library_from_images(uni), ctx.library_from_pixels(np.stack(uni))
    assert_bits_equal(a.lab(), b.lab())
    target = _rng(34).integers(0, 256, (3 * 10, 4 * 12, 3)).astype(np.uint8)
    grid = _rng(35).integers(0, 6, 12).astype(np.int32)
    np.testing.assert_array_equal(ctx.composite(a, target, grid, 4, 3, 0.5),
                                  ctx.composite(b, target, grid, 4, 3, 0.5))
    a.close()
    b.close()
    empty = ctx.library_from_images([])
    assert empty.n == 0
    empty.close()


@pytest.mark.parametrize("strength", [0.0, 0.4])
def test_composite_mixed_size_materials(ctx, orc, strength):
    # each material is resized from its own size (resize_image,
    # src/main.rs:710-738); the oracle composites the tiles it resized one by
    # one, which its identity resize leaves unchanged
    mats = _mixed_materials(36)
    gw, gh, tw, th = 10, 8, 16, 15
    target = _rng(37).integers(0, 256, (gh * th, gw * tw, 3)).astype(np.uint8)
    grid = _rng(38).integers(0, len(mats), gw * gh).astype(np.int32)
    grid[3] = -1
    lib = ctx.library_from_images(mats)
    got = ctx.composite(lib, target, grid, gw, gh, strength)
    tiles = np.stack([orc.resize(m, tw, th) for m in mats])
    np.testing.assert_array_equal(got, orc.composite(target, tiles, grid, gw, gh, strength))
    lib.close()


def test_generate_mixed_size_materials(ctx, orc, synth):
    mats = _mixed_materials(39, _MIXED * 4)
    gw, gh, tw, th = 12, 9, 16, 15
    target = synth.blob_target(gw * tw, gh * th, 6, _rng(40))
    lib = ctx.library_from_images(mats)
    got = ctx.generate(lib, target, gw, gh, 2, 0.3, 0.4, True, 300, seed=3)
    lab = np.stack([orc.average_lab(m) for m in mats])
    grid = orc.place(lab, orc.cell_lab_means(target, gw, gh), gw, gh, 2, 0.3)["grid"]
    grid = orc.anneal(lab, grid, gw, gh, 300, 100.0, 0.99995, seed=3)["grid"]
    tiles = np.stack([orc.resize(m, tw, th) for m in mats])
    np.testing.assert_array_equal(got["grid"], grid)
    np.testing.assert_array_equal(got["image"], orc.composite(target, tiles, grid, gw, gh, 0.4))
    lib.close()


# ------------------------------------------------------------ whole job
@pytest.mark.parametrize("name,score_mode", [("small", 0), ("medium", 0), ("small", 1),
                                             ("medium", 1)])
def test_generate_matches_oracle(ctx, orc, synth, name, score_mode):
    p, target, mats = synth.make(name)
    lib = ctx.library_from_pixels(mats)
    got = ctx.generate(lib, target, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                       p["color"], True, p["iters"], seed=0, score_mode=score_mode)
    want = orc.generate(target, mats, p["grid_w"], p["grid_h"], p["max_usage"], p["weight"],
                        p["color"], True, p["iters"], seed=0, score_mode=score_mode)
    np.testing.assert_array_equal(got["grid"], want["grid"])
    assert got["place"] == want["place"]["stats"]
    for key in ["initial_cost", "final_cost", "best_cost"]:
        assert _bits(np.float32(got["anneal"][key])) == _bits(np.float32(want["anneal"]["result"][key]))
    assert np.max(np.abs(got["image"].astype(int) - want["image"].astype(int))) <= 1
    lib.close()
