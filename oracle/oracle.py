"""numpy wrappers over the C oracle (oracle/_build/libmosaic_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmosaic_oracle.so")
EMPTY = -1

_lib = None


def build():
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


class PlaceStats(ctypes.Structure):
    _fields_ = [("placed", ctypes.c_int32), ("main_path", ctypes.c_int32),
                ("fallback_ban", ctypes.c_int32), ("fallback_near", ctypes.c_int32),
                ("resets", ctypes.c_int32), ("empty", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AnnealResult(ctypes.Structure):
    _fields_ = [("initial_cost", ctypes.c_float), ("final_cost", ctypes.c_float),
                ("best_cost", ctypes.c_float), ("improved_count", ctypes.c_int64),
                ("accepted_count", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("evaluated", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, f32, u64, u8 = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_uint64,
                              ctypes.c_uint8)
    sig = {
        "orc_srgb_lut": (None, [P]),
        "orc_cbrtf": (f32, [f32]),
        "orc_expf": (f32, [f32]),
        "orc_rgb8_to_lab": (None, [u8, u8, u8, P]),
        "orc_average_lab": (None, [P, i32, i32, i64, P]),
        "orc_cell_lab_means": (None, [P, i32, i32, i32, i32, P]),
        "orc_stack_lab_means": (None, [P, i32, i32, i32, P]),
        "orc_lab_distance": (f32, [P, P]),
        "orc_sim": (f32, [P, i32, i32]),
        "orc_sim_full": (None, [P, i32, P]),
        "orc_sim_packed": (None, [P, i32, P]),
        "orc_nearest_n": (None, [P, i32, P, i32, P, P]),
        "orc_nearest_n_kd": (None, [P, i32, P, i32, i32, P, P]),
        "orc_sim_rows": (None, [P, i32, P, i32, P]),
        "orc_composite_tiles": (None, [P, i32, i32, P, P, i32, i32, f32, P]),
        "orc_place": (None, [P, i32, P, i32, i32, i32, f32, i32, P, i32, P, P, P, P, P,
                             ctypes.POINTER(PlaceStats)]),
        "orc_total_cost": (f32, [P, i32, P, i32, i32]),
        "orc_swap_delta": (f32, [P, i32, P, i32, i32, i32, i32, i32, i32]),
        "orc_draw": (ctypes.c_uint32, [u64, u64, ctypes.c_uint32]),
        "orc_anneal": (None, [P, i32, P, i32, i32, i64, f32, f32, u64, i32,
                              ctypes.POINTER(AnnealResult), P, P]),
        "orc_average_rgb": (None, [P, i32, i32, i64, P]),
        "orc_rgb_to_hsv": (None, [f32, f32, f32, P]),
        "orc_hsv_to_rgb": (None, [f32, f32, f32, P]),
        "orc_adjustment_new": (None, [f32, f32, f32, f32, P]),
        "orc_optimal_adjustment": (None, [P, P, f32, P]),
        "orc_adjust_pixel": (None, [P, P, P]),
        "orc_apply_adjustment": (None, [P, P, i64, P]),
        "orc_resize_lanczos3": (None, [P, i32, i32, P, i32, i32]),
        "orc_composite": (None, [P, i32, i32, P, i32, i32, i32, P, i32, i32, f32, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def srgb_lut():
    out = np.zeros(256, np.float32)
    load().orc_srgb_lut(_p(out))
    return out


def cbrtf(x):
    return np.float32(load().orc_cbrtf(float(x)))


def expf(x):
    return np.float32(load().orc_expf(float(x)))


def rgb8_to_lab(r, g, b):
    out = np.zeros(3, np.float32)
    load().orc_rgb8_to_lab(int(r), int(g), int(b), _p(out))
    return out


def average_lab(rgb):
    rgb = _c(rgb, np.uint8)
    h, w = rgb.shape[:2]
    out = np.zeros(3, np.float32)
    load().orc_average_lab(_p(rgb), w, h, w * 3, _p(out))
    return out


def cell_lab_means(rgb, grid_w, grid_h):
    rgb = _c(rgb, np.uint8)
    h, w = rgb.shape[:2]
    out = np.zeros((grid_w * grid_h, 3), np.float32)
    load().orc_cell_lab_means(_p(rgb), w, h, grid_w, grid_h, _p(out))
    return out


def stack_lab_means(rgb):
    rgb = _c(rgb, np.uint8)
    n, h, w = rgb.shape[:3]
    out = np.zeros((n, 3), np.float32)
    load().orc_stack_lab_means(_p(rgb), n, w, h, _p(out))
    return out


def lab_distance(a, b):
    a = _c(a, np.float32)
    b = _c(b, np.float32)
    return np.float32(load().orc_lab_distance(_p(a), _p(b)))


def sim_full(lab):
    lab = _c(lab, np.float32).reshape(-1, 3)
    n = lab.shape[0]
    out = np.zeros((n, n), np.float32)
    load().orc_sim_full(_p(lab), n, _p(out))
    return out


def sim_packed(lab):
    lab = _c(lab, np.float32).reshape(-1, 3)
    n = lab.shape[0]
    out = np.zeros(n * (n - 1) // 2, np.float32)
    load().orc_sim_packed(_p(lab), n, _p(out))
    return out


def nearest_n(lab, query, k):
    lab = _c(lab, np.float32).reshape(-1, 3)
    q = _c(query, np.float32).reshape(-1, 3)
    n = lab.shape[0]
    kk = min(k, n)
    idx = np.zeros((q.shape[0], kk), np.int32)
    d = np.zeros((q.shape[0], kk), np.float32)
    L = load()
    for i in range(q.shape[0]):
        L.orc_nearest_n(_p(lab), n, _p(q[i]), kk, _p(idx[i]), _p(d[i]))
    return idx, d


def nearest_n_kd(lab, query, k):
    """nearest_n from the oracle's k-d tree (the reference's structure)."""
    lab = _c(lab, np.float32).reshape(-1, 3)
    q = _c(query, np.float32).reshape(-1, 3)
    n = lab.shape[0]
    kk = min(k, n)
    idx = np.zeros((q.shape[0], kk), np.int32)
    d = np.zeros((q.shape[0], kk), np.float32)
    load().orc_nearest_n_kd(_p(lab), n, _p(q), q.shape[0], k, _p(idx), _p(d))
    return idx, d


def sim_rows(lab, rows):
    """Rows `rows` of the full similarity matrix (len(rows) x n)."""
    lab = _c(lab, np.float32).reshape(-1, 3)
    r = _c(rows, np.int32)
    out = np.zeros((r.size, lab.shape[0]), np.float32)
    load().orc_sim_rows(_p(lab), lab.shape[0], _p(r), r.size, _p(out))
    return out


def place(lab, cell_lab, grid_w, grid_h, max_usage, weight, order=None, grid=None, usage=None,
          score_mode=0):
    """Placement chain; score_mode 0 = CLI (sqdist + penalty, src/main.rs:385-398),
    1 = GUI (sqrt(sqdist) + penalty, src/gui/app_full.rs:1436-1448)."""
    lab = _c(lab, np.float32).reshape(-1, 3)
    n = lab.shape[0]
    cells = grid_w * grid_h
    cl = _c(cell_lab, np.float32).reshape(cells, 3)
    g = np.full(cells, EMPTY, np.int32) if grid is None else _c(grid, np.int32).copy()
    u = np.zeros(max(n, 1), np.uint32) if usage is None else _c(usage, np.uint32).copy()
    o = None if order is None else _c(order, np.int32)
    stage = np.full(cells, 255, np.uint8)
    dist = np.full(cells, np.nan, np.float32)
    score = np.full(cells, np.nan, np.float32)
    st = PlaceStats()
    load().orc_place(_p(lab), n, _p(cl), grid_w, grid_h, max_usage, weight, score_mode, _p(o),
                     0 if o is None else len(o), _p(g), _p(u), _p(stage), _p(dist), _p(score),
                     ctypes.byref(st))
    return {"grid": g, "usage": u[:n], "stage": stage, "dist": dist, "score": score,
            "stats": st.as_dict()}


def total_cost(lab, grid, grid_w, grid_h):
    lab = _c(lab, np.float32).reshape(-1, 3)
    g = _c(grid, np.int32)
    return np.float32(load().orc_total_cost(_p(lab), lab.shape[0], _p(g), grid_w, grid_h))


def swap_delta(lab, grid, grid_w, grid_h, x1, y1, x2, y2):
    lab = _c(lab, np.float32).reshape(-1, 3)
    g = _c(grid, np.int32)
    return np.float32(load().orc_swap_delta(_p(lab), lab.shape[0], _p(g), grid_w, grid_h,
                                            x1, y1, x2, y2))


def draw(seed, it, k):
    return load().orc_draw(seed, it, k)


def anneal(lab, grid, grid_w, grid_h, max_iterations, initial_temperature=100.0,
           temperature_decay=0.99995, seed=0, greedy=False, trace=False):
    lab = _c(lab, np.float32).reshape(-1, 3)
    g = _c(grid, np.int32).copy()
    res = AnnealResult()
    td = np.zeros(max(max_iterations, 1), np.float32) if trace else None
    ta = np.zeros(max(max_iterations, 1), np.uint8) if trace else None
    load().orc_anneal(_p(lab), lab.shape[0], _p(g), grid_w, grid_h, max_iterations,
                      initial_temperature, temperature_decay, seed, 1 if greedy else 0,
                      ctypes.byref(res), _p(td), _p(ta))
    out = {"grid": g, "result": res.as_dict()}
    if trace:
        out["trace_delta"] = td[:max_iterations]
        out["trace_accept"] = ta[:max_iterations]
    return out


def average_rgb(rgb):
    rgb = _c(rgb, np.uint8)
    h, w = rgb.shape[:2]
    out = np.zeros(3, np.uint8)
    load().orc_average_rgb(_p(rgb), w, h, w * 3, _p(out))
    return out


def rgb_to_hsv(r, g, b):
    out = np.zeros(3, np.float32)
    load().orc_rgb_to_hsv(r, g, b, _p(out))
    return out


def hsv_to_rgb(h, s, v):
    out = np.zeros(3, np.float32)
    load().orc_hsv_to_rgb(h, s, v, _p(out))
    return out


def adjustment_new(brightness, contrast, hue_shift, saturation):
    out = np.zeros(4, np.float32)
    load().orc_adjustment_new(brightness, contrast, hue_shift, saturation, _p(out))
    return out


def optimal_adjustment(tile_rgb, target_rgb, strength):
    t = _c(tile_rgb, np.uint8)
    g = _c(target_rgb, np.uint8)
    out = np.zeros(4, np.float32)
    load().orc_optimal_adjustment(_p(t), _p(g), strength, _p(out))
    return out


def adjust_pixel(adj, rgb):
    a = _c(adj, np.float32)
    p = _c(rgb, np.uint8)
    out = np.zeros(3, np.uint8)
    load().orc_adjust_pixel(_p(a), _p(p), _p(out))
    return out


def apply_adjustment(adj, rgb):
    a = _c(adj, np.float32)
    p = _c(rgb, np.uint8)
    out = np.zeros_like(p)
    load().orc_apply_adjustment(_p(a), _p(p), p.size // 3, _p(out))
    return out


def resize(rgb, dw, dh):
    s = _c(rgb, np.uint8)
    sh, sw = s.shape[:2]
    out = np.zeros((dh, dw, 3), np.uint8)
    load().orc_resize_lanczos3(_p(s), sw, sh, _p(out), dw, dh)
    return out


def composite(target, materials, grid, grid_w, grid_h, strength):
    t = _c(target, np.uint8)
    m = _c(materials, np.uint8)
    h, w = t.shape[:2]
    n, mh, mw = m.shape[:3]
    tw, th = w // grid_w, h // grid_h
    out = np.zeros((grid_h * th, grid_w * tw, 3), np.uint8)
    g = _c(grid, np.int32)
    load().orc_composite(_p(t), w, h, _p(m), n, mw, mh, _p(g), grid_w, grid_h, strength, _p(out))
    return out


def generate(target, materials, grid_w, grid_h, max_usage, weight, color_strength,
             enable_optimization=True, optimization_iterations=1000, initial_temperature=100.0,
             temperature_decay=0.99995, seed=0, image=True, score_mode=0):
    """generate_mosaic (src/main.rs:477-685) without file I/O; score_mode 1 places
    with the GUI generator's scoring (src/gui/app_full.rs:1396-1563)."""
    lab = stack_lab_means(materials)
    cell_lab = cell_lab_means(target, grid_w, grid_h)
    pl = place(lab, cell_lab, grid_w, grid_h, max_usage, weight, score_mode=score_mode)
    grid = pl["grid"]
    ann = None
    if enable_optimization and weight > 0.0:
        ann = anneal(lab, grid, grid_w, grid_h, optimization_iterations, initial_temperature,
                     temperature_decay, seed)
        grid = ann["grid"]
    strength = min(max(color_strength, 0.0), 1.0)
    img = composite(target, materials, grid, grid_w, grid_h, strength) if image else None
    return {"grid": grid, "image": img, "place": pl, "anneal": ann, "lab": lab,
            "cell_lab": cell_lab}


# ----------------------------------------------------------- prepared jobs
def _pool(fn, items, threads):
    """fn over items on `threads` threads (the C calls release the GIL)."""
    import concurrent.futures
    if threads <= 1 or len(items) <= 1:
        return [fn(x) for x in items]
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(fn, items))


def stack_lab_means_par(materials, threads=8):
    """stack_lab_means over chunks of materials on several threads (the
    same values: each material is computed on its own)."""
    m = _c(materials, np.uint8)
    n = m.shape[0]
    step = max(1, (n + threads - 1) // threads)
    parts = _pool(lambda s: stack_lab_means(m[s:s + step]), list(range(0, n, step)), threads)
    return np.concatenate(parts) if parts else np.zeros((0, 3), np.float32)


def cell_lab_means_par(target, grid_w, grid_h, threads=8):
    """cell_lab_means over bands of cell rows on several threads (each band is
    the same cells of the same pixels)."""
    t = _c(target, np.uint8)
    th = t.shape[0] // grid_h
    step = max(1, (grid_h + threads - 1) // threads)
    bands = list(range(0, grid_h, step))
    parts = _pool(lambda y0: cell_lab_means(t[y0 * th:min(grid_h, y0 + step) * th], grid_w,
                                            min(grid_h, y0 + step) - y0), bands, threads)
    return np.concatenate(parts)


def resize_tiles(materials, tw, th, threads=8):
    """Every material resized to tw x th (n x th x tw x 3), once per cell
    size, as a resident library caches them."""
    m = _c(materials, np.uint8)
    n = m.shape[0]
    out = np.zeros((n, th, tw, 3), np.uint8)
    step = max(1, (n + threads - 1) // threads)

    def part(s):
        L = load()
        mh, mw = m.shape[1:3]
        for i in range(s, min(n, s + step)):
            L.orc_resize_lanczos3(_p(m[i]), mw, mh, _p(out[i]), tw, th)

    _pool(part, list(range(0, n, step)), threads)
    return out


def composite_tiles(target, tiles, grid, grid_w, grid_h, strength):
    """composite from prepared tiles (n x th x tw x 3)."""
    t = _c(target, np.uint8)
    tl = _c(tiles, np.uint8)
    h, w = t.shape[:2]
    tw, th = w // grid_w, h // grid_h
    out = np.zeros((grid_h * th, grid_w * tw, 3), np.uint8)
    g = _c(grid, np.int32)
    load().orc_composite_tiles(_p(t), w, h, _p(tl), _p(g), grid_w, grid_h, strength, _p(out))
    return out


def prepare_library(materials, cell_sizes=(), threads=8):
    """A library prepared once, as the GPU's resident library is: material
    Lab means (calculate_average_lab of each, src/main.rs:263-301) and the
    materials resized to each cell size (tw, th)."""
    lab = stack_lab_means_par(materials, threads)
    return {"lab": lab, "tiles": {tuple(c): resize_tiles(materials, c[0], c[1], threads)
                                   for c in cell_sizes}}


def generate_prepared(target, lib, grid_w, grid_h, max_usage, weight, color_strength,
                      enable_optimization=True, optimization_iterations=1000,
                      initial_temperature=100.0, temperature_decay=0.99995, seed=0, image=True,
                      score_mode=0, times=None, threads=1):
    """generate (src/main.rs:529-665) on a prepared library: the per-job work
    only.  `times`, when a dict, accumulates seconds per phase: cell_lab,
    place (K nearest + chain), anneal, composite."""
    import time
    h, w = target.shape[:2]
    tw, th = w // grid_w, h // grid_h
    t0 = time.perf_counter()
    cell_lab = (cell_lab_means_par(target, grid_w, grid_h, threads) if threads > 1
                else cell_lab_means(target, grid_w, grid_h))
    t1 = time.perf_counter()
    pl = place(lib["lab"], cell_lab, grid_w, grid_h, max_usage, weight, score_mode=score_mode)
    grid = pl["grid"]
    t2 = time.perf_counter()
    ann = None
    if enable_optimization and weight > 0.0:
        ann = anneal(lib["lab"], grid, grid_w, grid_h, optimization_iterations,
                     initial_temperature, temperature_decay, seed)
        grid = ann["grid"]
    t3 = time.perf_counter()
    strength = min(max(color_strength, 0.0), 1.0)
    img = None
    if image:
        tiles = lib["tiles"].get((tw, th))
        if tiles is None:
            raise KeyError(f"library not prepared for {tw}x{th} cells")
        img = composite_tiles(target, tiles, grid, grid_w, grid_h, strength)
    t4 = time.perf_counter()
    if times is not None:
        for k, v in (("cell_lab", t1 - t0), ("place", t2 - t1), ("anneal", t3 - t2),
                     ("composite", t4 - t3)):
            times[k] = times.get(k, 0.0) + v
    return {"grid": grid, "image": img, "place": pl, "anneal": ann, "lab": lib["lab"],
            "cell_lab": cell_lab}
