"""ctypes binding of libmosaic_hip.so (include/mosaic_hip.h) with numpy I/O.

This is plumbing for tests, the benchmark and Python callers; the product is
the C ABI and its gfx950 kernels.  Loading fails loudly when the library is
missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmosaic_hip.so")

MH_OK = 0
MH_ERR_INVALID_ARGUMENT = -1
MH_ERR_NO_DEVICE = -2
MH_ERR_HIP = -3
MH_ERR_OUT_OF_MEMORY = -4
MH_ERR_NOT_FOUND = -5
MH_ERR_UNSUPPORTED = -6
MH_EMPTY = -1
# Placement score: CLI squared distance + penalty, or the GUI generator's
# Euclidean distance + penalty (src/gui/app_full.rs:1436-1448).
SCORE_SQDIST = 0
SCORE_DISTANCE = 1
MH_PHASES = 8

# Every symbol include/mosaic_hip.h declares.
EXPORTS = [
    "mh_last_error", "mh_abi_version", "mh_context_create", "mh_context_destroy",
    "mh_context_synchronize", "mh_context_stream", "mh_context_set_timing",
    "mh_context_set_timing_phases",
    "mh_context_timing", "mh_phase_name", "mh_device_alloc", "mh_device_free",
    "mh_copy_to_device", "mh_copy_to_host", "mh_copy_on_device", "mh_copy_2d_on_device",
    "mh_library_from_lab",
    "mh_library_from_pixels", "mh_average_lab_many", "mh_library_from_images",
    "mh_library_destroy", "mh_library_size", "mh_library_lab",
    "mh_library_build_similarities", "mh_library_similarities", "mh_library_get_similarity",
    "mh_average_lab", "mh_cell_lab_means", "mh_nearest_n", "mh_place", "mh_total_cost",
    "mh_penalties", "mh_swap_deltas", "mh_anneal", "mh_average_rgb", "mh_optimal_adjustments",
    "mh_adjust_pixels", "mh_resize", "mh_composite", "mh_generate", "mh_generate_device",
    "mh_generate_device_result", "mh_generate_device_noalloc", "mh_job_footprint", "mh_batch_create", "mh_batch_set_graph",
    "mh_batch_run", "mh_batch_results", "mh_batch_job_outputs", "mh_batch_destroy",
    "mh_batch_tune", "mh_batch_get_tuning", "mh_batch_set_tuning", "mh_server_create", "mh_server_submit", "mh_server_wait", "mh_server_poll", "mh_server_drain",
    "mh_server_submit_many", "mh_server_job_times", "mh_server_get_stats",
    "mh_server_set_timing_phases",
    "mh_server_timing", "mh_server_destroy",
]


class MhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mosaic_hip error {code}: {msg}")
        self.code = code


class PlaceParams(ctypes.Structure):
    _fields_ = [("grid_w", ctypes.c_int32), ("grid_h", ctypes.c_int32),
                ("max_usage", ctypes.c_int32), ("weight", ctypes.c_float),
                ("score_mode", ctypes.c_int32)]


class PlaceStats(ctypes.Structure):
    _fields_ = [("placed", ctypes.c_int32), ("main_path", ctypes.c_int32),
                ("fallback_ban", ctypes.c_int32), ("fallback_near", ctypes.c_int32),
                ("resets", ctypes.c_int32), ("empty", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AnnealParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("initial_temperature", ctypes.c_float),
                ("temperature_decay", ctypes.c_float), ("seed", ctypes.c_uint64),
                ("greedy", ctypes.c_int32)]


class AnnealResult(ctypes.Structure):
    _fields_ = [("initial_cost", ctypes.c_float), ("final_cost", ctypes.c_float),
                ("best_cost", ctypes.c_float), ("improved_count", ctypes.c_int64),
                ("accepted_count", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("evaluated", ctypes.c_int64), ("rounds", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobParams(ctypes.Structure):
    _fields_ = [("grid_w", ctypes.c_int32), ("grid_h", ctypes.c_int32),
                ("max_usage", ctypes.c_int32), ("weight", ctypes.c_float),
                ("color_strength", ctypes.c_float), ("enable_optimization", ctypes.c_int32),
                ("optimization_iterations", ctypes.c_int64),
                ("initial_temperature", ctypes.c_float), ("temperature_decay", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("score_mode", ctypes.c_int32)]


class JobResult(ctypes.Structure):
    _fields_ = [("place", PlaceStats), ("anneal", AnnealResult),
                ("out_width", ctypes.c_int32), ("out_height", ctypes.c_int32)]


class Job(ctypes.Structure):
    _fields_ = [("lib", ctypes.c_void_p), ("d_target", ctypes.c_void_p),
                ("target_pitch", ctypes.c_int64), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("params", JobParams),
                ("d_grid_out", ctypes.c_void_p), ("d_image_out", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("n_groups", ctypes.c_int32), ("topk_cells_per_wave", ctypes.c_int32 * 2),
                ("composite_threads", ctypes.c_int32 * 2), ("topk_ms", ctypes.c_float * 2),
                ("composite_ms", ctypes.c_float * 2)]

    def as_dict(self):
        g = self.n_groups
        return {"groups": g,
                "topk_cells_per_wave": list(self.topk_cells_per_wave)[:g],
                "composite_threads": list(self.composite_threads)[:g],
                "topk_ms": [round(v, 5) for v in list(self.topk_ms)[:g]],
                "composite_ms": [round(v, 5) for v in list(self.composite_ms)[:g]]}


class ServerParams(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_int64), ("max_batch", ctypes.c_int32),
                ("streams", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("max_batch_jobs", ctypes.c_int32),
                ("memory_in_use", ctypes.c_int64), ("memory_peak", ctypes.c_int64),
                ("memory_budget", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _job_result_dict(res):
    return {"place": res.place.as_dict(), "anneal": res.anneal.as_dict(),
            "out_size": (res.out_width, res.out_height)}


_lib = None


def load(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i32, i64, f32, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_uint64
    sig = {
        "mh_last_error": (ctypes.c_char_p, []),
        "mh_abi_version": (i32, []),
        "mh_context_create": (ctypes.c_int, [i32, ctypes.POINTER(P)]),
        "mh_context_destroy": (None, [P]),
        "mh_context_synchronize": (ctypes.c_int, [P]),
        "mh_context_stream": (P, [P]),
        "mh_context_set_timing": (ctypes.c_int, [P, i32]),
        "mh_context_set_timing_phases": (ctypes.c_int, [P, ctypes.c_uint32]),
        "mh_context_timing": (ctypes.c_int, [P, P, P]),
        "mh_phase_name": (ctypes.c_char_p, [i32]),
        "mh_device_alloc": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
        "mh_device_free": (ctypes.c_int, [P, P]),
        "mh_copy_to_device": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
        "mh_copy_to_host": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
        "mh_copy_on_device": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
        "mh_copy_2d_on_device": (ctypes.c_int, [P, P, ctypes.c_size_t, P, ctypes.c_size_t,
                                                ctypes.c_size_t, ctypes.c_size_t]),
        "mh_library_from_lab": (ctypes.c_int, [P, P, i32, ctypes.POINTER(P)]),
        "mh_library_from_pixels": (ctypes.c_int, [P, P, i32, i32, i32, ctypes.POINTER(P)]),
        "mh_average_lab_many": (ctypes.c_int, [P, P, P, P, P, i32, P]),
        "mh_library_from_images": (ctypes.c_int, [P, P, P, P, P, i32, ctypes.POINTER(P)]),
        "mh_library_destroy": (None, [P]),
        "mh_library_size": (i32, [P]),
        "mh_library_lab": (ctypes.c_int, [P, P]),
        "mh_library_build_similarities": (ctypes.c_int, [P]),
        "mh_library_similarities": (ctypes.c_int, [P, P, i32]),
        "mh_library_get_similarity": (ctypes.c_int, [P, i32, i32, ctypes.POINTER(f32)]),
        "mh_average_lab": (ctypes.c_int, [P, P, i32, i32, P]),
        "mh_cell_lab_means": (ctypes.c_int, [P, P, i32, i32, i32, i32, P]),
        "mh_nearest_n": (ctypes.c_int, [P, P, P, i32, i32, P, P]),
        "mh_place": (ctypes.c_int, [P, P, P, ctypes.POINTER(PlaceParams), P, i32, P, P, P, P, P,
                                    ctypes.POINTER(PlaceStats)]),
        "mh_total_cost": (ctypes.c_int, [P, P, P, i32, i32, ctypes.POINTER(f32)]),
        "mh_penalties": (ctypes.c_int, [P, P, P, i32, i32, f32, P, i32, P]),
        "mh_swap_deltas": (ctypes.c_int, [P, P, P, i32, i32, P, i32, P]),
        "mh_anneal": (ctypes.c_int, [P, P, P, i32, i32, ctypes.POINTER(AnnealParams),
                                     ctypes.POINTER(AnnealResult), P, P]),
        "mh_average_rgb": (ctypes.c_int, [P, P, i32, i32, P]),
        "mh_optimal_adjustments": (ctypes.c_int, [P, P, P, i32, f32, P]),
        "mh_adjust_pixels": (ctypes.c_int, [P, P, P, i64, P]),
        "mh_resize": (ctypes.c_int, [P, P, i32, i32, P, i32, i32]),
        "mh_composite": (ctypes.c_int, [P, P, P, i32, i32, P, i32, i32, f32, P]),
        "mh_generate": (ctypes.c_int, [P, P, P, i32, i32, ctypes.POINTER(JobParams), P, P,
                                       ctypes.POINTER(JobResult)]),
        "mh_generate_device": (ctypes.c_int, [P, P, P, i32, i32, ctypes.POINTER(JobParams), P, P]),
        "mh_generate_device_result": (ctypes.c_int, [P, ctypes.POINTER(JobResult)]),
        "mh_generate_device_noalloc": (ctypes.c_int, [P, P, P, i32, i32,
                                                      ctypes.POINTER(JobParams), P, P]),
        "mh_job_footprint": (i64, [P, ctypes.POINTER(Job)]),
        "mh_batch_create": (ctypes.c_int, [P, P, i32, ctypes.POINTER(P)]),
        "mh_batch_set_graph": (ctypes.c_int, [P, i32]),
        "mh_batch_run": (ctypes.c_int, [P]),
        "mh_batch_results": (ctypes.c_int, [P, P]),
        "mh_batch_job_outputs": (ctypes.c_int, [P, i32, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "mh_batch_destroy": (None, [P]),
        "mh_batch_tune": (ctypes.c_int, [P, i32, P]),
        "mh_batch_get_tuning": (ctypes.c_int, [P, P]),
        "mh_batch_set_tuning": (ctypes.c_int, [P, P]),
        "mh_server_create": (ctypes.c_int, [P, ctypes.POINTER(ServerParams), ctypes.POINTER(P)]),
        "mh_server_submit": (ctypes.c_int, [P, ctypes.POINTER(Job), ctypes.POINTER(i64)]),
        "mh_server_submit_many": (ctypes.c_int, [P, P, i32, P]),
        "mh_server_wait": (ctypes.c_int, [P, i64, ctypes.POINTER(JobResult),
                                          ctypes.POINTER(ctypes.c_double)]),
        "mh_server_poll": (ctypes.c_int, [P, i64, ctypes.POINTER(i32)]),
        "mh_server_drain": (ctypes.c_int, [P]),
        "mh_server_job_times": (ctypes.c_int, [P, i64, P]),
        "mh_server_get_stats": (ctypes.c_int, [P, ctypes.POINTER(ServerStats)]),
        "mh_server_set_timing_phases": (ctypes.c_int, [P, ctypes.c_uint32]),
        "mh_server_timing": (ctypes.c_int, [P, P, P, P]),
        "mh_server_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != MH_OK:
        raise MhError(rc, load().mh_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """One HIP device + stream (mh_context)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        _check(self.lib.mh_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.mh_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(self.lib.mh_context_synchronize(self.h))

    @property
    def stream(self):
        return self.lib.mh_context_stream(self.h)

    def set_timing(self, enabled):
        _check(self.lib.mh_context_set_timing(self.h, 1 if enabled else 0))

    def set_timing_phases(self, phases):
        """Times only the named phases (each adds two event records per launch)."""
        mask = 0
        for name in phases:
            mask |= 1 << [self.lib.mh_phase_name(i).decode() for i in range(MH_PHASES)].index(name)
        _check(self.lib.mh_context_set_timing_phases(self.h, mask))

    def timing(self):
        """{phase: (total_ms, launches)} since the last call (synchronises)."""
        ms = np.zeros(MH_PHASES, np.float64)
        n = np.zeros(MH_PHASES, np.int64)
        _check(self.lib.mh_context_timing(self.h, _ptr(ms), _ptr(n)))
        return {self.lib.mh_phase_name(i).decode(): (float(ms[i]), int(n[i]))
                for i in range(MH_PHASES)}

    # ------------------------------------------------------ device buffers
    def device_alloc(self, nbytes):
        p = ctypes.c_void_p()
        _check(self.lib.mh_device_alloc(self.h, nbytes, ctypes.byref(p)))
        return p.value

    def device_free(self, ptr):
        _check(self.lib.mh_device_free(self.h, ctypes.c_void_p(ptr)))

    def to_device(self, array):
        """Copies a numpy array into a new device buffer; returns its address."""
        a = np.ascontiguousarray(array)
        ptr = self.device_alloc(a.nbytes)
        _check(self.lib.mh_copy_to_device(self.h, ctypes.c_void_p(ptr), _ptr(a), a.nbytes))
        return ptr

    def copy_to_device(self, ptr, array):
        """Copies a numpy array into an existing device buffer."""
        a = np.ascontiguousarray(array)
        _check(self.lib.mh_copy_to_device(self.h, ctypes.c_void_p(ptr), _ptr(a), a.nbytes))

    def copy_on_device(self, dst, src, nbytes):
        _check(self.lib.mh_copy_on_device(self.h, ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes))

    def copy_2d_on_device(self, dst, dst_pitch, src, src_pitch, row_bytes, rows):
        _check(self.lib.mh_copy_2d_on_device(self.h, ctypes.c_void_p(dst), dst_pitch,
                                             ctypes.c_void_p(src), src_pitch, row_bytes, rows))

    def to_host(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        _check(self.lib.mh_copy_to_host(self.h, _ptr(out), ctypes.c_void_p(ptr), out.nbytes))
        return out

    # ---------------------------------------------------------------- a1
    def average_lab(self, rgb):
        rgb = _c(rgb, np.uint8)
        h, w = rgb.shape[:2]
        out = np.zeros(3, np.float32)
        _check(self.lib.mh_average_lab(self.h, _ptr(rgb), w, h, _ptr(out)))
        return out

    def cell_lab_means(self, rgb, grid_w, grid_h):
        rgb = _c(rgb, np.uint8)
        h, w = rgb.shape[:2]
        out = np.zeros((grid_w * grid_h, 3), np.float32)
        _check(self.lib.mh_cell_lab_means(self.h, _ptr(rgb), w, h, grid_w, grid_h, _ptr(out)))
        return out

    # ------------------------------------------------------------ library
    def library_from_lab(self, lab):
        return Library(self, lab=lab)

    def library_from_pixels(self, rgb):
        return Library(self, pixels=rgb)

    def library_from_images(self, images):
        """Library from RGB8 images of different sizes (list of h x w x 3)."""
        return Library(self, images=images)

    def average_lab_many(self, images):
        """calculate_average_lab of each RGB8 image (list of h x w x 3)."""
        buf, off, w, h = _pack_images(images)
        out = np.zeros((len(images), 3), np.float32)
        _check(self.lib.mh_average_lab_many(self.h, _ptr(buf), _ptr(off), _ptr(w), _ptr(h),
                                            len(images), _ptr(out)))
        return out

    # --------------------------------------------------------------- a3
    def nearest_n(self, lib, query, k):
        q = _c(query, np.float32).reshape(-1, 3)
        kk = min(k, lib.n)
        idx = np.zeros((q.shape[0], kk), np.int32)
        d = np.zeros((q.shape[0], kk), np.float32)
        _check(self.lib.mh_nearest_n(self.h, lib.h, _ptr(q), q.shape[0], k, _ptr(idx), _ptr(d)))
        return idx, d

    # ------------------------------------------------------------ a4-a9
    def place(self, lib, cell_lab, grid_w, grid_h, max_usage, weight, order=None, grid=None,
              usage=None, score_mode=SCORE_SQDIST):
        cells = grid_w * grid_h
        cl = _c(cell_lab, np.float32).reshape(cells, 3)
        g = np.full(cells, MH_EMPTY, np.int32) if grid is None else _c(grid, np.int32).copy()
        u = np.zeros(max(lib.n, 1), np.uint32) if usage is None else _c(usage, np.uint32).copy()
        o = None if order is None else _c(order, np.int32)
        stage = np.full(cells, 255, np.uint8)
        dist = np.full(cells, np.nan, np.float32)
        score = np.full(cells, np.nan, np.float32)
        st = PlaceStats()
        prm = PlaceParams(grid_w, grid_h, max_usage, weight, score_mode)
        _check(self.lib.mh_place(self.h, lib.h, _ptr(cl), ctypes.byref(prm), _ptr(o),
                                 0 if o is None else len(o), _ptr(g), _ptr(u), _ptr(stage),
                                 _ptr(dist), _ptr(score), ctypes.byref(st)))
        return {"grid": g, "usage": u[:lib.n], "stage": stage, "dist": dist, "score": score,
                "stats": st.as_dict()}

    # ----------------------------------------------------------- a10-a12
    def total_cost(self, lib, grid, grid_w, grid_h):
        g = _c(grid, np.int32)
        out = ctypes.c_float()
        _check(self.lib.mh_total_cost(self.h, lib.h, _ptr(g), grid_w, grid_h, ctypes.byref(out)))
        return out.value

    def penalties(self, lib, grid, grid_w, grid_h, weight, queries):
        g = _c(grid, np.int32)
        q = _c(queries, np.int32).reshape(-1, 3)
        out = np.zeros(q.shape[0], np.float32)
        _check(self.lib.mh_penalties(self.h, lib.h, _ptr(g), grid_w, grid_h, weight, _ptr(q),
                                     q.shape[0], _ptr(out)))
        return out

    def swap_deltas(self, lib, grid, grid_w, grid_h, pairs):
        g = _c(grid, np.int32)
        p = _c(pairs, np.int32).reshape(-1, 4)
        out = np.zeros(p.shape[0], np.float32)
        _check(self.lib.mh_swap_deltas(self.h, lib.h, _ptr(g), grid_w, grid_h, _ptr(p),
                                       p.shape[0], _ptr(out)))
        return out

    def anneal(self, lib, grid, grid_w, grid_h, max_iterations, initial_temperature=100.0,
               temperature_decay=0.99995, seed=0, greedy=False, trace=False):
        g = _c(grid, np.int32).copy()
        prm = AnnealParams(max_iterations, initial_temperature, temperature_decay, seed,
                           1 if greedy else 0)
        res = AnnealResult()
        td = np.zeros(max(max_iterations, 1), np.float32) if trace else None
        ta = np.zeros(max(max_iterations, 1), np.uint8) if trace else None
        _check(self.lib.mh_anneal(self.h, lib.h, _ptr(g), grid_w, grid_h, ctypes.byref(prm),
                                  ctypes.byref(res), _ptr(td), _ptr(ta)))
        out = {"grid": g, "result": res.as_dict()}
        if trace:
            out["trace_delta"] = td[:max_iterations]
            out["trace_accept"] = ta[:max_iterations]
        return out

    # ----------------------------------------------------------- a13-a16
    def average_rgb(self, rgb):
        rgb = _c(rgb, np.uint8)
        h, w = rgb.shape[:2]
        out = np.zeros(3, np.uint8)
        _check(self.lib.mh_average_rgb(self.h, _ptr(rgb), w, h, _ptr(out)))
        return out

    def optimal_adjustments(self, tile_rgb, target_rgb, strength):
        t = _c(tile_rgb, np.uint8).reshape(-1, 3)
        g = _c(target_rgb, np.uint8).reshape(-1, 3)
        out = np.zeros((t.shape[0], 4), np.float32)
        _check(self.lib.mh_optimal_adjustments(self.h, _ptr(t), _ptr(g), t.shape[0], strength,
                                               _ptr(out)))
        return out

    def adjust_pixels(self, adj, rgb):
        a = _c(adj, np.float32)
        p = _c(rgb, np.uint8)
        out = np.zeros_like(p)
        _check(self.lib.mh_adjust_pixels(self.h, _ptr(a), _ptr(p), p.size // 3, _ptr(out)))
        return out

    def resize(self, rgb, dw, dh):
        s = _c(rgb, np.uint8)
        sh, sw = s.shape[:2]
        out = np.zeros((dh, dw, 3), np.uint8)
        _check(self.lib.mh_resize(self.h, _ptr(s), sw, sh, _ptr(out), dw, dh))
        return out

    def composite(self, lib, target, grid, grid_w, grid_h, strength):
        t = _c(target, np.uint8)
        h, w = t.shape[:2]
        tw, th = w // grid_w, h // grid_h
        out = np.zeros((grid_h * th, grid_w * tw, 3), np.uint8)
        g = _c(grid, np.int32)
        _check(self.lib.mh_composite(self.h, lib.h, _ptr(t), w, h, _ptr(g), grid_w, grid_h,
                                     strength, _ptr(out)))
        return out

    # ---------------------------------------------------------- whole job
    def generate(self, lib, target, grid_w, grid_h, max_usage, weight, color_strength,
                 enable_optimization=True, optimization_iterations=1000,
                 initial_temperature=100.0, temperature_decay=0.99995, seed=0, image=True,
                 score_mode=SCORE_SQDIST):
        t = _c(target, np.uint8)
        h, w = t.shape[:2]
        tw, th = w // grid_w, h // grid_h
        grid = np.zeros(grid_w * grid_h, np.int32)
        img = np.zeros((grid_h * th, grid_w * tw, 3), np.uint8) if image else None
        prm = JobParams(grid_w, grid_h, max_usage, weight, color_strength,
                        1 if enable_optimization else 0, optimization_iterations,
                        initial_temperature, temperature_decay, seed, score_mode)
        res = JobResult()
        _check(self.lib.mh_generate(self.h, lib.h, _ptr(t), w, h, ctypes.byref(prm), _ptr(grid),
                                    _ptr(img), ctypes.byref(res)))
        return {"grid": grid, "image": img, "place": res.place.as_dict(),
                "anneal": res.anneal.as_dict(), "out_size": (res.out_width, res.out_height)}

    def generate_device(self, lib, d_target, width, height, params, d_grid, d_image):
        """Device-pointer form; d_* are integer device addresses (e.g. torch
        tensor.data_ptr()).  Enqueues on the context stream, no sync."""
        _check(self.lib.mh_generate_device(self.h, lib.h, ctypes.c_void_p(d_target), width, height,
                                           ctypes.byref(params), ctypes.c_void_p(d_grid),
                                           ctypes.c_void_p(d_image) if d_image else None))

    def generate_device_noalloc(self, lib, d_target, width, height, params, d_grid, d_image):
        """mh_generate_device_noalloc: generate_device with its working memory
        in the library's static per-GPU scratch array (no device allocation)."""
        _check(self.lib.mh_generate_device_noalloc(
            self.h, lib.h, ctypes.c_void_p(d_target), width, height, ctypes.byref(params),
            ctypes.c_void_p(d_grid) if d_grid else None,
            ctypes.c_void_p(d_image) if d_image else None))

    def generate_device_result(self):
        res = JobResult()
        _check(self.lib.mh_generate_device_result(self.h, ctypes.byref(res)))
        return _job_result_dict(res)

    # ------------------------------------------------- batches and serving
    def job(self, lib, d_target, width, height, params, d_grid=None, d_image=None, pitch=0):
        """An mh_job over device buffers (integer addresses)."""
        return Job(lib.h.value, d_target, pitch, width, height, params, d_grid, d_image)

    def job_footprint(self, job):
        v = self.lib.mh_job_footprint(self.h, ctypes.byref(job))
        if v < 0:
            _check(int(v))
        return int(v)

    def batch(self, jobs):
        return Batch(self, jobs)

    def server(self, memory_budget=0, max_batch=0, streams=0):
        return Server(self, memory_budget, max_batch, streams)


def _pack_images(images):
    """Concatenates RGB8 images into one buffer: (bytes, offsets, widths, heights)."""
    arrs = [_c(im, np.uint8) for im in images]
    off = np.zeros(len(arrs), np.int64)
    w = np.array([a.shape[1] for a in arrs], np.int32)
    h = np.array([a.shape[0] for a in arrs], np.int32)
    sizes = np.array([a.size for a in arrs], np.int64)
    if len(arrs):
        off[1:] = np.cumsum(sizes)[:-1]
    buf = np.concatenate([a.reshape(-1) for a in arrs]) if arrs else np.zeros(1, np.uint8)
    return buf, off, w, h


class Library:
    """Device-resident material library (SimilarityDatabase + Lab table)."""

    def __init__(self, ctx, lab=None, pixels=None, images=None):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        if images is not None:
            buf, off, w, hh = _pack_images(images)
            _check(self.lib.mh_library_from_images(ctx.h, _ptr(buf), _ptr(off), _ptr(w), _ptr(hh),
                                                   len(images), ctypes.byref(h)))
        elif pixels is not None:
            px = _c(pixels, np.uint8)
            n, mh, mw = px.shape[:3]
            _check(self.lib.mh_library_from_pixels(ctx.h, _ptr(px), n, mw, mh, ctypes.byref(h)))
        else:
            lab = _c(lab, np.float32).reshape(-1, 3)
            _check(self.lib.mh_library_from_lab(ctx.h, _ptr(lab), lab.shape[0], ctypes.byref(h)))
        self.h = h
        self.n = self.lib.mh_library_size(h)

    def close(self):
        if self.h:
            self.lib.mh_library_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lab(self):
        out = np.zeros((self.n, 3), np.float32)
        _check(self.lib.mh_library_lab(self.h, _ptr(out)))
        return out

    def build_similarities(self):
        _check(self.lib.mh_library_build_similarities(self.h))

    def similarities(self, packed=False):
        n = self.n
        out = np.zeros(n * (n - 1) // 2 if packed else n * n, np.float32)
        _check(self.lib.mh_library_similarities(self.h, _ptr(out), 1 if packed else 0))
        return out if packed else out.reshape(n, n)

    def get_similarity(self, i, j):
        v = ctypes.c_float()
        rc = self.lib.mh_library_get_similarity(self.h, i, j, ctypes.byref(v))
        if rc == MH_ERR_NOT_FOUND:
            return None
        _check(rc)
        return v.value


class Batch:
    """A fixed set of jobs run as one fused launch per stage (mh_batch)."""

    def __init__(self, ctx, jobs):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = len(jobs)
        self._jobs = (Job * self.n)(*jobs)
        h = ctypes.c_void_p()
        _check(self.lib.mh_batch_create(ctx.h, ctypes.cast(self._jobs, ctypes.c_void_p), self.n,
                                        ctypes.byref(h)))
        self.h = h

    def set_graph(self, enabled):
        _check(self.lib.mh_batch_set_graph(self.h, 1 if enabled else 0))

    def run(self):
        _check(self.lib.mh_batch_run(self.h))

    def tune(self, repeats=5):
        """Times each candidate launch shape on this batch and keeps the fastest."""
        t = BatchTuning()
        _check(self.lib.mh_batch_tune(self.h, repeats, ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = BatchTuning()
        _check(self.lib.mh_batch_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, topk_cells_per_wave=(0, 0), composite_threads=(0, 0)):
        t = BatchTuning()
        for g in range(2):
            t.topk_cells_per_wave[g] = topk_cells_per_wave[g]
            t.composite_threads[g] = composite_threads[g]
        _check(self.lib.mh_batch_set_tuning(self.h, ctypes.byref(t)))

    def results(self):
        res = (JobResult * self.n)()
        _check(self.lib.mh_batch_results(self.h, ctypes.cast(res, ctypes.c_void_p)))
        return [_job_result_dict(r) for r in res]

    def outputs(self, i):
        g, im = ctypes.c_void_p(), ctypes.c_void_p()
        _check(self.lib.mh_batch_job_outputs(self.h, i, ctypes.byref(g), ctypes.byref(im)))
        return g.value, im.value

    def close(self):
        if self.h:
            self.lib.mh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (mh_server)."""

    def __init__(self, ctx, memory_budget=0, max_batch=0, streams=0):
        self.ctx = ctx
        self.lib = ctx.lib
        self._keep = {}
        h = ctypes.c_void_p()
        prm = ServerParams(memory_budget, max_batch, streams)
        _check(self.lib.mh_server_create(ctx.h, ctypes.byref(prm), ctypes.byref(h)))
        self.h = h

    def submit(self, job):
        t = ctypes.c_int64()
        _check(self.lib.mh_server_submit(self.h, ctypes.byref(job), ctypes.byref(t)))
        return t.value

    def submit_many(self, jobs):
        """Queues every job at once (one call); returns their tickets."""
        arr = (Job * len(jobs))(*jobs)
        t = np.zeros(len(jobs), np.int64)
        _check(self.lib.mh_server_submit_many(self.h, arr, len(jobs), _ptr(t)))
        return t.tolist()

    def wait(self, ticket):
        res = JobResult()
        lat = ctypes.c_double()
        _check(self.lib.mh_server_wait(self.h, ticket, ctypes.byref(res), ctypes.byref(lat)))
        out = _job_result_dict(res)
        out["latency_ms"] = lat.value
        return out

    def poll(self, ticket):
        d = ctypes.c_int32()
        _check(self.lib.mh_server_poll(self.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def drain(self):
        _check(self.lib.mh_server_drain(self.h))

    def times(self, ticket):
        t = np.zeros(3, np.float64)
        _check(self.lib.mh_server_job_times(self.h, ticket, _ptr(t)))
        return t

    def stats(self):
        st = ServerStats()
        _check(self.lib.mh_server_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def set_timing_phases(self, phases):
        """Brackets the named phases of every batch launched from now on with
        HIP events on its stream slot."""
        names = [self.lib.mh_phase_name(i).decode() for i in range(MH_PHASES)]
        mask = 0
        for name in phases:
            mask |= 1 << names.index(name)
        _check(self.lib.mh_server_set_timing_phases(self.h, mask))

    def timing(self):
        """{phase: (total_ms, launches, cells)} of the timed batches completed
        since the last call."""
        ms = np.zeros(MH_PHASES, np.float64)
        n = np.zeros(MH_PHASES, np.int64)
        c = np.zeros(MH_PHASES, np.int64)
        _check(self.lib.mh_server_timing(self.h, _ptr(ms), _ptr(n), _ptr(c)))
        return {self.lib.mh_phase_name(i).decode(): (float(ms[i]), int(n[i]), int(c[i]))
                for i in range(MH_PHASES)}

    def close(self):
        if self.h:
            self.lib.mh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
