//! Raw bindings to libmosaic_hip.so, field for field and argument for
//! argument from include/mosaic_hip.h (tests/test_rust_binding.py checks the
//! two against each other).  Every function returns MH_OK or a negative
//! MH_ERR_* code and leaves a message for the calling thread in
//! mh_last_error(); `check` turns that into a Result.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::{c_char, c_int, c_void};

pub const MH_OK: i32 = 0;
pub const MH_ERR_INVALID_ARGUMENT: i32 = -1;
pub const MH_ERR_NO_DEVICE: i32 = -2;
pub const MH_ERR_HIP: i32 = -3;
pub const MH_ERR_OUT_OF_MEMORY: i32 = -4;
pub const MH_ERR_NOT_FOUND: i32 = -5;
pub const MH_ERR_UNSUPPORTED: i32 = -6;
pub const MH_EMPTY: i32 = -1;
pub const MH_MAX_CANDIDATES: i32 = 100;
pub const MH_PHASE_CELL_LAB: i32 = 0;
pub const MH_PHASE_TOPK: i32 = 1;
pub const MH_PHASE_PLACE: i32 = 2;
pub const MH_PHASE_ANNEAL: i32 = 3;
pub const MH_PHASE_RESIZE: i32 = 4;
pub const MH_PHASE_ADJUST: i32 = 5;
pub const MH_PHASE_COMPOSITE: i32 = 6;
pub const MH_PHASE_SIM: i32 = 7;
pub const MH_PHASES: usize = 8;
pub const MH_SCORE_SQDIST: i32 = 0;
pub const MH_SCORE_DISTANCE: i32 = 1;
pub const MH_NOALLOC_SCRATCH_BYTES: u64 = 67108864;

#[repr(C)]
pub struct mh_batch {
    _p: [u8; 0],
}

#[repr(C)]
pub struct mh_context {
    _p: [u8; 0],
}

#[repr(C)]
pub struct mh_library {
    _p: [u8; 0],
}

#[repr(C)]
pub struct mh_server {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_place_params {
    pub grid_w: i32,
    pub grid_h: i32,
    pub max_usage: i32,
    pub weight: f32,
    pub score_mode: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_place_stats {
    pub placed: i32,
    pub main_path: i32,
    pub fallback_ban: i32,
    pub fallback_near: i32,
    pub resets: i32,
    pub empty: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_anneal_params {
    pub max_iterations: i64,
    pub initial_temperature: f32,
    pub temperature_decay: f32,
    pub seed: u64,
    pub greedy: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_anneal_result {
    pub initial_cost: f32,
    pub final_cost: f32,
    pub best_cost: f32,
    pub improved_count: i64,
    pub accepted_count: i64,
    pub iterations: i64,
    pub evaluated: i64,
    pub rounds: i64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_job_params {
    pub grid_w: i32,
    pub grid_h: i32,
    pub max_usage: i32,
    pub weight: f32,
    pub color_strength: f32,
    pub enable_optimization: i32,
    pub optimization_iterations: i64,
    pub initial_temperature: f32,
    pub temperature_decay: f32,
    pub seed: u64,
    pub score_mode: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_job_result {
    pub place: mh_place_stats,
    pub anneal: mh_anneal_result,
    pub out_width: i32,
    pub out_height: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct mh_job {
    pub lib: *const mh_library,
    pub d_target: *const u8,
    pub target_pitch: i64,
    pub width: i32,
    pub height: i32,
    pub params: mh_job_params,
    pub d_grid_out: *mut i32,
    pub d_image_out: *mut u8,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_batch_tuning {
    pub n_groups: i32,
    pub topk_cells_per_wave: [i32; 2],
    pub composite_threads: [i32; 2],
    pub topk_ms: [f32; 2],
    pub composite_ms: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_server_params {
    pub memory_budget: i64,
    pub max_batch: i32,
    pub streams: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct mh_server_stats {
    pub submitted: i64,
    pub completed: i64,
    pub batches: i64,
    pub max_batch_jobs: i32,
    pub memory_in_use: i64,
    pub memory_peak: i64,
    pub memory_budget: i64,
}

extern "C" {
    pub fn mh_last_error() -> *const c_char;
    pub fn mh_abi_version() -> i32;
    pub fn mh_context_create(device: i32, out: *mut *mut mh_context) -> c_int;
    pub fn mh_context_destroy(ctx: *mut mh_context);
    pub fn mh_context_synchronize(ctx: *mut mh_context) -> c_int;
    pub fn mh_context_stream(ctx: *mut mh_context) -> *mut c_void;
    pub fn mh_context_set_timing(ctx: *mut mh_context, enabled: i32) -> c_int;
    pub fn mh_context_set_timing_phases(ctx: *mut mh_context, phase_mask: u32) -> c_int;
    pub fn mh_context_timing(ctx: *mut mh_context, ms: *mut f64, launches: *mut i64) -> c_int;
    pub fn mh_phase_name(phase: i32) -> *const c_char;
    pub fn mh_device_alloc(ctx: *mut mh_context, bytes: usize, out: *mut *mut c_void) -> c_int;
    pub fn mh_device_free(ctx: *mut mh_context, ptr: *mut c_void) -> c_int;
    pub fn mh_copy_to_device(ctx: *mut mh_context, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn mh_copy_to_host(ctx: *mut mh_context, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn mh_copy_on_device(ctx: *mut mh_context, dst: *mut c_void, src: *const c_void, bytes: usize) -> c_int;
    pub fn mh_copy_2d_on_device(ctx: *mut mh_context, dst: *mut c_void, dst_pitch: usize, src: *const c_void, src_pitch: usize, row_bytes: usize, rows: usize) -> c_int;
    pub fn mh_library_from_lab(ctx: *mut mh_context, lab: *const f32, n: i32, out: *mut *mut mh_library) -> c_int;
    pub fn mh_library_from_pixels(ctx: *mut mh_context, rgb: *const u8, n: i32, w: i32, h: i32, out: *mut *mut mh_library) -> c_int;
    pub fn mh_average_lab_many(ctx: *mut mh_context, rgb: *const u8, offsets: *const i64, widths: *const i32, heights: *const i32, n: i32, out: *mut f32) -> c_int;
    pub fn mh_library_from_images(ctx: *mut mh_context, rgb: *const u8, offsets: *const i64, widths: *const i32, heights: *const i32, n: i32, out: *mut *mut mh_library) -> c_int;
    pub fn mh_library_destroy(lib: *mut mh_library);
    pub fn mh_library_size(lib: *const mh_library) -> i32;
    pub fn mh_library_lab(lib: *const mh_library, out: *mut f32) -> c_int;
    pub fn mh_library_build_similarities(lib: *mut mh_library) -> c_int;
    pub fn mh_library_similarities(lib: *mut mh_library, out: *mut f32, packed: i32) -> c_int;
    pub fn mh_library_get_similarity(lib: *mut mh_library, i: i32, j: i32, out: *mut f32) -> c_int;
    pub fn mh_average_lab(ctx: *mut mh_context, rgb: *const u8, w: i32, h: i32, out: *mut f32) -> c_int;
    pub fn mh_cell_lab_means(ctx: *mut mh_context, rgb: *const u8, width: i32, height: i32, grid_w: i32, grid_h: i32, out: *mut f32) -> c_int;
    pub fn mh_nearest_n(ctx: *mut mh_context, lib: *const mh_library, query: *const f32, nq: i32, k: i32, idx: *mut i32, sqdist: *mut f32) -> c_int;
    pub fn mh_place(ctx: *mut mh_context, lib: *const mh_library, cell_lab: *const f32, params: *const mh_place_params, order: *const i32, n_order: i32, grid: *mut i32, usage: *mut u32, stage: *mut u8, dist: *mut f32, score: *mut f32, stats: *mut mh_place_stats) -> c_int;
    pub fn mh_total_cost(ctx: *mut mh_context, lib: *const mh_library, grid: *const i32, grid_w: i32, grid_h: i32, out: *mut f32) -> c_int;
    pub fn mh_penalties(ctx: *mut mh_context, lib: *const mh_library, grid: *const i32, grid_w: i32, grid_h: i32, weight: f32, queries: *const i32, q: i32, out: *mut f32) -> c_int;
    pub fn mh_swap_deltas(ctx: *mut mh_context, lib: *const mh_library, grid: *const i32, grid_w: i32, grid_h: i32, pairs: *const i32, p: i32, out: *mut f32) -> c_int;
    pub fn mh_anneal(ctx: *mut mh_context, lib: *const mh_library, grid: *mut i32, grid_w: i32, grid_h: i32, params: *const mh_anneal_params, result: *mut mh_anneal_result, trace_delta: *mut f32, trace_accept: *mut u8) -> c_int;
    pub fn mh_average_rgb(ctx: *mut mh_context, rgb: *const u8, w: i32, h: i32, out: *mut u8) -> c_int;
    pub fn mh_optimal_adjustments(ctx: *mut mh_context, tile_rgb: *const u8, target_rgb: *const u8, n: i32, strength: f32, out: *mut f32) -> c_int;
    pub fn mh_adjust_pixels(ctx: *mut mh_context, adj: *const f32, in_: *const u8, npx: i64, out: *mut u8) -> c_int;
    pub fn mh_resize(ctx: *mut mh_context, src: *const u8, sw: i32, sh: i32, dst: *mut u8, dw: i32, dh: i32) -> c_int;
    pub fn mh_composite(ctx: *mut mh_context, lib: *const mh_library, target: *const u8, width: i32, height: i32, grid: *const i32, grid_w: i32, grid_h: i32, strength: f32, out: *mut u8) -> c_int;
    pub fn mh_generate(ctx: *mut mh_context, lib: *const mh_library, target: *const u8, width: i32, height: i32, params: *const mh_job_params, grid_out: *mut i32, image_out: *mut u8, result: *mut mh_job_result) -> c_int;
    pub fn mh_generate_device(ctx: *mut mh_context, lib: *const mh_library, d_target: *const u8, width: i32, height: i32, params: *const mh_job_params, d_grid_out: *mut i32, d_image_out: *mut u8) -> c_int;
    pub fn mh_generate_device_result(ctx: *mut mh_context, result: *mut mh_job_result) -> c_int;
    pub fn mh_generate_device_noalloc(ctx: *mut mh_context, lib: *const mh_library, d_target: *const u8, width: i32, height: i32, params: *const mh_job_params, d_grid_out: *mut i32, d_image_out: *mut u8) -> c_int;
    pub fn mh_job_footprint(ctx: *mut mh_context, job: *const mh_job) -> i64;
    pub fn mh_batch_create(ctx: *mut mh_context, jobs: *const mh_job, n_jobs: i32, out: *mut *mut mh_batch) -> c_int;
    pub fn mh_batch_set_graph(batch: *mut mh_batch, enabled: i32) -> c_int;
    pub fn mh_batch_run(batch: *mut mh_batch) -> c_int;
    pub fn mh_batch_results(batch: *mut mh_batch, results: *mut mh_job_result) -> c_int;
    pub fn mh_batch_job_outputs(batch: *mut mh_batch, i: i32, d_grid: *mut *mut i32, d_image: *mut *mut u8) -> c_int;
    pub fn mh_batch_destroy(batch: *mut mh_batch);
    pub fn mh_batch_tune(batch: *mut mh_batch, repeats: i32, tuning: *mut mh_batch_tuning) -> c_int;
    pub fn mh_batch_get_tuning(batch: *mut mh_batch, tuning: *mut mh_batch_tuning) -> c_int;
    pub fn mh_batch_set_tuning(batch: *mut mh_batch, tuning: *const mh_batch_tuning) -> c_int;
    pub fn mh_server_create(ctx: *mut mh_context, params: *const mh_server_params, out: *mut *mut mh_server) -> c_int;
    pub fn mh_server_submit(server: *mut mh_server, job: *const mh_job, ticket: *mut i64) -> c_int;
    pub fn mh_server_submit_many(server: *mut mh_server, jobs: *const mh_job, n: i32, tickets: *mut i64) -> c_int;
    pub fn mh_server_wait(server: *mut mh_server, ticket: i64, result: *mut mh_job_result, latency_ms: *mut f64) -> c_int;
    pub fn mh_server_poll(server: *mut mh_server, ticket: i64, done: *mut i32) -> c_int;
    pub fn mh_server_drain(server: *mut mh_server) -> c_int;
    pub fn mh_server_job_times(server: *mut mh_server, ticket: i64, times_ms: *mut f64) -> c_int;
    pub fn mh_server_get_stats(server: *mut mh_server, stats: *mut mh_server_stats) -> c_int;
    pub fn mh_server_set_timing_phases(server: *mut mh_server, phase_mask: u32) -> c_int;
    pub fn mh_server_timing(server: *mut mh_server, ms: *mut f64, launches: *mut i64, cells: *mut i64) -> c_int;
    pub fn mh_server_destroy(server: *mut mh_server);
}

/// MH_OK -> Ok(()); anything else -> the code with mh_last_error()'s message.
pub fn check(rc: c_int) -> Result<(), crate::Error> {
    if rc == MH_OK {
        return Ok(());
    }
    let msg = unsafe { std::ffi::CStr::from_ptr(mh_last_error()) }
        .to_string_lossy()
        .into_owned();
    Err(crate::Error { code: rc, message: msg })
}
