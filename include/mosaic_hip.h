/*
 * mosaic_hip.h -- C ABI of the MI355X (gfx950) tile-matching path of
 * mosaic-art-rust.
 *
 * Every entry point returns an int status (MH_OK = 0, negative on error;
 * mh_last_error() gives a thread-local message).  Host pointers are
 * caller-owned; device memory is owned by the context / library handles.
 * One in-flight call per context.  Material identity is an index
 * 0..n-1 (the Rust side keeps its PathBuf <-> index maps, as
 * SimilarityDatabase::path_to_index does, src/similarity.rs:33-36).
 * Grids are row-major int32 cells, grid[y*grid_w + x], MH_EMPTY = -1 for
 * None (the reference's Vec<Vec<Option<PathBuf>>>).
 *
 * Reference interfaces replaced (paths relative to the mosaic-art-rust
 * repository):
 *   mh_library_*              SimilarityDatabase (src/similarity.rs:30-137)
 *                             + the kd-tree built at src/main.rs:145-151
 *   mh_average_lab,
 *   mh_cell_lab_means         MosaicGeneratorImpl::calculate_average_lab
 *                             (src/lib.rs:94-119) per cell (src/main.rs:536-541)
 *   mh_nearest_n              kiddo nearest_n::<SquaredEuclidean>
 *                             (src/main.rs:348-352)
 *   mh_place                  find_and_use_best_tile_with_position +
 *                             fallback_tile_selection + UsageTracker
 *                             (src/main.rs:309-475, src/lib.rs:57-89)
 *   mh_total_cost             AdjacencyPenaltyCalculator::calculate_total_cost
 *                             (src/adjacency.rs:91-130)
 *   mh_penalties              AdjacencyPenaltyCalculator::calculate_penalty
 *                             (src/adjacency.rs:63-88)
 *   mh_swap_deltas            AdjacencyPenaltyCalculator::calculate_swap_delta
 *                             (src/adjacency.rs:133-214)
 *   mh_anneal                 MosaicOptimizer::optimize / optimize_greedy
 *                             (src/optimizer.rs:41-198)
 *   mh_average_rgb            MosaicGenerator::calculate_average_rgb
 *                             (src/main.rs:687-708)
 *   mh_optimal_adjustments    calculate_optimal_adjustment
 *                             (src/color_adjustment.rs:99-154)
 *   mh_adjust_pixels          ColorAdjustment::apply_to_image / adjust_pixel
 *                             (src/color_adjustment.rs:40-96)
 *   mh_resize                 MosaicGenerator::resize_image (src/main.rs:710-738)
 *   mh_composite              tile compositing (src/main.rs:548-590, :625-664)
 *   mh_generate               MosaicGenerator::generate_mosaic (src/main.rs:477-685)
 *   mh_batch_*, mh_server_*   no counterpart: the reference runs one job per
 *                             process (src/main.rs:117-163); these run many
 *                             generate_mosaic jobs per launch / as a request loop
 */
#ifndef MOSAIC_HIP_H
#define MOSAIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_ERR_INVALID_ARGUMENT (-1)
#define MH_ERR_NO_DEVICE (-2)
#define MH_ERR_HIP (-3)
#define MH_ERR_OUT_OF_MEMORY (-4)
#define MH_ERR_NOT_FOUND (-5)
#define MH_ERR_UNSUPPORTED (-6)

#define MH_EMPTY (-1)

/* Largest candidate list the placement keeps per cell: K = min(n, 100),
 * src/main.rs:348. */
#define MH_MAX_CANDIDATES 100

typedef struct mh_context mh_context;
typedef struct mh_library mh_library;

/* ---------------------------------------------------------------- errors */
const char *mh_last_error(void);
/* Version of the structure layouts below (2: score_mode in the placement and
 * job parameters). */
int32_t mh_abi_version(void);

/* --------------------------------------------------------------- context */
/* Binds a HIP device and a stream; owns grow-only scratch buffers. */
int mh_context_create(int32_t device, mh_context **out);
void mh_context_destroy(mh_context *ctx);
int mh_context_synchronize(mh_context *ctx);
/* The context's hipStream_t (as void*), for callers that order their own
 * work against it. */
void *mh_context_stream(mh_context *ctx);

/* Per-phase device timing: when enabled, every phase the library enqueues
 * is bracketed by HIP events on the context stream.  mh_context_timing
 * synchronises, returns the accumulated milliseconds and launch counts per
 * phase since the last call, and clears them.  Phase names: mh_phase_name.
 * (No counterpart in the reference, whose time_tracker.rs times on the host.) */
#define MH_PHASE_CELL_LAB 0
#define MH_PHASE_TOPK 1
#define MH_PHASE_PLACE 2
#define MH_PHASE_ANNEAL 3
#define MH_PHASE_RESIZE 4
#define MH_PHASE_ADJUST 5
#define MH_PHASE_COMPOSITE 6
#define MH_PHASE_SIM 7
#define MH_PHASES 8
int mh_context_set_timing(mh_context *ctx, int32_t enabled);
/* Times only the phases whose bit (1 << MH_PHASE_x) is set in phase_mask;
 * each timed phase adds two event records to the stream. */
int mh_context_set_timing_phases(mh_context *ctx, uint32_t phase_mask);
int mh_context_timing(mh_context *ctx, double ms[MH_PHASES], int64_t launches[MH_PHASES]);
const char *mh_phase_name(int32_t phase);

/* Device buffers for the device-resident entry points (mh_generate_device):
 * plain hipMalloc'd memory on the context's device, and copies ordered on the
 * context stream (mh_copy_* block until the copy is done). */
int mh_device_alloc(mh_context *ctx, size_t bytes, void **out);
int mh_device_free(mh_context *ctx, void *ptr);
int mh_copy_to_device(mh_context *ctx, void *dst, const void *src, size_t bytes);
int mh_copy_to_host(mh_context *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel, enqueued on the context
 * stream (also the known-size copy used to calibrate HBM byte counters). */
int mh_copy_on_device(mh_context *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy of rows x row_bytes between pitched buffers (e.g. a
 * crop of a larger frame), enqueued on the context stream. */
int mh_copy_2d_on_device(mh_context *ctx, void *dst, size_t dst_pitch, const void *src,
                         size_t src_pitch, size_t row_bytes, size_t rows);

/* --------------------------------------------------------------- library */
/* SimilarityDatabase::new + add_tile x n (src/similarity.rs:51-66): the Lab
 * table (n x {l,a,b} f32) becomes device resident. n may be 0. */
int mh_library_from_lab(mh_context *ctx, const float *lab, int32_t n, mh_library **out);
/* Material ingest (src/main.rs:263-301): n RGB8 images of w x h
 * (n*h*w*3 bytes, row-major); Lab = calculate_average_lab of each, computed
 * on the GPU.  The pixels stay resident for compositing. */
int mh_library_from_pixels(mh_context *ctx, const uint8_t *rgb, int32_t n, int32_t w, int32_t h,
                           mh_library **out);
/* Material ingest of images of different sizes (load_tiles' per-file
 * process_tile, src/main.rs:166-301): image i is widths[i] x heights[i]
 * RGB8, row-major, at rgb + offsets[i].  Each dimension is >= 1 and an image
 * holds at most 2^28 pixels.  mh_average_lab_many writes
 * calculate_average_lab of each image (out = n x 3) from one batched
 * launch; mh_library_from_images keeps the pixels resident and resizes each
 * material from its own size when tiles are composited
 * (resize_image, src/main.rs:710-738). */
int mh_average_lab_many(mh_context *ctx, const uint8_t *rgb, const int64_t *offsets,
                        const int32_t *widths, const int32_t *heights, int32_t n, float *out);
int mh_library_from_images(mh_context *ctx, const uint8_t *rgb, const int64_t *offsets,
                           const int32_t *widths, const int32_t *heights, int32_t n,
                           mh_library **out);
void mh_library_destroy(mh_library *lib);
int32_t mh_library_size(const mh_library *lib);
/* SimilarityDatabase::get_lab_color for every material (n x 3 f32). */
int mh_library_lab(const mh_library *lib, float *out);
/* build_similarities (src/similarity.rs:69-82): the n x n f32 matrix,
 * resident in HBM. */
int mh_library_build_similarities(mh_library *lib);
/* Read the matrix back: packed = 0 -> full n x n row-major; packed = 1 ->
 * the reference's upper-triangular Vec<f32>, n(n-1)/2 entries,
 * position i*n - i*(i+1)/2 + j - i - 1 (src/similarity.rs:100). */
int mh_library_similarities(mh_library *lib, float *out, int32_t packed);
/* get_similarity by index (src/similarity.rs:85-103): 0.0 on the diagonal;
 * MH_ERR_NOT_FOUND when an index is outside 0..n-1 (the reference's None). */
int mh_library_get_similarity(mh_library *lib, int32_t i, int32_t j, float *out);

/* -------------------------------------------------------------- a1: Lab */
int mh_average_lab(mh_context *ctx, const uint8_t *rgb, int32_t w, int32_t h, float out[3]);
/* Average Lab of every cell of a width x height RGB8 target split into
 * grid_w x grid_h cells of (width/grid_w) x (height/grid_h) pixels;
 * out = grid_w*grid_h x 3, raster order. */
int mh_cell_lab_means(mh_context *ctx, const uint8_t *rgb, int32_t width, int32_t height,
                      int32_t grid_w, int32_t grid_h, float *out);

/* --------------------------------------------------- a3: nearest materials */
/* For each of nq query Labs, the k = min(k, n) nearest materials by
 * squared Lab distance, ascending; exact ties ordered by material index.
 * idx/sqdist: nq x k. */
int mh_nearest_n(mh_context *ctx, const mh_library *lib, const float *query, int32_t nq, int32_t k,
                 int32_t *idx, float *sqdist);

/* ------------------------------------------------------ a4-a9: placement */
/* Candidate score of the placement scan.  MH_SCORE_SQDIST is the CLI's
 * squared Lab distance + penalty (src/main.rs:385-398); MH_SCORE_DISTANCE is
 * the GUI generator's Euclidean distance + penalty (sqrt of the same sum,
 * src/gui/app_full.rs:1436-1448).  Candidates, bans, usage limits and
 * fallbacks are the same in both. */
#define MH_SCORE_SQDIST 0
#define MH_SCORE_DISTANCE 1

typedef struct {
    int32_t grid_w, grid_h;
    int32_t max_usage;   /* UsageTracker::new(max_usage_per_image)          */
    float weight;        /* adjacency_penalty_weight; <= 0 disables penalty */
    int32_t score_mode;  /* MH_SCORE_SQDIST or MH_SCORE_DISTANCE             */
} mh_place_params;

typedef struct {
    int32_t placed, main_path, fallback_ban, fallback_near, resets, empty;
} mh_place_stats;

/* Places cells in `order` (n_order cell indices y*grid_w+x; NULL = raster
 * order over the whole grid, as generate_mosaic does) using cell_lab
 * (grid_w*grid_h x 3).  grid (grid_w*grid_h) and usage (n) are in/out
 * state; a fresh mosaic passes an all-MH_EMPTY grid and zero usage.
 * Optional outputs per cell: stage (0 scored scan, 1 fallback
 * first-placeable, 2 nearest_one, 255 untouched), dist (chosen candidate's
 * colour distance as the score mode measures it), score (its total score). */
int mh_place(mh_context *ctx, const mh_library *lib, const float *cell_lab,
             const mh_place_params *params, const int32_t *order, int32_t n_order, int32_t *grid,
             uint32_t *usage, uint8_t *stage, float *dist, float *score, mh_place_stats *stats);

/* ----------------------------------------------- a10-a12: cost, SA */
int mh_total_cost(mh_context *ctx, const mh_library *lib, const int32_t *grid, int32_t grid_w,
                  int32_t grid_h, float *out);
/* calculate_penalty for q (candidate material, x, y) queries against grid;
 * queries = q x 3 int32 {material, x, y}; out = q f32 (already weighted). */
int mh_penalties(mh_context *ctx, const mh_library *lib, const int32_t *grid, int32_t grid_w,
                 int32_t grid_h, float weight, const int32_t *queries, int32_t q, float *out);
/* calculate_swap_delta for p proposals {x1, y1, x2, y2} (p x 4 int32). */
int mh_swap_deltas(mh_context *ctx, const mh_library *lib, const int32_t *grid, int32_t grid_w,
                   int32_t grid_h, const int32_t *pairs, int32_t p, float *out);

typedef struct {
    int64_t max_iterations;     /* OptimizationConfig::max_iterations        */
    float initial_temperature;  /* default 100.0                             */
    float temperature_decay;    /* default 0.99995                           */
    uint64_t seed;              /* counter-RNG key                           */
    int32_t greedy;             /* 1 = optimize_greedy                       */
} mh_anneal_params;

typedef struct {
    float initial_cost, final_cost, best_cost;
    int64_t improved_count, accepted_count, iterations;
    int64_t evaluated;  /* proposals that reached the delta (not skipped)   */
    int64_t rounds;     /* speculative rounds the device used               */
} mh_anneal_result;

/* Simulated annealing on grid (in/out).  Proposal it draws x1,y1,x2,y2 and
 * u from a counter RNG keyed by (seed, it); the device evaluates many
 * proposals per round and commits them in order up to the first one whose
 * neighbourhood an earlier accepted swap of the round touched, so the
 * result equals the one-at-a-time chain bit for bit.  Optional traces
 * (max_iterations entries): delta (NaN when skipped), accept flag. */
int mh_anneal(mh_context *ctx, const mh_library *lib, int32_t *grid, int32_t grid_w,
              int32_t grid_h, const mh_anneal_params *params, mh_anneal_result *result,
              float *trace_delta, uint8_t *trace_accept);

/* ----------------------------------- a13-a16: colour adjustment, compositing */
int mh_average_rgb(mh_context *ctx, const uint8_t *rgb, int32_t w, int32_t h, uint8_t out[3]);
/* calculate_optimal_adjustment for n (tile, target) RGB8 pairs; out = n x 4
 * {brightness, contrast, hue_shift, saturation} after ColorAdjustment::new's
 * clamps.  strength is used as given (the generator clamps it to [0,1]). */
int mh_optimal_adjustments(mh_context *ctx, const uint8_t *tile_rgb, const uint8_t *target_rgb,
                           int32_t n, float strength, float *out);
/* ColorAdjustment::apply_to_image over npx RGB8 pixels with adj[4]. */
int mh_adjust_pixels(mh_context *ctx, const float adj[4], const uint8_t *in, int64_t npx,
                     uint8_t *out);
/* Lanczos3 resize of one RGB8 image sw x sh -> dw x dh. */
int mh_resize(mh_context *ctx, const uint8_t *src, int32_t sw, int32_t sh, uint8_t *dst, int32_t dw,
              int32_t dh);
/* Composite a placed grid: lib must hold pixels; out is
 * (grid_w*tw) x (grid_h*th) x 3 with tw = width/grid_w, th = height/grid_h.
 * strength > 0 applies the per-cell colour adjustment. */
int mh_composite(mh_context *ctx, const mh_library *lib, const uint8_t *target, int32_t width,
                 int32_t height, const int32_t *grid, int32_t grid_w, int32_t grid_h,
                 float strength, uint8_t *out);

/* ------------------------------------------------------ whole mosaic job */
typedef struct {
    int32_t grid_w, grid_h;
    int32_t max_usage;
    float weight;                   /* adjacency_penalty_weight              */
    float color_strength;           /* clamped to [0,1] (src/main.rs:162)    */
    int32_t enable_optimization;    /* SA runs when set and weight > 0       */
    int64_t optimization_iterations;
    float initial_temperature, temperature_decay;
    uint64_t seed;
    int32_t score_mode;             /* MH_SCORE_SQDIST or MH_SCORE_DISTANCE  */
} mh_job_params;

typedef struct {
    mh_place_stats place;
    mh_anneal_result anneal;
    int32_t out_width, out_height;
} mh_job_result;

/* generate_mosaic minus file I/O: cell Lab means -> candidates ->
 * placement -> (SA) -> composite.  grid_out (grid_w*grid_h) and image_out
 * (out_width*out_height*3, out_width = grid_w*(width/grid_w)) are optional. */
int mh_generate(mh_context *ctx, const mh_library *lib, const uint8_t *target, int32_t width,
                int32_t height, const mh_job_params *params, int32_t *grid_out, uint8_t *image_out,
                mh_job_result *result);

/* Device-resident form of mh_generate: target, grid_out and image_out are
 * device pointers; work is enqueued on the context stream and the call
 * returns without synchronising.  mh_job_result (if non-NULL) is filled by
 * mh_generate_device_result after the stream is synchronised. */
int mh_generate_device(mh_context *ctx, const mh_library *lib, const uint8_t *d_target,
                       int32_t width, int32_t height, const mh_job_params *params,
                       int32_t *d_grid_out, uint8_t *d_image_out);
int mh_generate_device_result(mh_context *ctx, mh_job_result *result);

/* Allocation-free form of mh_generate_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, outputs and result (read with
 * mh_generate_device_result) as mh_generate_device, bit for bit.  Its working
 * memory (cell Labs, candidate rows, usage counts, adjustments, temperature
 * table, result words, and the grid when d_grid_out is NULL) lives in one
 * static device array of MH_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a job that needs more fails with MH_ERR_UNSUPPORTED (about
 * 1.1 KB per cell: 60,000 cells fit).  It calls no hipMalloc or hipFree, so
 * the library's resized tiles for the job's cell size must already exist
 * when d_image_out is given (any mh_generate / mh_generate_device /
 * mh_composite at that cell size builds them; else MH_ERR_UNSUPPORTED), and
 * the adjacency terms come from the Lab table instead of the library's term
 * matrix (the same bits).  Phase timing is not recorded.  Every call on one
 * GPU shares that GPU's array and leaves its contents behind: calls that
 * overlap on one GPU (two contexts, or two host threads) race, so callers
 * order them. */
#define MH_NOALLOC_SCRATCH_BYTES (64ull << 20)
int mh_generate_device_noalloc(mh_context *ctx, const mh_library *lib, const uint8_t *d_target,
                               int32_t width, int32_t height, const mh_job_params *params,
                               int32_t *d_grid_out, uint8_t *d_image_out);

/* ------------------------------------------- batched and served jobs */
/* One mosaic job for the batch and server entry points.  d_target is a
 * device RGB8 image with target_pitch bytes per row (0 = width*3).  Outputs
 * are device buffers (grid_w*grid_h int32; (grid_w*tw) x (grid_h*th) x 3
 * bytes); when NULL the library composites into working memory of its own
 * (readable through mh_batch_job_outputs for batches, discarded for served
 * jobs).  The library must hold pixels and belong to the same context.
 * Each job's result equals mh_generate_device's for the same job alone. */
typedef struct {
    const mh_library *lib;
    const uint8_t *d_target;
    int64_t target_pitch;
    int32_t width, height;
    mh_job_params params;
    int32_t *d_grid_out;
    uint8_t *d_image_out;
} mh_job;

/* Device working memory the job needs while it runs (bytes), or < 0 on a
 * validation error. */
int64_t mh_job_footprint(mh_context *ctx, const mh_job *job);

/* A fixed set of jobs run as one fused launch per stage (cell Lab, nearest
 * materials, placement chains, annealing chains, adjustment, composite) on
 * the context stream.  Creation validates the jobs, builds the libraries'
 * caches and allocates all working memory; runs allocate nothing.  With the
 * graph enabled the first run captures the launch sequence into a hipGraph
 * and later runs replay it (inputs are re-read from the same addresses). */
typedef struct mh_batch mh_batch;
int mh_batch_create(mh_context *ctx, const mh_job *jobs, int32_t n_jobs, mh_batch **out);
int mh_batch_set_graph(mh_batch *batch, int32_t enabled);
int mh_batch_run(mh_batch *batch);
/* Synchronises the context stream; results has n_jobs entries. */
int mh_batch_results(mh_batch *batch, mh_job_result *results);
/* Device grid and image of job i (the caller's buffers or the batch's own). */
int mh_batch_job_outputs(mh_batch *batch, int32_t i, int32_t **d_grid, uint8_t **d_image);
void mh_batch_destroy(mh_batch *batch);

/* Launch shapes of a batch's wide stages, per job group (0: the long jobs,
 * 1: the rest; zeros where a group or stage is absent) and the average
 * launch time the shape measured. */
typedef struct {
    int32_t n_groups;
    int32_t topk_cells_per_wave[2]; /* query cells one nearest-material wave scans for */
    int32_t composite_threads[2];   /* output pixels per composite workgroup */
    float topk_ms[2], composite_ms[2];
} mh_batch_tuning;
/* Autotuner (no counterpart in the reference, which has no launch shapes):
 * runs the batch once, then times `repeats` launches of each candidate shape
 * of the nearest-material stage (1, 2, 4, 8 cells per wave) and the composite
 * stage (64, 128, 256 threads) on the batch's own data, keeps the fastest
 * and drops a captured graph so the next run captures the tuned shapes.
 * Every shape computes the same outputs.  tuning may be NULL; 1 <= repeats
 * <= 100.  Synchronises the context stream. */
int mh_batch_tune(mh_batch *batch, int32_t repeats, mh_batch_tuning *tuning);
/* The shapes in use (the heuristics' until mh_batch_tune runs; times 0). */
int mh_batch_get_tuning(mh_batch *batch, mh_batch_tuning *tuning);
/* Use the given shapes (say, a tuned batch's for a later batch of the same
 * kind): topk_cells_per_wave in [1, 8] (0 keeps the heuristic's),
 * composite_threads 64, 128 or 256 (0 keeps the current); times and
 * n_groups are ignored.  Drops a captured graph. */
int mh_batch_set_tuning(mh_batch *batch, const mh_batch_tuning *tuning);

/* Continuous-batching request loop.  A scheduler thread admits queued jobs
 * while their working memory fits memory_budget, and launches the admitted
 * jobs as one batch on a free slot (at most `streams` batches in flight, at
 * most max_batch jobs each); jobs that arrive while batches run join the next
 * one as slots and memory free up.  A batch takes the oldest queued job and
 * the queued jobs among the next 8 * max_batch nearest to it in cell count,
 * so that one batch's placement chains are of similar length; once the whole
 * queue fits those 8 * max_batch, the job with the most cells takes the
 * oldest one's place (until the oldest has been passed over 8 times), so a
 * backlog ends on its shortest chains.  A job's result does not depend on
 * its batch; with at most max_batch jobs queued, admission is first-come
 * first-served.  The batches' wide stages
 * (cell Labs, K nearest) run on two streams, consecutive batches
 * alternating, their placement chains and annealing on up to four more
 * beside later batches' wide stages, and their output stages (adjustment,
 * composite) on an output stream per chain stream.  Working memory is
 * allocated in chunks of 1 GiB (or one job's size, if larger) as admitted
 * jobs need it, never beyond memory_budget; a job that finds no device
 * memory while nothing else runs fails with MH_ERR_OUT_OF_MEMORY.  Libraries
 * stay resident with their term matrices and resized tiles.  The context's
 * other entry points must not run concurrently with a live server. */
typedef struct {
    int64_t memory_budget;  /* bytes of device working memory (0 = 64 GiB) */
    int32_t max_batch;      /* jobs per launch (0 = 128)                    */
    int32_t streams;        /* batches in flight, 1..16 (0 = 8)             */
} mh_server_params;

typedef struct {
    int64_t submitted, completed, batches;
    int32_t max_batch_jobs;
    int64_t memory_in_use, memory_peak, memory_budget;
} mh_server_stats;

typedef struct mh_server mh_server;
int mh_server_create(mh_context *ctx, const mh_server_params *params, mh_server **out);
/* Queues a job (its buffers must stay valid until it completes); returns a
 * ticket.  Fails at once if the job is invalid or larger than the budget. */
int mh_server_submit(mh_server *server, const mh_job *job, int64_t *ticket);
/* Queues n jobs at once (tickets[i] for jobs[i]), in order: one lock and one
 * wake-up of the scheduler instead of n.  Every job is checked first; if any
 * is invalid or larger than the budget, none is queued and the error names
 * it. */
int mh_server_submit_many(mh_server *server, const mh_job *jobs, int32_t n, int64_t *tickets);
/* Blocks until the job is done; latency_ms = completion - submission. */
int mh_server_wait(mh_server *server, int64_t ticket, mh_job_result *result, double *latency_ms);
int mh_server_poll(mh_server *server, int64_t ticket, int32_t *done);
/* Blocks until every submitted job is done. */
int mh_server_drain(mh_server *server);
/* Host steady-clock times (ms) of submission, launch and completion. */
int mh_server_job_times(mh_server *server, int64_t ticket, double times_ms[3]);
int mh_server_get_stats(mh_server *server, mh_server_stats *stats);
/* Per-phase device timing of the server's batches (no counterpart in the
 * reference): phases whose bit (1 << MH_PHASE_x) is set are bracketed by HIP
 * events on the stream slot that runs each batch launched from now on.
 * mh_server_timing returns, per phase, the milliseconds, launches and cells
 * (summed over each timed launch's jobs) of the batches completed since the
 * last call, and clears them; any output may be NULL. */
int mh_server_set_timing_phases(mh_server *server, uint32_t phase_mask);
int mh_server_timing(mh_server *server, double ms[MH_PHASES], int64_t launches[MH_PHASES],
                     int64_t cells[MH_PHASES]);
/* Finishes queued and running jobs, then frees the server. */
void mh_server_destroy(mh_server *server);

#ifdef __cplusplus
}
#endif
#endif
