/*
 * mosaic_oracle.h -- CPU restatement of mosaic-art-rust's tile-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the mosaic-art-rust repository root).  Arithmetic contract: IEEE f32
 * with the reference's operation order, no FMA contraction
 * (-ffp-contract=off), correctly rounded f32 divide and sqrt.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_srgb_lut / orc_cbrtf / orc_expf : closed forms vs libm (<= 1 ulp)
 *   orc_rgb8_to_lab, orc_average_lab    : src/lib.rs:133-158 known answers
 *                                          (red, grey) + CIE closed form;
 *                                          bit parity with palette 0.7.6 is
 *                                          "parity unpinned" (crate absent)
 *   orc_lab_distance / orc_sim_*        : src/similarity.rs:207-218, 336-362
 *                                          + closed form sqrt(300)
 *   orc_nearest_n                       : brute-force sort of all N keys;
 *                                          src/lib.rs:213-227 known answer;
 *                                          kiddo tie order "parity unpinned"
 *   orc_place                           : hand-worked grids (ban, usage,
 *                                          reset, nearest_one), brute force
 *   orc_total_cost / orc_swap_delta     : delta == cost(after)-cost(before),
 *                                          src/adjacency.rs:274-290
 *   orc_anneal                          : schedule invariants
 *                                          (src/optimizer.rs:440-457);
 *                                          RNG streams are not reproducible
 *                                          in the reference (thread_rng)
 *   orc_average_rgb                     : src/main.rs:877-889 ([127,127,127])
 *   orc_optimal_adjustment, orc_adjust_pixel :
 *                                          src/color_adjustment.rs:161-499
 *   orc_resize_lanczos3                 : solid colour stays solid,
 *                                          src/main.rs:892-921; bit parity
 *                                          with fast_image_resize 5.1.4 is
 *                                          "parity unpinned"
 */
#ifndef MOSAIC_ORACLE_H
#define MOSAIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EMPTY (-1)

/* ---- math building blocks shared by the Lab and SA restatements ---- */
void orc_srgb_lut(float out[256]);
float orc_cbrtf(float x);
float orc_expf(float x);

/* ---- a1: colour conversion and average Lab ---- */
void orc_rgb8_to_lab(uint8_t r, uint8_t g, uint8_t b, float out[3]);
void orc_average_lab(const uint8_t *rgb, int32_t w, int32_t h, int64_t row_stride, float out[3]);
void orc_cell_lab_means(const uint8_t *rgb, int32_t width, int32_t height, int32_t grid_w,
                        int32_t grid_h, float *out);
void orc_stack_lab_means(const uint8_t *rgb, int32_t n, int32_t w, int32_t h, float *out);

/* ---- a2: similarity database ---- */
float orc_lab_distance(const float *lab1, const float *lab2);
float orc_sim(const float *lab, int32_t i, int32_t j);
void orc_sim_full(const float *lab, int32_t n, float *out);
void orc_sim_packed(const float *lab, int32_t n, float *out);
void orc_sim_rows(const float *lab, int32_t n, const int32_t *rows, int32_t nrows, float *out);

/* ---- a3: K nearest materials by (sqdist, index) ---- */
void orc_nearest_n(const float *lab, int32_t n, const float *query, int32_t k, int32_t *out_idx,
                   float *out_sqdist);
/* The same K nearest from a k-d tree (kiddo's structure, src/main.rs:145-151):
 * identical answers, ties included; orc_place queries one. */
typedef struct orc_kdtree orc_kdtree;
orc_kdtree *orc_kdtree_build(const float *lab, int32_t n);
void orc_kdtree_free(orc_kdtree *t);
void orc_kdtree_nearest_n(const orc_kdtree *t, const float *query, int32_t k, int32_t *out_idx,
                          float *out_sqdist);
void orc_nearest_n_kd(const float *lab, int32_t n, const float *queries, int32_t nq, int32_t k,
                      int32_t *out_idx, float *out_sqdist);

/* ---- a4-a9: placement ---- */
typedef struct {
    int32_t placed;         /* cells that received a material            */
    int32_t main_path;      /* cells placed by the scored candidate scan */
    int32_t fallback_ban;   /* cells placed by fallback stage 1          */
    int32_t fallback_near;  /* cells placed by nearest_one               */
    int32_t resets;         /* usage-tracker resets                      */
    int32_t empty;          /* cells left empty (empty library)          */
} orc_place_stats;

void orc_place(const float *lab, int32_t n, const float *cell_lab, int32_t grid_w, int32_t grid_h,
               int32_t max_usage, float weight, int32_t score_mode, const int32_t *order,
               int32_t n_order, int32_t *grid, uint32_t *usage, uint8_t *stage, float *dist,
               float *score, orc_place_stats *stats);

/* ---- a10-a12: cost, swap delta, annealing ---- */
float orc_total_cost(const float *lab, int32_t n, const int32_t *grid, int32_t grid_w,
                     int32_t grid_h);
float orc_swap_delta(const float *lab, int32_t n, const int32_t *grid, int32_t grid_w,
                     int32_t grid_h, int32_t x1, int32_t y1, int32_t x2, int32_t y2);

uint32_t orc_draw(uint64_t seed, uint64_t it, uint32_t k);

typedef struct {
    float initial_cost, final_cost, best_cost;
    int64_t improved_count, accepted_count, iterations, evaluated;
} orc_anneal_result;

void orc_anneal(const float *lab, int32_t n, int32_t *grid, int32_t grid_w, int32_t grid_h,
                int64_t max_iterations, float initial_temperature, float temperature_decay,
                uint64_t seed, int32_t greedy, orc_anneal_result *res, float *trace_delta,
                uint8_t *trace_accept);

/* ---- a13-a16: average RGB, colour adjustment, resize, composite ---- */
void orc_average_rgb(const uint8_t *rgb, int32_t w, int32_t h, int64_t row_stride, uint8_t out[3]);
void orc_rgb_to_hsv(float r, float g, float b, float out[3]);
void orc_hsv_to_rgb(float h, float s, float v, float out[3]);
void orc_adjustment_new(float brightness, float contrast, float hue_shift, float saturation,
                        float out[4]);
void orc_optimal_adjustment(const uint8_t tile_rgb[3], const uint8_t target_rgb[3], float strength,
                            float out[4]);
void orc_adjust_pixel(const float adj[4], const uint8_t in[3], uint8_t out[3]);
void orc_apply_adjustment(const float adj[4], const uint8_t *in, int64_t npx, uint8_t *out);
void orc_resize_lanczos3(const uint8_t *src, int32_t sw, int32_t sh, uint8_t *dst, int32_t dw,
                         int32_t dh);
void orc_composite(const uint8_t *target, int32_t width, int32_t height, const uint8_t *materials,
                   int32_t n, int32_t mw, int32_t mh, const int32_t *grid, int32_t grid_w,
                   int32_t grid_h, float strength, uint8_t *out);
void orc_composite_tiles(const uint8_t *target, int32_t width, int32_t height,
                         const uint8_t *tiles, const int32_t *grid, int32_t gw, int32_t gh,
                         float strength, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
