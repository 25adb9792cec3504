// mh_kernels.hpp -- job descriptors and host-side launchers of the gfx950
// kernels (one translation unit per kernel family; the C ABI in mh_api.hip
// only calls these launchers).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mh {

constexpr int32_t kEmpty = -1;       // MH_EMPTY
constexpr int32_t kMaxCand = 100;    // K = min(n, 100), src/main.rs:348
constexpr int32_t kCandStride = 128; // per-cell candidate row (padded)

// One candidate of a cell: material and squared distance, 8 bytes (one
// 64-bit load).  Rows of kCandStride, ranks ascending (sqdist, idx); ranks
// >= K hold {-1, +inf}.  A candidate's Lab is read from the library's lab4
// where a penalty needs it.
struct CandRec {
    int32_t m;
    float d;
};
static_assert(sizeof(CandRec) == 8, "one 64-bit word per candidate");

// One placement chain (one wave).  All pointers are device pointers.
struct PlaceJob {
    const float4* lab4;      // n material Labs {l,a,b,0}
    int32_t n;
    const CandRec* cand;      // cells x kCandStride
    const int32_t* order;     // n_order cell indices, or nullptr = raster
    int32_t n_order;
    int32_t* grid;            // cells, in/out
    uint32_t* usage;          // n, in/out
    int32_t zero_usage;       // 1: start from zero counts instead of reading `usage`
    int32_t grid_w, grid_h;
    int32_t max_usage;
    float weight;
    int32_t score_sqrt;       // 1: candidates score sqrt(sqdist) + penalty (GUI variant)
    uint8_t* stage;           // optional per-cell outputs
    float* dist;
    float* score;
    int32_t* stats;           // 6 ints: placed, main, ban, near, resets, empty
};

struct AnnealOut {
    float initial_cost, final_cost, best_cost;
    int64_t improved, accepted, iterations, evaluated, rounds;
};

// One annealing chain (one wave).
struct AnnealJob {
    const float4* lab4;
    const float* term;        // n x n adjacency terms, or nullptr (computed from lab4)
    int32_t n;
    int32_t* grid;            // cells, in/out
    int32_t grid_w, grid_h;
    int64_t max_iter;
    const float* temps;       // temperature after k non-skipped proposals
    uint64_t seed;
    int32_t greedy;
    float* trace_delta;       // optional, max_iter entries
    uint8_t* trace_accept;
    AnnealOut* result;
};

// Fixed-point Lanczos3 taps for one axis (device pointers): output o sums
// cnt[o] source pixels from lo[o] with the 16-bit weights k[o * ksize + t],
// which carry `prec` fraction bits.
struct ResizeTaps {
    const int32_t* lo;
    const int32_t* cnt;
    const int32_t* k;
    int32_t ksize;
    int32_t prec;
};

// Cells of one image: cell c's pixel (px, py) is at
//   rgb + (c / per_row) * block_row_stride + (c % per_row) * block_col_stride
//       + py * row_stride + px * 3.
struct CellSeg {
    const uint8_t* rgb;
    int64_t row_stride, block_row_stride, block_col_stride;
    int32_t per_row, cw, ch, n_cells;
    float* out;        // Lab means, n_cells x 3 (cell Lab kernels)
    uint8_t* out_rgb;  // integer RGB means, n_cells x 3 (average-RGB kernels)
};

// K nearest materials of nq query Labs against one library (SoA Lab table).
// A library's materials in blocks of 64 (median splits on the widest Lab
// axis, so each block is a compact box): slot 64 b + i of the permuted
// arrays is member i of block b (slots past n are padding), pidx holds each
// slot's material index, box[2 b] / box[2 b + 1] the block's low / high
// corner (L, a, b; w unused).
struct MatBlocks {
    const float* pL;
    const float* pA;
    const float* pB;
    const int32_t* pidx;
    const float4* box;
    int32_t nb;
};

struct TopkSeg {
    const float* mL;
    const float* mA;
    const float* mB;
    const float* query;  // nq x 3
    int32_t* out_idx;    // optional
    float* out_d;        // optional
    CandRec* out_cand;   // optional; rows padded with {-1, +inf} up to out_stride
    int32_t n, nq, k, out_stride;
    MatBlocks blk;
};

// Finishing one placed grid: per-cell adjustment and compositing.
struct CompSeg {
    const int32_t* grid;
    int32_t gw, gh, tw, th;
    const uint8_t* tiles;       // n x th x tw x 3 resized materials
    const uint8_t* tile_avg;    // n x 3
    const uint8_t* target_avg;  // cells x 3
    float strength;
    int32_t adjust;             // 0: plain tiles (adj unused)
    float4* adj;                // cells
    uint8_t* out;               // (gw*tw) x (gh*th) x 3
    int32_t narrow;             // composite_narrow(n, tw, th)
};

// Cell Lab means; with out_rgb, also the cells' integer average RGB
// (calculate_average_rgb) from the same pixel pass.
void launch_cell_lab(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                     int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                     int32_t n_cells, const float* lut, float* out, hipStream_t stream,
                     uint8_t* out_rgb = nullptr);
void launch_sim_full(const float4* lab, int32_t n, float* out, hipStream_t stream);
void launch_sim_packed(const float4* lab, int32_t n, float* out, hipStream_t stream);
void launch_term_full(const float4* lab, int32_t n, float* out, hipStream_t stream);
// out_idx / out_d / out_cand may be null.
void launch_topk(const float* mL, const float* mA, const float* mB, int32_t n, const MatBlocks& blk,
                 const float* query, int32_t nq, int32_t k, int32_t* out_idx, float* out_d,
                 CandRec* out_cand, int32_t out_stride, hipStream_t stream);
// Batched launches over segment tables: starts[j] is the first block of
// segment j (starts[n_segs] = total).  *_blocks give each segment's count.
// cells per block (4 or 16) for a launch over total_cells cells
int32_t cell_lab_cells_per_block(int64_t total_cells);
int32_t cell_lab_blocks(int32_t n_cells, int32_t cells_per_block);
void launch_cell_lab_batch(const CellSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                           int32_t total_blocks, int32_t cells_per_block, const float* lut,
                           hipStream_t stream);
// one wave per run of `run` consecutive queries of a segment
int32_t topk_run_length(int64_t total_queries);
int32_t topk_blocks(int32_t nq, int32_t run);
void launch_topk_batch(const TopkSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                       int32_t total_blocks, int32_t run, hipStream_t stream);
int32_t cell_adjust_blocks(int32_t cells);
// the batched composite's blocks of `threads` threads (64, 128 or 256), 8
// pixels of one output row each
int32_t composite_blocks(int32_t ow, int32_t oh, int32_t threads);
// 1 when n tiles of tw x th fit the composite's 32-bit tile offsets
int32_t composite_narrow(int32_t n_tiles, int32_t tw, int32_t th);
void launch_cell_adjust_batch(const CompSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                              int32_t total_blocks, hipStream_t stream);
void launch_composite_batch(const CompSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                            int32_t total_blocks, int32_t threads, hipStream_t stream);
// General chain: any cell order, pre-filled grid and usage (state in LDS).
size_t place_lds_bytes(int32_t cells, int32_t n);
// Raster chain: generate_mosaic's raster order from an empty grid, where
// only the up and left neighbours can be filled.  `staged`: the LDS also
// holds two groups of candidate rows (52 KB more), copied in ahead of the
// chain; the chain runs unstaged where that does not fit.
size_t place_raster_lds_bytes(int32_t grid_w, int32_t n, bool staged);
// One job passed by value (kernel argument), or n_jobs jobs from a device
// array (one wave each); `staged` as for place_raster_lds_bytes.
void launch_place_single(const PlaceJob& job, bool raster, bool staged, size_t lds_bytes,
                         hipStream_t stream);
// score_kind: kScoreSq or kScoreDist when every job scores alike, else
// kScoreMixed (each job's score_sqrt is read).
enum : int32_t { kScoreSq = 0, kScoreDist = 1, kScoreMixed = 2 };
void launch_place_batch(const PlaceJob* d_jobs, int32_t n_jobs, bool raster, bool staged,
                        int32_t score_kind, size_t lds_bytes, hipStream_t stream);
// Annealing variants: the grid in LDS with 16-bit material ids (libraries
// under 32,768 materials) or 32-bit ids, or in HBM with hashed LDS stamps
// (grids too large for LDS; sides below 65,536).
enum : int32_t { kAnnealNarrow = 0, kAnnealWide = 1, kAnnealHbm = 2 };
// The first variant that holds the job in max_lds bytes of LDS, -1 if none.
int32_t anneal_kind(int32_t cells, int32_t n, int32_t gw, int32_t gh, size_t max_lds);
size_t anneal_lds_bytes(int32_t cells, int32_t kind);
void launch_anneal_single(const AnnealJob& job, int32_t kind, size_t lds_bytes, hipStream_t stream);
void launch_anneal_batch(const AnnealJob* d_jobs, int32_t n_jobs, int32_t kind, size_t lds_bytes,
                         hipStream_t stream);
void launch_total_cost(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4,
                       const float* term, int32_t n, float* scratch, int32_t chunk, float* out,
                       hipStream_t stream);
void launch_swap_deltas(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4,
                        const float* term, int32_t n, const int32_t* pairs, int32_t p, float* out,
                        hipStream_t stream);
void launch_penalties(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4, float weight,
                      const int32_t* queries, int32_t nq, float* out, hipStream_t stream);
// n images of sw x sh (contiguous) -> dw x dh tiles; image m lands in tile
// slot dst_index[m] (nullptr: slot m).
void launch_resize(const uint8_t* src, int32_t n, int32_t sw, int32_t sh, int32_t dw, int32_t dh,
                   const ResizeTaps& h, const ResizeTaps& v, uint8_t* tmp, uint8_t* dst,
                   hipStream_t stream, const int32_t* dst_index = nullptr);
void launch_avg_rgb(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                    int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                    int32_t n_cells, uint8_t* out, hipStream_t stream);
void launch_cell_adjust(const int32_t* grid, int32_t cells, const uint8_t* tile_avg,
                        const uint8_t* target_avg, float strength, float4* out,
                        hipStream_t stream);
void launch_composite(const int32_t* grid, int32_t gw, int32_t gh, int32_t tw, int32_t th,
                      const uint8_t* tiles, int32_t n_tiles, const float4* adj, int32_t adjust,
                      uint8_t* out, hipStream_t stream);
void launch_adjust_pixels(float4 adj, const uint8_t* in, int64_t npx, uint8_t* out,
                          hipStream_t stream);
void launch_optimal_adjust(const uint8_t* tile, const uint8_t* target, int32_t n, float k,
                           float4* out, hipStream_t stream);

// Kernels whose dynamic LDS may exceed 64 KB get their limit raised once
// (to the per-block maximum minus the kernel's static LDS).
hipError_t raise_dynamic_lds(const void* kernel, int max_lds);
hipError_t configure_kernel_limits(int max_lds);
hipError_t place_configure(int max_lds);
hipError_t anneal_configure(int max_lds);
hipError_t resize_configure(int max_lds);

}  // namespace mh
