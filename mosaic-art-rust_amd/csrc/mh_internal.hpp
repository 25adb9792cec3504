// mh_internal.hpp -- host-side state shared by the C ABI translation units
// (mh_api.hip: single-job entry points; mh_batch.hip: batches and serving).
#pragma once

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mosaic_hip.h"
#include "mh_kernels.hpp"

namespace mh {
namespace host {

int fail(int code, const char* fmt, ...);

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP,        \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,     \
                        __LINE__);                                                           \
    } while (0)

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

enum Slot {
    S_IN0, S_IN1, S_IN2, S_OUT0, S_OUT1, S_OUT2, S_OUT3,
    S_CELL_LAB, S_CAND, S_GRID, S_USAGE, S_ORDER,
    S_TARGET_AVG, S_ADJ, S_STATS, S_ANNEAL_OUT,
    S_COUNT
};

struct TempTable {
    float* p = nullptr;
    int64_t len = 0;
};

// Resized materials (n x th x tw x 3) and their integer average RGB (n x 3).
struct TileSet {
    uint8_t* tiles = nullptr;
    uint8_t* avg = nullptr;
};

struct TapsBuf {
    DevBuf lo, cnt, k;
    int32_t ksize = 0;
    int32_t prec = 0;
};

struct PhaseMark {
    int32_t phase;
    hipEvent_t a, b;
};

}  // namespace host
}  // namespace mh

struct mh_context {
    int32_t device = 0;
    // libraries and batches still open on this context: destroying the
    // context releases their device memory and detaches them, so destroying
    // them afterwards only frees host memory
    std::vector<mh_library*> libs;
    std::vector<mh_batch*> batches;
    std::mutex dep_mu;
    // per-phase timing (HIP events on `stream`)
    uint32_t timing = 0;  // bit p: phase p is bracketed by events
    std::vector<hipEvent_t> event_pool;
    std::vector<mh::host::PhaseMark> marks;
    hipEvent_t open_event[MH_PHASES] = {};
    double phase_ms[MH_PHASES] = {};
    int64_t phase_n[MH_PHASES] = {};
    hipStream_t stream = nullptr;
    float* d_lut = nullptr;
    size_t max_lds = 65536;
    mh::host::DevBuf slots[mh::host::S_COUNT];
    // temperature tables keyed by the bits of (T0, decay); a table is never
    // overwritten while the context lives (queued kernels may read it), a
    // longer one replaces it and the old one is retired
    std::map<std::pair<uint32_t, uint32_t>, mh::host::TempTable> temps;
    std::vector<void*> retired;
    std::mutex temps_mu;
    // Lanczos taps keyed by (in, out)
    std::map<std::pair<int32_t, int32_t>, mh::host::TapsBuf> taps;
    std::mutex taps_mu;
    // last device job geometry and result buffers for mh_generate_device_result
    int32_t last_out_w = 0, last_out_h = 0;
    int32_t last_ran_anneal = 0;
    int64_t last_iters = 0;
    const int32_t* last_stats = nullptr;
    const mh::AnnealOut* last_aout = nullptr;
    // this device's copy of mh_generate_device_noalloc's scratch array
    // (hipGetSymbolAddress at context creation)
    uint8_t* noalloc_scratch = nullptr;
};

// Material pixels of one size: `count` RGB8 images of w x h stored back to
// back from byte `offset` of mh_library::d_pixels; image k of the group is
// material d_index[k] (d_index == nullptr: the group is materials 0..count-1).
struct PixelGroup {
    int32_t w = 0, h = 0, count = 0;
    int64_t offset = 0;
    int32_t* d_index = nullptr;
};

struct mh_library {
    mh_context* ctx = nullptr;
    int32_t n = 0;
    std::vector<float> h_lab;  // n x 3
    float4* d_lab4 = nullptr;
    float* d_L = nullptr;
    float* d_A = nullptr;
    float* d_B = nullptr;
    // the K-nearest scan's block layout (MatBlocks)
    float* d_pL = nullptr;
    float* d_pA = nullptr;
    float* d_pB = nullptr;
    int32_t* d_pidx = nullptr;
    float4* d_box = nullptr;
    int32_t nb = 0;
    float* d_sim = nullptr;    // n x n after build_similarities
    float* d_term = nullptr;   // n x n 1/(sim+1), built on first use by the cost kernels
    uint8_t* d_pixels = nullptr;
    std::vector<PixelGroup> groups;  // materials grouped by size (empty: Labs only)
    // lazily built device caches (term matrix, resized tiles per cell size)
    std::mutex cache_mu;
    std::map<std::pair<int32_t, int32_t>, mh::host::TileSet> tiles;
};

#define CHK(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_) return rc_;      \
    } while (0)

namespace mh {
namespace host {

int set_device(mh_context* ctx);
void phase_begin(mh_context* ctx, int phase);
void phase_end(mh_context* ctx, int phase);
int ensure(mh_context* ctx, DevBuf& b, size_t bytes);
bool finite3(const float* v, int64_t count);
int get_taps(mh_context* ctx, int32_t in_size, int32_t out_size, ResizeTaps* t);
// Device memory of a library / batch freed, the handle detached from its
// context (the context is being destroyed).
void library_release(mh_library* lib);
// The library's block layout for the K-nearest kernels.
MatBlocks library_blocks(const mh_library* lib);
void batch_release(mh_batch* b);
int get_temps(mh_context* ctx, float t0, float decay, int64_t len, const float** out);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
// Term matrix 1/(sim+1) of the library (nullptr above kTermMatrixMaxN
// materials); built once on `stream`, which is then synchronised so any
// stream may read it.
constexpr int32_t kTermMatrixMaxN = 16384;
int library_terms(const mh_library* lib, hipStream_t stream, const float** out);
// Resized tiles + average RGB for cell size tw x th, built once on `stream`
// (then synchronised).
int library_tiles(const mh_library* lib, int32_t tw, int32_t th, hipStream_t stream,
                  const uint8_t** tiles, const uint8_t** avg);
int sync(mh_context* ctx);
// Looks up the device's scratch array of the allocation-free entry point
// (ctx->noalloc_scratch); the context's device must be current.
int noalloc_scratch_init(mh_context* ctx);
int check_geometry(int32_t gw, int32_t gh);
int check_score_mode(int32_t mode);
void to_result(const AnnealOut& o, mh_anneal_result* r);

template <typename T>
int slot(mh_context* ctx, Slot s, size_t count, T** out) {
    int rc = ensure(ctx, ctx->slots[s], sizeof(T) * count);
    if (rc) return rc;
    *out = reinterpret_cast<T*>(ctx->slots[s].p);
    return MH_OK;
}

}  // namespace host
}  // namespace mh
