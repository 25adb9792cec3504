// mh_api.hip -- the C ABI (include/mosaic_hip.h) over the gfx950 kernels.
//
// Host entry points take caller-owned host buffers, stage them through
// context-owned device scratch (grow-only), enqueue the kernels on the
// context stream and synchronise.  mh_generate_device is the asynchronous,
// device-resident form used by the benchmark and by batching callers.
#include <algorithm>
#include <numeric>

#include "mh_internal.hpp"

using namespace mh;
using namespace mh::host;

namespace {
thread_local std::string g_err;
}  // namespace

namespace mh {
namespace host {

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int set_device(mh_context* ctx) {
    // a library or batch whose context was destroyed carries no context
    if (!ctx) return fail(MH_ERR_INVALID_ARGUMENT, "the handle's context was destroyed");
    HIPCHK(hipSetDevice(ctx->device));
    return MH_OK;
}

hipEvent_t take_event(mh_context* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Bracket the kernels of one phase with events (no-ops unless timing is on).
void phase_begin(mh_context* ctx, int phase) {
    if (!((ctx->timing >> phase) & 1u)) return;
    hipEvent_t e = take_event(ctx);
    if (e && hipEventRecord(e, ctx->stream) == hipSuccess) ctx->open_event[phase] = e;
}

void phase_end(mh_context* ctx, int phase) {
    if (!((ctx->timing >> phase) & 1u) || !ctx->open_event[phase]) return;
    hipEvent_t e = take_event(ctx);
    if (e && hipEventRecord(e, ctx->stream) == hipSuccess)
        ctx->marks.push_back(PhaseMark{phase, ctx->open_event[phase], e});
    ctx->open_event[phase] = nullptr;
}

int ensure(mh_context* ctx, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return MH_OK;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 4;
    HIPCHK(hipMalloc(&b.p, want));
    b.cap = want;
    (void)ctx;
    return MH_OK;
}


bool finite3(const float* v, int64_t count) {
    for (int64_t i = 0; i < count; ++i)
        if (!isfinite(v[i])) return false;
    return true;
}

// Lanczos3 fixed-point taps for one axis: the convolution coefficients of
// fast_image_resize 5.1.4's default Convolution(Lanczos3) for u8 pixels
// (src/main.rs:710-738 calls it; the crate follows Pillow-SIMD):
//  * scale = in/out, filter scale fs = max(scale, 1), radius r = 3 fs;
//  * output o centres on c = (o + 0.5) scale and reads source pixels
//    [max(floor(c - r), 0), min(ceil(c + r), in)), pixel x weighted
//    lanczos3((x - (c - 0.5)) * (1 / fs)), the weights divided by their sum;
//  * one precision p per axis: the largest p < 22 for which every weight
//    times 2^(p + 1), rounded, would still be below 2^15 -- or the first p
//    for which it no longer is; weights are rounded half away from zero to
//    p fraction bits (16-bit integers).
// The passes add (1 << (p - 1)) + sum(pixel * weight) in 32 bits and keep
// (sum >> p) clamped to 0..255.
double sinc(double x) {
    if (x == 0.0) return 1.0;
    x = x * M_PI;
    return sin(x) / x;
}

double lanczos3(double x) {
    if (x > -3.0 && x < 3.0) return sinc(x) * sinc(x / 3.0);
    return 0.0;
}

int get_taps(mh_context* ctx, int32_t in_size, int32_t out_size, ResizeTaps* t) {
    std::lock_guard<std::mutex> g(ctx->taps_mu);
    auto key = std::make_pair(in_size, out_size);
    auto it = ctx->taps.find(key);
    if (it == ctx->taps.end()) {
        const double scale = (double)in_size / (double)out_size;
        const double fs = scale < 1.0 ? 1.0 : scale;
        const double radius = 3.0 * fs;
        const double recip = 1.0 / fs;
        const int32_t ks = (int32_t)ceil(radius) * 2 + 1;
        std::vector<int32_t> lo(out_size), cnt(out_size), k((size_t)out_size * ks, 0);
        std::vector<double> w((size_t)out_size * ks, 0.0);
        double max_w = 0.0;
        for (int32_t o = 0; o < out_size; ++o) {
            const double in_center = ((double)o + 0.5) * scale;
            const double fmin = floor(in_center - radius), fmax = ceil(in_center + radius);
            const int32_t xmin = fmin > 0.0 ? (int32_t)fmin : 0;
            const int32_t xmax = fmax < (double)in_size ? (int32_t)fmax : in_size;
            const double center = in_center - 0.5;
            double* wo = &w[(size_t)o * ks];
            double ww = 0.0;
            for (int32_t x = xmin; x < xmax; ++x) {
                wo[x - xmin] = lanczos3(((double)x - center) * recip);
                ww += wo[x - xmin];
            }
            if (ww != 0.0)
                for (int32_t t = 0; t < xmax - xmin; ++t) wo[t] /= ww;
            for (int32_t t = 0; t < xmax - xmin; ++t) max_w = wo[t] > max_w ? wo[t] : max_w;
            lo[o] = xmin;
            cnt[o] = xmax - xmin;
        }
        int32_t prec = 0;
        for (int32_t p = 0; p < 22; ++p) {
            prec = p;
            if (std::round(max_w * (double)(1 << (p + 1))) >= 32768.0) break;
        }
        for (size_t i = 0; i < k.size(); ++i)
            k[i] = (int32_t)std::round(w[i] * (double)(1 << prec));  // half away from zero
        TapsBuf tb;
        tb.ksize = ks;
        tb.prec = prec;
        int rc;
        if ((rc = ensure(ctx, tb.lo, sizeof(int32_t) * out_size))) return rc;
        if ((rc = ensure(ctx, tb.cnt, sizeof(int32_t) * out_size))) return rc;
        if ((rc = ensure(ctx, tb.k, sizeof(int32_t) * k.size()))) return rc;
        HIPCHK(hipMemcpy(tb.lo.p, lo.data(), sizeof(int32_t) * out_size, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(tb.cnt.p, cnt.data(), sizeof(int32_t) * out_size, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(tb.k.p, k.data(), sizeof(int32_t) * k.size(), hipMemcpyHostToDevice));
        it = ctx->taps.emplace(key, tb).first;
    }
    t->lo = (const int32_t*)it->second.lo.p;
    t->cnt = (const int32_t*)it->second.cnt.p;
    t->k = (const int32_t*)it->second.k.p;
    t->ksize = it->second.ksize;
    t->prec = it->second.prec;
    return MH_OK;
}

// Temperature after k non-skipped proposals: T_0 = T0, T_{k+1} = T_k * decay
// in f32 (src/optimizer.rs:54, :102).
int get_temps(mh_context* ctx, float t0, float decay, int64_t len, const float** out) {
    if (len < 1) len = 1;
    std::lock_guard<std::mutex> g(ctx->temps_mu);
    uint32_t kt, kd;
    memcpy(&kt, &t0, 4);
    memcpy(&kd, &decay, 4);
    TempTable& tt = ctx->temps[std::make_pair(kt, kd)];
    if (tt.len < len) {
        std::vector<float> t((size_t)len);
        float v = t0;
        for (int64_t i = 0; i < len; ++i) {
            t[(size_t)i] = v;
            v = v * decay;
        }
        float* p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(float) * (size_t)len));
        hipError_t e = hipMemcpy(p, t.data(), sizeof(float) * (size_t)len, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(p);
            return fail(MH_ERR_HIP, "temperature table: %s", hipGetErrorString(e));
        }
        if (tt.p) ctx->retired.push_back(tt.p);
        tt.p = p;
        tt.len = len;
    }
    *out = tt.p;
    return MH_OK;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// Device-resident placement from cell Labs already on the device.  raster:
// the whole grid in raster order from an empty grid (generate_mosaic's case,
// the fast chain); otherwise `order` (or raster) over the given grid/usage.
// The candidate rows go to cand_buf (cells x kCandStride) when given, else
// to the context's scratch.
int run_place(mh_context* ctx, const mh_library* lib, const float* d_cell_lab, int32_t gw,
              int32_t gh, int32_t max_usage, float weight, int32_t score_mode, bool raster,
              const int32_t* d_order,
              int32_t n_order, int32_t* d_grid, uint32_t* d_usage, bool zero_usage,
              uint8_t* d_stage, float* d_dist, float* d_score, int32_t* d_stats,
              CandRec* cand_buf = nullptr) {
    const int32_t cells = gw * gh;
    // the raster chain stages its candidate rows in LDS when they fit
    const bool staged = raster && place_raster_lds_bytes(gw, lib->n, true) <= ctx->max_lds;
    size_t lds = raster ? place_raster_lds_bytes(gw, lib->n, staged)
                        : place_lds_bytes(cells, lib->n);
    // state beyond LDS: the general chain with grid and counts in HBM (it
    // reads the neighbours from the grid, so a raster run starts it empty)
    const bool hbm = lds > ctx->max_lds;
    if (hbm) {
        lds = 0;
        if (raster) fill_i32(d_grid, cells, kEmpty, ctx->stream);
    }
    CandRec* cand = cand_buf;
    int rc;
    if (!cand && (rc = slot(ctx, S_CAND, (size_t)cells * kCandStride, &cand))) return rc;
    const int32_t K = lib->n < kMaxCand ? lib->n : kMaxCand;
    phase_begin(ctx, MH_PHASE_TOPK);
    launch_topk(lib->d_L, lib->d_A, lib->d_B, lib->n, library_blocks(lib), d_cell_lab, cells, K,
                nullptr, nullptr, cand, kCandStride, ctx->stream);
    phase_end(ctx, MH_PHASE_TOPK);
    PlaceJob job;
    job.lab4 = lib->d_lab4;
    job.n = lib->n;
    job.cand = cand;
    job.order = d_order;
    job.n_order = n_order;
    job.grid = d_grid;
    job.usage = d_usage;
    job.zero_usage = zero_usage ? 1 : 0;
    job.grid_w = gw;
    job.grid_h = gh;
    job.max_usage = max_usage;
    job.weight = weight;
    job.score_sqrt = score_mode == MH_SCORE_DISTANCE ? 1 : 0;
    job.stage = d_stage;
    job.dist = d_dist;
    job.score = d_score;
    job.stats = d_stats;
    phase_begin(ctx, MH_PHASE_PLACE);
    launch_place_single(job, raster && !hbm, staged, lds, ctx->stream);
    phase_end(ctx, MH_PHASE_PLACE);
    HIPCHK(hipGetLastError());
    return MH_OK;
}

int library_terms(const mh_library* clib, hipStream_t stream, const float** out) {
    mh_library* lib = const_cast<mh_library*>(clib);
    *out = nullptr;
    if (lib->n == 0 || lib->n > kTermMatrixMaxN) return MH_OK;
    std::lock_guard<std::mutex> g(lib->cache_mu);
    if (!lib->d_term) {
        float* t = nullptr;
        HIPCHK(hipMalloc(&t, sizeof(float) * (size_t)lib->n * lib->n));
        launch_term_full(lib->d_lab4, lib->n, t, stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            hipFree(t);
            return fail(MH_ERR_HIP, "term matrix: %s", hipGetErrorString(e));
        }
        lib->d_term = t;
    }
    *out = lib->d_term;
    return MH_OK;
}

// Resized tiles + their average RGB for cell size tw x th, cached per
// library (the resize depends only on material and cell size).
int library_tiles(const mh_library* clib, int32_t tw, int32_t th, hipStream_t stream,
                  const uint8_t** tiles, const uint8_t** avg) {
    mh_library* lib = const_cast<mh_library*>(clib);
    mh_context* ctx = lib->ctx;
    std::lock_guard<std::mutex> g(lib->cache_mu);
    auto key = std::make_pair(tw, th);
    auto it = lib->tiles.find(key);
    if (it == lib->tiles.end()) {
        // Lanczos taps per source size; the two-pass fallback's intermediate
        // is sized for the largest group
        std::vector<std::pair<ResizeTaps, ResizeTaps>> taps(lib->groups.size());
        size_t tmp_bytes = 16;
        for (size_t g = 0; g < lib->groups.size(); ++g) {
            const PixelGroup& pg = lib->groups[g];
            CHK(get_taps(ctx, pg.w, tw, &taps[g].first));
            CHK(get_taps(ctx, pg.h, th, &taps[g].second));
            tmp_bytes = std::max(tmp_bytes, (size_t)pg.count * pg.h * tw * 3);
        }
        const size_t n = (size_t)(lib->n > 0 ? lib->n : 1);
        TileSet ts;
        uint8_t* tmp = nullptr;
        hipError_t e = hipMalloc(&ts.tiles, n * th * tw * 3);
        if (e == hipSuccess) e = hipMalloc(&ts.avg, n * 3);
        if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
        if (e == hipSuccess && lib->n > 0) {
            const bool timed = stream == ctx->stream;
            if (timed) phase_begin(ctx, MH_PHASE_RESIZE);
            // one launch per source size into disjoint tile slots (the
            // two-pass fallback's intermediate is reused in stream order)
            for (size_t g = 0; g < lib->groups.size(); ++g) {
                const PixelGroup& pg = lib->groups[g];
                launch_resize(lib->d_pixels + pg.offset, pg.count, pg.w, pg.h, tw, th,
                              taps[g].first, taps[g].second, tmp, ts.tiles, stream, pg.d_index);
            }
            if (timed) phase_end(ctx, MH_PHASE_RESIZE);
            launch_avg_rgb(ts.tiles, (int64_t)tw * 3, 1, (int64_t)tw * th * 3, 0, tw, th, lib->n, ts.avg,
                           stream);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (tmp) hipFree(tmp);
        if (e != hipSuccess) {
            if (ts.tiles) hipFree(ts.tiles);
            if (ts.avg) hipFree(ts.avg);
            return fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP,
                        "resized tiles %d x %d: %s", tw, th, hipGetErrorString(e));
        }
        it = lib->tiles.emplace(key, ts).first;
    }
    *tiles = it->second.tiles;
    *avg = it->second.avg;
    return MH_OK;
}

int library_tiles_cached(const mh_library* clib, int32_t tw, int32_t th, const uint8_t** tiles,
                         const uint8_t** avg) {
    mh_library* lib = const_cast<mh_library*>(clib);
    std::lock_guard<std::mutex> g(lib->cache_mu);
    auto it = lib->tiles.find(std::make_pair(tw, th));
    if (it == lib->tiles.end())
        return fail(MH_ERR_UNSUPPORTED,
                    "the library holds no resized tiles for %d x %d cells (an mh_generate, "
                    "mh_generate_device or mh_composite call at this cell size builds them)",
                    tw, th);
    *tiles = it->second.tiles;
    *avg = it->second.avg;
    return MH_OK;
}

// temps_ready: the temperature table, already on the device (else the
// context's cached table).  noalloc: allocate nothing -- the adjacency terms
// come from the Lab table (bit-identical to the library's term matrix) and
// temps_ready must be given whenever a proposal can read it.
int run_anneal(mh_context* ctx, const mh_library* lib, int32_t* d_grid, int32_t gw, int32_t gh,
               const mh_anneal_params* p, float* d_trace_delta, uint8_t* d_trace_accept,
               AnnealOut* d_out, const float* temps_ready, bool noalloc) {
    const int32_t cells = gw * gh;
    const int32_t kind = anneal_kind(cells, lib->n, gw, gh, ctx->max_lds);
    if (kind < 0)
        return fail(MH_ERR_UNSUPPORTED, "annealing a %d x %d grid: a side of 65,536 or more", gw,
                    gh);
    const size_t lds = anneal_lds_bytes(cells, kind);
    const float* temps = temps_ready;
    int rc;
    if (!p->greedy && !temps && !noalloc) {
        if ((rc = get_temps(ctx, p->initial_temperature, p->temperature_decay, p->max_iterations,
                            &temps)))
            return rc;
    }
    if (!p->greedy && !temps && p->max_iterations > 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "annealing without a temperature table");
    const float* term = nullptr;
    if (!noalloc && (rc = library_terms(lib, ctx->stream, &term))) return rc;
    AnnealJob job;
    job.lab4 = lib->d_lab4;
    job.term = term;
    job.n = lib->n;
    job.grid = d_grid;
    job.grid_w = gw;
    job.grid_h = gh;
    job.max_iter = p->max_iterations;
    job.temps = temps;
    job.seed = p->seed;
    job.greedy = p->greedy;
    job.trace_delta = d_trace_delta;
    job.trace_accept = d_trace_accept;
    job.result = d_out;
    phase_begin(ctx, MH_PHASE_ANNEAL);
    launch_anneal_single(job, kind, lds, ctx->stream);
    phase_end(ctx, MH_PHASE_ANNEAL);
    HIPCHK(hipGetLastError());
    return MH_OK;
}


// tavg_ready: the target cells' average RGB, already computed (by the cell
// Lab pass), or nullptr to compute it here.  adj_buf: room for the cells'
// adjustments (else the context's scratch).  cached_only: the library's
// resized tiles for this cell size must exist already (nothing is built).
int run_composite(mh_context* ctx, const mh_library* lib, const uint8_t* d_target, int32_t width,
                  int32_t height, const int32_t* d_grid, int32_t gw, int32_t gh, float strength,
                  uint8_t* d_out, const uint8_t* tavg_ready, float4* adj_buf, bool cached_only) {
    const int32_t tw = width / gw, th = height / gh;
    if (tw <= 0 || th <= 0) return MH_OK;
    const uint8_t *tiles, *tile_avg;
    int rc;
    if (cached_only) {
        if ((rc = library_tiles_cached(lib, tw, th, &tiles, &tile_avg))) return rc;
    } else if ((rc = library_tiles(lib, tw, th, ctx->stream, &tiles, &tile_avg))) {
        return rc;
    }
    phase_begin(ctx, MH_PHASE_ADJUST);
    const int32_t cells = gw * gh;
    float4* adj = nullptr;
    const int32_t adjust = strength > 0.0f ? 1 : 0;
    if (adjust) {
        const uint8_t* tavg = tavg_ready;
        if (!tavg) {
            uint8_t* t;
            if ((rc = slot(ctx, S_TARGET_AVG, (size_t)cells * 3, &t))) return rc;
            launch_avg_rgb(d_target, (int64_t)width * 3, gw, (int64_t)th * width * 3,
                           (int64_t)tw * 3, tw, th, cells, t, ctx->stream);
            tavg = t;
        }
        adj = adj_buf;
        if (!adj && (rc = slot(ctx, S_ADJ, (size_t)cells, &adj))) return rc;
        launch_cell_adjust(d_grid, cells, tile_avg, tavg, strength, adj, ctx->stream);
    }
    phase_end(ctx, MH_PHASE_ADJUST);
    phase_begin(ctx, MH_PHASE_COMPOSITE);
    launch_composite(d_grid, gw, gh, tw, th, tiles, lib->n, adj, adjust, d_out, ctx->stream);
    phase_end(ctx, MH_PHASE_COMPOSITE);
    HIPCHK(hipGetLastError());
    return MH_OK;
}

template <typename T>
int upload(mh_context* ctx, Slot s, const T* host, size_t count, T** dev) {
    int rc = slot(ctx, s, count, dev);
    if (rc) return rc;
    if (count) HIPCHK(hipMemcpyAsync(*dev, host, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
    return MH_OK;
}

template <typename T>
int download(mh_context* ctx, T* host, const T* dev, size_t count) {
    if (host && count)
        HIPCHK(hipMemcpyAsync(host, dev, sizeof(T) * count, hipMemcpyDeviceToHost, ctx->stream));
    return MH_OK;
}

int sync(mh_context* ctx) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}


int check_grid_values(const int32_t* grid, int64_t cells, int32_t n) {
    for (int64_t i = 0; i < cells; ++i)
        if (grid[i] < -1 || grid[i] >= n)
            return fail(MH_ERR_INVALID_ARGUMENT, "grid[%lld] = %d is not MH_EMPTY or a material index < %d",
                        (long long)i, grid[i], n);
    return MH_OK;
}

int check_geometry(int32_t gw, int32_t gh) {
    if (gw <= 0 || gh <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "grid %d x %d must be positive", gw, gh);
    if ((int64_t)gw * gh > (int64_t)1 << 30)
        return fail(MH_ERR_INVALID_ARGUMENT, "grid %d x %d too large", gw, gh);
    return MH_OK;
}

int check_score_mode(int32_t mode) {
    if (mode != MH_SCORE_SQDIST && mode != MH_SCORE_DISTANCE)
        return fail(MH_ERR_INVALID_ARGUMENT, "score_mode %d is neither MH_SCORE_SQDIST nor "
                    "MH_SCORE_DISTANCE", mode);
    return MH_OK;
}

void to_result(const AnnealOut& o, mh_anneal_result* r) {
    r->initial_cost = o.initial_cost;
    r->final_cost = o.final_cost;
    r->best_cost = o.best_cost;
    r->improved_count = o.improved;
    r->accepted_count = o.accepted;
    r->iterations = o.iterations;
    r->evaluated = o.evaluated;
    r->rounds = o.rounds;
}

void library_release(mh_library* lib) {
    mh_context* ctx = lib->ctx;
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (lib->d_lab4) hipFree(lib->d_lab4);
    if (lib->d_L) hipFree(lib->d_L);
    if (lib->d_A) hipFree(lib->d_A);
    if (lib->d_B) hipFree(lib->d_B);
    if (lib->d_pL) hipFree(lib->d_pL);
    if (lib->d_pA) hipFree(lib->d_pA);
    if (lib->d_pB) hipFree(lib->d_pB);
    if (lib->d_pidx) hipFree(lib->d_pidx);
    if (lib->d_box) hipFree(lib->d_box);
    lib->d_pL = lib->d_pA = lib->d_pB = nullptr;
    lib->d_pidx = nullptr;
    lib->d_box = nullptr;
    if (lib->d_sim) hipFree(lib->d_sim);
    if (lib->d_term) hipFree(lib->d_term);
    if (lib->d_pixels) hipFree(lib->d_pixels);
    lib->d_lab4 = nullptr;
    lib->d_L = lib->d_A = lib->d_B = lib->d_sim = lib->d_term = nullptr;
    lib->d_pixels = nullptr;
    for (PixelGroup& g : lib->groups) {
        if (g.d_index) hipFree(g.d_index);
        g.d_index = nullptr;
    }
    for (auto& kv : lib->tiles) {
        if (kv.second.tiles) hipFree(kv.second.tiles);
        if (kv.second.avg) hipFree(kv.second.avg);
    }
    lib->tiles.clear();
    lib->ctx = nullptr;
}

MatBlocks library_blocks(const mh_library* lib) {
    return MatBlocks{lib->d_pL, lib->d_pA, lib->d_pB, lib->d_pidx, lib->d_box, lib->nb};
}

}  // namespace host
}  // namespace mh

// =================================================================== ABI
extern "C" {

const char* mh_last_error(void) { return g_err.c_str(); }

int32_t mh_abi_version(void) { return 2; }

int mh_context_create(int32_t device, mh_context** out) {
    if (!out) return fail(MH_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MH_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return fail(MH_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, count);
    mh_context* ctx = new mh_context();
    ctx->device = device;
    if (set_device(ctx)) {
        delete ctx;
        return MH_ERR_HIP;
    }
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(MH_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        lds > 0)
        ctx->max_lds = (size_t)lds;
    float lut[256];
    for (int v = 0; v < 256; ++v) {
        const float c = (float)v / 255.0f;
        lut[v] = c <= 0.04045f ? c / 12.92f : (float)pow(((double)c + 0.055) / 1.055, 2.4);
    }
    e = hipMalloc(&ctx->d_lut, sizeof(lut));
    if (e == hipSuccess) e = hipMemcpy(ctx->d_lut, lut, sizeof(lut), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mh_context_destroy(ctx);
        return fail(MH_ERR_HIP, "context init: %s", hipGetErrorString(e));
    }
    e = configure_kernel_limits((int)ctx->max_lds);
    if (e != hipSuccess) {
        mh_context_destroy(ctx);
        return fail(MH_ERR_HIP, "raising the dynamic LDS limit to %zu B: %s", ctx->max_lds,
                    hipGetErrorString(e));
    }
    // this device's copy of the allocation-free entry point's scratch array
    if (const int rc = noalloc_scratch_init(ctx)) {
        mh_context_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return MH_OK;
}

void mh_context_destroy(mh_context* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    {
        // handles still open on this context lose their device memory now;
        // destroying them later frees only their host memory
        std::vector<mh_batch*> batches;
        std::vector<mh_library*> libs;
        {
            std::lock_guard<std::mutex> g(ctx->dep_mu);
            batches.swap(ctx->batches);
            libs.swap(ctx->libs);
        }
        for (mh_batch* b : batches) batch_release(b);
        for (mh_library* lib : libs) library_release(lib);
    }
    for (auto& b : ctx->slots)
        if (b.p) hipFree(b.p);
    for (auto& kv : ctx->temps)
        if (kv.second.p) hipFree(kv.second.p);
    for (void* p : ctx->retired) hipFree(p);
    for (auto& kv : ctx->taps) {
        if (kv.second.lo.p) hipFree(kv.second.lo.p);
        if (kv.second.cnt.p) hipFree(kv.second.cnt.p);
        if (kv.second.k.p) hipFree(kv.second.k.p);
    }
    if (ctx->d_lut) hipFree(ctx->d_lut);
    for (const PhaseMark& m : ctx->marks) {
        hipEventDestroy(m.a);
        hipEventDestroy(m.b);
    }
    for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
    for (hipEvent_t e : ctx->open_event)
        if (e) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int mh_context_synchronize(mh_context* ctx) {
    if (!ctx) return fail(MH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    CHK(set_device(ctx));
    return sync(ctx);
}

void* mh_context_stream(mh_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mh_context_set_timing(mh_context* ctx, int32_t enabled) {
    if (!ctx) return fail(MH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->timing = enabled != 0 ? (1u << MH_PHASES) - 1u : 0u;
    return MH_OK;
}

int mh_context_set_timing_phases(mh_context* ctx, uint32_t phase_mask) {
    if (!ctx) return fail(MH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->timing = phase_mask & ((1u << MH_PHASES) - 1u);
    return MH_OK;
}

int mh_context_timing(mh_context* ctx, double ms[MH_PHASES], int64_t launches[MH_PHASES]) {
    if (!ctx || !ms || !launches) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    CHK(set_device(ctx));
    CHK(sync(ctx));
    for (const PhaseMark& m : ctx->marks) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, m.a, m.b));
        ctx->phase_ms[m.phase] += t;
        ctx->phase_n[m.phase] += 1;
        ctx->event_pool.push_back(m.a);
        ctx->event_pool.push_back(m.b);
    }
    ctx->marks.clear();
    for (int i = 0; i < MH_PHASES; ++i) {
        ms[i] = ctx->phase_ms[i];
        launches[i] = ctx->phase_n[i];
        ctx->phase_ms[i] = 0.0;
        ctx->phase_n[i] = 0;
    }
    return MH_OK;
}

const char* mh_phase_name(int32_t phase) {
    static const char* names[MH_PHASES] = {"cell_lab", "topk", "place", "anneal",
                                           "resize", "adjust", "composite", "sim_matrix"};
    return (phase >= 0 && phase < MH_PHASES) ? names[phase] : "unknown";
}

int mh_device_alloc(mh_context* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    CHK(set_device(ctx));
    HIPCHK(hipMalloc(out, bytes ? bytes : 16));
    return MH_OK;
}

int mh_device_free(mh_context* ctx, void* ptr) {
    if (!ctx) return fail(MH_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ptr) return MH_OK;
    CHK(set_device(ctx));
    CHK(sync(ctx));
    HIPCHK(hipFree(ptr));
    return MH_OK;
}

int mh_copy_to_device(mh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    CHK(set_device(ctx));
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return sync(ctx);
}

__global__ void device_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                   int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void device_copy_tail_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                        int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int mh_copy_on_device(mh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (bytes == 0) return MH_OK;
    CHK(set_device(ctx));
    const bool aligned = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
    const int64_t n16 = aligned ? (int64_t)(bytes / 16) : 0;
    if (n16 > 0) {
        const int64_t blocks = (n16 + 255) / 256 < 8192 ? (n16 + 255) / 256 : 8192;
        hipLaunchKernelGGL(device_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           (const uint4*)src, (uint4*)dst, n16);
    }
    const int64_t done = n16 * 16, rest = (int64_t)bytes - done;
    if (rest > 0)
        hipLaunchKernelGGL(device_copy_tail_kernel, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0,
                           ctx->stream, (const uint8_t*)src + done, (uint8_t*)dst + done, rest);
    HIPCHK(hipGetLastError());
    return MH_OK;
}

int mh_copy_2d_on_device(mh_context* ctx, void* dst, size_t dst_pitch, const void* src,
                         size_t src_pitch, size_t row_bytes, size_t rows) {
    if (!ctx || (row_bytes && rows && (!dst || !src)))
        return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    if (row_bytes > dst_pitch || row_bytes > src_pitch)
        return fail(MH_ERR_INVALID_ARGUMENT, "row_bytes exceeds a pitch");
    if (row_bytes == 0 || rows == 0) return MH_OK;
    CHK(set_device(ctx));
    HIPCHK(hipMemcpy2DAsync(dst, dst_pitch, src, src_pitch, row_bytes, rows, hipMemcpyDeviceToDevice,
                            ctx->stream));
    return MH_OK;
}

int mh_copy_to_host(mh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    CHK(set_device(ctx));
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return sync(ctx);
}

// ---------------------------------------------------------------- library
static int library_alloc(mh_context* ctx, int32_t n, mh_library** out) {
    mh_library* lib = new mh_library();
    lib->ctx = ctx;
    {
        std::lock_guard<std::mutex> g(ctx->dep_mu);
        ctx->libs.push_back(lib);
    }
    lib->n = n;
    lib->h_lab.assign((size_t)n * 3, 0.0f);
    const size_t m = n > 0 ? (size_t)n : 1;
    hipError_t e = hipMalloc(&lib->d_lab4, sizeof(float4) * m);
    if (e == hipSuccess) e = hipMalloc(&lib->d_L, sizeof(float) * m);
    if (e == hipSuccess) e = hipMalloc(&lib->d_A, sizeof(float) * m);
    if (e == hipSuccess) e = hipMalloc(&lib->d_B, sizeof(float) * m);
    if (e != hipSuccess) {
        mh_library_destroy(lib);
        return fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP, "library alloc: %s",
                    hipGetErrorString(e));
    }
    *out = lib;
    return MH_OK;
}

// Blocks of 64 materials for the K-nearest scan (MatBlocks): ranges of the
// permutation are split at a multiple of 64 (about half way) on the Lab axis
// of widest extent, until a range is one block; ties on the axis value go
// by material index, so the layout is a function of the Labs alone.  Only
// the last block can be partial.
static int library_build_blocks(mh_library* lib) {
    const int32_t n = lib->n;
    const int32_t nb = (n + 63) / 64;
    const float* lab = lib->h_lab.data();
    std::vector<int32_t> perm((size_t)n);
    for (int32_t i = 0; i < n; ++i) perm[(size_t)i] = i;
    std::vector<std::pair<int32_t, int32_t>> todo{{0, n}};
    while (!todo.empty()) {
        const auto [lo, hi] = todo.back();
        todo.pop_back();
        if (hi - lo <= 64) continue;
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t i = lo; i < hi; ++i)
            for (int a = 0; a < 3; ++a) {
                const float v = lab[3 * (size_t)perm[(size_t)i] + a];
                mn[a] = std::min(mn[a], v);
                mx[a] = std::max(mx[a], v);
            }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
        const int32_t blocks = (hi - lo + 63) / 64;
        const int32_t mid = lo + 64 * ((blocks + 1) / 2);
        std::nth_element(perm.begin() + lo, perm.begin() + mid, perm.begin() + hi,
                         [&](int32_t x, int32_t y) {
                             const float u = lab[3 * (size_t)x + ax], v = lab[3 * (size_t)y + ax];
                             return u < v || (u == v && x < y);
                         });
        todo.push_back({lo, mid});
        todo.push_back({mid, hi});
    }
    const size_t slots = (size_t)(nb > 0 ? nb : 1) * 64;
    std::vector<float> pl(slots, 0.0f), pa(slots, 0.0f), pb(slots, 0.0f);
    std::vector<int32_t> pidx(slots, -1);
    std::vector<float4> box((size_t)(nb > 0 ? nb : 1) * 2, make_float4(0, 0, 0, 0));
    for (int32_t b = 0; b < nb; ++b) {
        float4 lo4 = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
        float4 hi4 = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        for (int32_t s = 64 * b; s < std::min(n, 64 * b + 64); ++s) {
            const int32_t m = perm[(size_t)s];
            const float* v = lab + 3 * (size_t)m;
            pl[(size_t)s] = v[0];
            pa[(size_t)s] = v[1];
            pb[(size_t)s] = v[2];
            pidx[(size_t)s] = m;
            lo4.x = std::min(lo4.x, v[0]);
            lo4.y = std::min(lo4.y, v[1]);
            lo4.z = std::min(lo4.z, v[2]);
            hi4.x = std::max(hi4.x, v[0]);
            hi4.y = std::max(hi4.y, v[1]);
            hi4.z = std::max(hi4.z, v[2]);
        }
        box[2 * (size_t)b] = lo4;
        box[2 * (size_t)b + 1] = hi4;
    }
    hipError_t e = hipMalloc(&lib->d_pL, sizeof(float) * slots);
    if (e == hipSuccess) e = hipMalloc(&lib->d_pA, sizeof(float) * slots);
    if (e == hipSuccess) e = hipMalloc(&lib->d_pB, sizeof(float) * slots);
    if (e == hipSuccess) e = hipMalloc(&lib->d_pidx, sizeof(int32_t) * slots);
    if (e == hipSuccess) e = hipMalloc(&lib->d_box, sizeof(float4) * box.size());
    if (e == hipSuccess) e = hipMemcpy(lib->d_pL, pl.data(), sizeof(float) * slots, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(lib->d_pA, pa.data(), sizeof(float) * slots, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(lib->d_pB, pb.data(), sizeof(float) * slots, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(lib->d_pidx, pidx.data(), sizeof(int32_t) * slots, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(lib->d_box, box.data(), sizeof(float4) * box.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP,
                    "material blocks: %s", hipGetErrorString(e));
    lib->nb = nb;
    return MH_OK;
}

static int library_upload_lab(mh_library* lib) {
    const int32_t n = lib->n;
    if (n == 0) return MH_OK;
    std::vector<float4> l4((size_t)n);
    std::vector<float> L((size_t)n), A((size_t)n), B((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        L[i] = lib->h_lab[3 * i];
        A[i] = lib->h_lab[3 * i + 1];
        B[i] = lib->h_lab[3 * i + 2];
        l4[i] = make_float4(L[i], A[i], B[i], 0.0f);
    }
    HIPCHK(hipMemcpy(lib->d_lab4, l4.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(lib->d_L, L.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(lib->d_A, A.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(lib->d_B, B.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    return library_build_blocks(lib);
}

int mh_library_from_lab(mh_context* ctx, const float* lab, int32_t n, mh_library** out) {
    if (!ctx || !out || n < 0 || (n > 0 && !lab))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_library_from_lab: bad arguments");
    if (!finite3(lab, (int64_t)n * 3))
        return fail(MH_ERR_INVALID_ARGUMENT, "material Lab values must be finite");
    CHK(set_device(ctx));
    mh_library* lib = nullptr;
    CHK(library_alloc(ctx, n, &lib));
    if (n) memcpy(lib->h_lab.data(), lab, sizeof(float) * 3 * (size_t)n);
    int rc = library_upload_lab(lib);
    if (rc) {
        mh_library_destroy(lib);
        return rc;
    }
    *out = lib;
    return MH_OK;
}

int mh_library_from_pixels(mh_context* ctx, const uint8_t* rgb, int32_t n, int32_t w, int32_t h,
                           mh_library** out) {
    if (!ctx || !out || n < 0 || w <= 0 || h <= 0 || (n > 0 && !rgb))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_library_from_pixels: bad arguments");
    CHK(set_device(ctx));
    mh_library* lib = nullptr;
    CHK(library_alloc(ctx, n, &lib));
    const size_t bytes = (size_t)n * w * h * 3;
    if (n > 0) lib->groups.push_back(PixelGroup{w, h, n, 0, nullptr});
    int rc = MH_OK;
    hipError_t e = hipMalloc(&lib->d_pixels, bytes ? bytes : 16);
    if (e == hipSuccess && bytes) e = hipMemcpy(lib->d_pixels, rgb, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mh_library_destroy(lib);
        return fail(MH_ERR_HIP, "library pixels: %s", hipGetErrorString(e));
    }
    if (n > 0) {
        float* d_lab = nullptr;
        rc = slot(ctx, S_OUT0, (size_t)n * 3, &d_lab);
        if (!rc) {
            launch_cell_lab(lib->d_pixels, (int64_t)w * 3, 1, (int64_t)w * h * 3, 0, w, h, n, ctx->d_lut,
                            d_lab, ctx->stream);
            e = hipMemcpyAsync(lib->h_lab.data(), d_lab, sizeof(float) * 3 * n, hipMemcpyDeviceToHost,
                               ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = fail(MH_ERR_HIP, "material Lab: %s", hipGetErrorString(e));
        }
        if (!rc) rc = library_upload_lab(lib);
    }
    if (rc) {
        mh_library_destroy(lib);
        return rc;
    }
    *out = lib;
    return MH_OK;
}

// ------------------------------------------- material ingest, mixed sizes
// Images of different sizes are grouped by size (in order of first
// appearance) and packed group after group, so one Lab segment and one
// resize launch serve a whole group.
constexpr int64_t kMaxImagePixels = (int64_t)1 << 28;

struct RaggedPlan {
    std::vector<PixelGroup> groups;    // offsets into the packed buffer
    std::vector<int32_t> group_first;  // first packed slot of each group
    std::vector<int32_t> slot_image;   // packed slot -> image index
    size_t bytes = 0;
};

static int ragged_plan(const uint8_t* rgb, const int64_t* offsets, const int32_t* widths,
                       const int32_t* heights, int32_t n, RaggedPlan* p) {
    if (n < 0 || (n > 0 && (!rgb || !offsets || !widths || !heights)))
        return fail(MH_ERR_INVALID_ARGUMENT, "images: NULL argument or n < 0");
    std::map<std::pair<int32_t, int32_t>, int32_t> group_of;
    std::vector<std::vector<int32_t>> members;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t w = widths[i], h = heights[i];
        if (w <= 0 || h <= 0 || (int64_t)w * h > kMaxImagePixels)
            return fail(MH_ERR_INVALID_ARGUMENT, "image %d: size %d x %d outside 1..2^28 pixels", i,
                        w, h);
        if (offsets[i] < 0) return fail(MH_ERR_INVALID_ARGUMENT, "image %d: negative offset", i);
        auto it = group_of.find(std::make_pair(w, h));
        if (it == group_of.end()) {
            it = group_of.emplace(std::make_pair(w, h), (int32_t)members.size()).first;
            members.emplace_back();
        }
        members[it->second].push_back(i);
    }
    size_t off = 0;
    for (const std::vector<int32_t>& m : members) {
        PixelGroup g;
        g.w = widths[m[0]];
        g.h = heights[m[0]];
        g.count = (int32_t)m.size();
        g.offset = (int64_t)off;
        p->group_first.push_back((int32_t)p->slot_image.size());
        p->slot_image.insert(p->slot_image.end(), m.begin(), m.end());
        off += (size_t)g.count * g.w * g.h * 3;
        p->groups.push_back(g);
    }
    p->bytes = off;
    return MH_OK;
}

// Packs the images in slot order and copies them to d_dst (p.bytes bytes).
static int ragged_upload(mh_context* ctx, const uint8_t* rgb, const int64_t* offsets,
                         const RaggedPlan& p, uint8_t* d_dst) {
    if (p.bytes == 0) return MH_OK;
    std::vector<uint8_t> pack(p.bytes);
    for (size_t g = 0; g < p.groups.size(); ++g) {
        const PixelGroup& pg = p.groups[g];
        const size_t img = (size_t)pg.w * pg.h * 3;
        for (int32_t k = 0; k < pg.count; ++k)
            memcpy(pack.data() + pg.offset + k * img, rgb + offsets[p.slot_image[p.group_first[g] + k]],
                   img);
    }
    HIPCHK(hipMemcpyAsync(d_dst, pack.data(), p.bytes, hipMemcpyHostToDevice, ctx->stream));
    return sync(ctx);  // `pack` is pageable and local
}

// calculate_average_lab of every packed image (one batched launch, one
// segment per size group); lab_out = n x 3 in image order.
static int ragged_lab(mh_context* ctx, const uint8_t* d_pixels, const RaggedPlan& p,
                      float* lab_out) {
    const int32_t n = (int32_t)p.slot_image.size();
    if (n == 0) return MH_OK;
    const int32_t n_segs = (int32_t)p.groups.size();
    std::vector<CellSeg> segs((size_t)n_segs);
    std::vector<int32_t> starts((size_t)n_segs + 1);
    float* d_lab;
    CHK(slot(ctx, S_OUT0, (size_t)n * 3, &d_lab));
    int32_t blocks = 0;
    const int32_t cpb = cell_lab_cells_per_block(n);
    for (int32_t g = 0; g < n_segs; ++g) {
        const PixelGroup& pg = p.groups[g];
        segs[g] = CellSeg{d_pixels + pg.offset, (int64_t)pg.w * 3, (int64_t)pg.w * pg.h * 3, 0, 1,
                          pg.w, pg.h, pg.count, d_lab + 3 * (int64_t)p.group_first[g], nullptr};
        starts[g] = blocks;
        blocks += cell_lab_blocks(pg.count, cpb);
    }
    starts[n_segs] = blocks;
    CellSeg* d_segs;
    int32_t* d_starts;
    CHK(upload(ctx, S_IN1, segs.data(), segs.size(), &d_segs));
    CHK(upload(ctx, S_IN2, starts.data(), starts.size(), &d_starts));
    phase_begin(ctx, MH_PHASE_CELL_LAB);
    launch_cell_lab_batch(d_segs, d_starts, n_segs, blocks, cpb, ctx->d_lut, ctx->stream);
    phase_end(ctx, MH_PHASE_CELL_LAB);
    HIPCHK(hipGetLastError());
    std::vector<float> packed((size_t)n * 3);
    CHK(download(ctx, packed.data(), d_lab, packed.size()));
    CHK(sync(ctx));
    for (int32_t k = 0; k < n; ++k)
        memcpy(lab_out + 3 * (size_t)p.slot_image[k], packed.data() + 3 * (size_t)k, 3 * sizeof(float));
    return MH_OK;
}

int mh_average_lab_many(mh_context* ctx, const uint8_t* rgb, const int64_t* offsets,
                        const int32_t* widths, const int32_t* heights, int32_t n, float* out) {
    if (!ctx || (n > 0 && !out)) return fail(MH_ERR_INVALID_ARGUMENT, "mh_average_lab_many: NULL argument");
    RaggedPlan p;
    CHK(ragged_plan(rgb, offsets, widths, heights, n, &p));
    if (n == 0) return MH_OK;
    CHK(set_device(ctx));
    uint8_t* d_in;
    CHK(slot(ctx, S_IN0, p.bytes, &d_in));
    CHK(ragged_upload(ctx, rgb, offsets, p, d_in));
    return ragged_lab(ctx, d_in, p, out);
}

int mh_library_from_images(mh_context* ctx, const uint8_t* rgb, const int64_t* offsets,
                           const int32_t* widths, const int32_t* heights, int32_t n,
                           mh_library** out) {
    if (!ctx || !out) return fail(MH_ERR_INVALID_ARGUMENT, "mh_library_from_images: NULL argument");
    RaggedPlan p;
    CHK(ragged_plan(rgb, offsets, widths, heights, n, &p));
    CHK(set_device(ctx));
    mh_library* lib = nullptr;
    CHK(library_alloc(ctx, n, &lib));
    int rc = MH_OK;
    hipError_t e = hipMalloc(&lib->d_pixels, p.bytes ? p.bytes : 16);
    if (e != hipSuccess)
        rc = fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP, "library pixels: %s",
                  hipGetErrorString(e));
    for (size_t g = 0; g < p.groups.size() && !rc; ++g) {
        PixelGroup pg = p.groups[g];
        const int32_t* idx = p.slot_image.data() + p.group_first[g];
        bool identity = p.group_first[g] == 0;
        for (int32_t k = 0; k < pg.count && identity; ++k) identity = idx[k] == k;
        if (!identity) {
            e = hipMalloc(&pg.d_index, sizeof(int32_t) * pg.count);
            if (e == hipSuccess)
                e = hipMemcpy(pg.d_index, idx, sizeof(int32_t) * pg.count, hipMemcpyHostToDevice);
            if (e != hipSuccess)
                rc = fail(MH_ERR_HIP, "library index: %s", hipGetErrorString(e));
        }
        lib->groups.push_back(pg);  // owned by the library even on failure
    }
    if (!rc) rc = ragged_upload(ctx, rgb, offsets, p, lib->d_pixels);
    if (!rc) rc = ragged_lab(ctx, lib->d_pixels, p, lib->h_lab.data());
    if (!rc) rc = library_upload_lab(lib);
    if (rc) {
        mh_library_destroy(lib);
        return rc;
    }
    *out = lib;
    return MH_OK;
}

void mh_library_destroy(mh_library* lib) {
    if (!lib) return;
    if (mh_context* ctx = lib->ctx) {
        {
            std::lock_guard<std::mutex> g(ctx->dep_mu);
            for (size_t i = 0; i < ctx->libs.size(); ++i)
                if (ctx->libs[i] == lib) {
                    ctx->libs[i] = ctx->libs.back();
                    ctx->libs.pop_back();
                    break;
                }
        }
        library_release(lib);
    }
    delete lib;
}

int32_t mh_library_size(const mh_library* lib) { return lib ? lib->n : -1; }

int mh_library_lab(const mh_library* lib, float* out) {
    if (!lib || (!out && lib->n > 0)) return fail(MH_ERR_INVALID_ARGUMENT, "bad arguments");
    if (lib->n) memcpy(out, lib->h_lab.data(), sizeof(float) * 3 * (size_t)lib->n);
    return MH_OK;
}

int mh_library_build_similarities(mh_library* lib) {
    if (!lib) return fail(MH_ERR_INVALID_ARGUMENT, "lib is NULL");
    mh_context* ctx = lib->ctx;
    CHK(set_device(ctx));
    if (lib->n == 0) return MH_OK;
    if (!lib->d_sim) HIPCHK(hipMalloc(&lib->d_sim, sizeof(float) * (size_t)lib->n * lib->n));
    phase_begin(ctx, MH_PHASE_SIM);
    launch_sim_full(lib->d_lab4, lib->n, lib->d_sim, ctx->stream);
    phase_end(ctx, MH_PHASE_SIM);
    HIPCHK(hipGetLastError());
    return sync(ctx);
}

int mh_library_similarities(mh_library* lib, float* out, int32_t packed) {
    if (!lib || (!out && lib->n > 1)) return fail(MH_ERR_INVALID_ARGUMENT, "bad arguments");
    mh_context* ctx = lib->ctx;
    CHK(set_device(ctx));
    const int64_t n = lib->n;
    if (n == 0) return MH_OK;
    if (packed) {
        const int64_t cnt = n * (n - 1) / 2;
        if (cnt == 0) return MH_OK;
        float* d = nullptr;
        CHK(slot(ctx, S_OUT0, (size_t)cnt, &d));
        launch_sim_packed(lib->d_lab4, lib->n, d, ctx->stream);
        HIPCHK(hipGetLastError());
        CHK(download(ctx, out, d, (size_t)cnt));
        return sync(ctx);
    }
    if (!lib->d_sim) CHK(mh_library_build_similarities(lib));
    CHK(download(ctx, out, lib->d_sim, (size_t)(n * n)));
    return sync(ctx);
}

int mh_library_get_similarity(mh_library* lib, int32_t i, int32_t j, float* out) {
    if (!lib || !out) return fail(MH_ERR_INVALID_ARGUMENT, "bad arguments");
    if (i < 0 || j < 0 || i >= lib->n || j >= lib->n)
        return fail(MH_ERR_NOT_FOUND, "material index (%d, %d) not in library of %d", i, j, lib->n);
    if (i == j) {
        *out = 0.0f;
        return MH_OK;
    }
    mh_context* ctx = lib->ctx;
    CHK(set_device(ctx));
    if (!lib->d_sim) CHK(mh_library_build_similarities(lib));
    const int32_t a = i < j ? i : j, b = i < j ? j : i;
    HIPCHK(hipMemcpy(out, lib->d_sim + (size_t)a * lib->n + b, sizeof(float), hipMemcpyDeviceToHost));
    return MH_OK;
}

// -------------------------------------------------------------- a1: Lab
int mh_average_lab(mh_context* ctx, const uint8_t* rgb, int32_t w, int32_t h, float out[3]) {
    if (!ctx || !rgb || !out || w <= 0 || h <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_average_lab: bad arguments");
    CHK(set_device(ctx));
    uint8_t* d_in;
    float* d_out;
    CHK(upload(ctx, S_IN0, rgb, (size_t)w * h * 3, &d_in));
    CHK(slot(ctx, S_OUT0, 3, &d_out));
    launch_cell_lab(d_in, (int64_t)w * 3, 1, 0, 0, w, h, 1, ctx->d_lut, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, 3));
    return sync(ctx);
}

int mh_cell_lab_means(mh_context* ctx, const uint8_t* rgb, int32_t width, int32_t height,
                      int32_t gw, int32_t gh, float* out) {
    if (!ctx || !rgb || !out) return fail(MH_ERR_INVALID_ARGUMENT, "mh_cell_lab_means: NULL argument");
    CHK(check_geometry(gw, gh));
    const int32_t tw = width / gw, th = height / gh;
    if (tw <= 0 || th <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "image %d x %d smaller than grid %d x %d", width, height,
                    gw, gh);
    CHK(set_device(ctx));
    uint8_t* d_in;
    float* d_out;
    const int32_t cells = gw * gh;
    CHK(upload(ctx, S_IN0, rgb, (size_t)width * height * 3, &d_in));
    CHK(slot(ctx, S_OUT0, (size_t)cells * 3, &d_out));
    launch_cell_lab(d_in, (int64_t)width * 3, gw, (int64_t)th * width * 3, (int64_t)tw * 3, tw, th,
                    cells, ctx->d_lut, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, (size_t)cells * 3));
    return sync(ctx);
}

// ------------------------------------------------------------- a3: top-K
int mh_nearest_n(mh_context* ctx, const mh_library* lib, const float* query, int32_t nq, int32_t k,
                 int32_t* idx, float* sqdist) {
    if (!ctx || !lib || nq < 0 || k < 0 || k > 128 || (nq > 0 && (!query || !idx || !sqdist)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_nearest_n: bad arguments (k must be <= 128)");
    if (!finite3(query, (int64_t)nq * 3))
        return fail(MH_ERR_INVALID_ARGUMENT, "query Lab values must be finite");
    const int32_t kk = k < lib->n ? k : lib->n;
    if (nq == 0 || kk == 0) return MH_OK;
    CHK(set_device(ctx));
    float* d_q;
    int32_t* d_i;
    float* d_d;
    CHK(upload(ctx, S_IN0, query, (size_t)nq * 3, &d_q));
    CHK(slot(ctx, S_OUT0, (size_t)nq * kk, &d_i));
    CHK(slot(ctx, S_OUT1, (size_t)nq * kk, &d_d));
    launch_topk(lib->d_L, lib->d_A, lib->d_B, lib->n, library_blocks(lib), d_q, nq, kk, d_i, d_d,
                nullptr, kk, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, idx, d_i, (size_t)nq * kk));
    CHK(download(ctx, sqdist, d_d, (size_t)nq * kk));
    return sync(ctx);
}

// ----------------------------------------------------------- a4-a9: place
int mh_place(mh_context* ctx, const mh_library* lib, const float* cell_lab,
             const mh_place_params* p, const int32_t* order, int32_t n_order, int32_t* grid,
             uint32_t* usage, uint8_t* stage, float* dist, float* score, mh_place_stats* stats) {
    if (!ctx || !lib || !p || !cell_lab || !grid || (!usage && lib->n > 0))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_place: NULL argument");
    if (!isfinite(p->weight)) return fail(MH_ERR_INVALID_ARGUMENT, "weight must be finite");
    CHK(check_geometry(p->grid_w, p->grid_h));
    if (p->max_usage < 0) return fail(MH_ERR_INVALID_ARGUMENT, "max_usage must be >= 0");
    CHK(check_score_mode(p->score_mode));
    const int32_t cells = p->grid_w * p->grid_h;
    if (!finite3(cell_lab, (int64_t)cells * 3))
        return fail(MH_ERR_INVALID_ARGUMENT, "cell Lab values must be finite");
    if (order) {
        if (n_order < 0) return fail(MH_ERR_INVALID_ARGUMENT, "n_order < 0");
        for (int32_t i = 0; i < n_order; ++i)
            if (order[i] < 0 || order[i] >= cells)
                return fail(MH_ERR_INVALID_ARGUMENT, "order[%d] = %d outside the grid", i, order[i]);
    }
    CHK(check_grid_values(grid, cells, lib->n));
    CHK(set_device(ctx));
    float* d_lab;
    int32_t *d_grid, *d_order = nullptr, *d_stats;
    uint32_t* d_usage;
    CHK(upload(ctx, S_CELL_LAB, cell_lab, (size_t)cells * 3, &d_lab));
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(upload(ctx, S_USAGE, usage, (size_t)lib->n, &d_usage));
    if (order) CHK(upload(ctx, S_ORDER, order, (size_t)(n_order > 0 ? n_order : 1), &d_order));
    uint8_t* d_stage = nullptr;
    float *d_dist = nullptr, *d_score = nullptr;
    if (stage) {
        CHK(upload(ctx, S_OUT0, stage, (size_t)cells, &d_stage));
    }
    if (dist) CHK(upload(ctx, S_OUT1, dist, (size_t)cells, &d_dist));
    if (score) CHK(upload(ctx, S_OUT2, score, (size_t)cells, &d_score));
    CHK(slot(ctx, S_STATS, 8, &d_stats));
    bool empty_grid = true;
    for (int32_t i = 0; i < cells && empty_grid; ++i) empty_grid = grid[i] == MH_EMPTY;
    const bool raster = !order && empty_grid;
    CHK(run_place(ctx, lib, d_lab, p->grid_w, p->grid_h, p->max_usage, p->weight, p->score_mode,
                  raster, d_order,
                  order ? n_order : cells, d_grid, d_usage, false, d_stage, d_dist, d_score,
                  d_stats));
    int32_t st[6];
    CHK(download(ctx, grid, d_grid, (size_t)cells));
    if (usage) CHK(download(ctx, usage, d_usage, (size_t)lib->n));
    CHK(download(ctx, stage, d_stage, (size_t)cells));
    CHK(download(ctx, dist, d_dist, (size_t)cells));
    CHK(download(ctx, score, d_score, (size_t)cells));
    CHK(download(ctx, st, d_stats, 6));
    CHK(sync(ctx));
    if (stats) {
        stats->placed = st[0];
        stats->main_path = st[1];
        stats->fallback_ban = st[2];
        stats->fallback_near = st[3];
        stats->resets = st[4];
        stats->empty = st[5];
    }
    return MH_OK;
}

// ------------------------------------------------------- a10-a12: cost/SA
int mh_total_cost(mh_context* ctx, const mh_library* lib, const int32_t* grid, int32_t gw,
                  int32_t gh, float* out) {
    if (!ctx || !lib || !out) return fail(MH_ERR_INVALID_ARGUMENT, "mh_total_cost: NULL argument");
    if (gw == 0 || gh == 0) {
        *out = 0.0f;
        return MH_OK;
    }
    CHK(check_geometry(gw, gh));
    if (!grid) return fail(MH_ERR_INVALID_ARGUMENT, "grid is NULL");
    const int32_t cells = gw * gh;
    CHK(check_grid_values(grid, cells, lib->n));
    CHK(set_device(ctx));
    int32_t* d_grid;
    float *d_scratch, *d_out;
    const int32_t chunk = cells < 65536 ? cells : 65536;
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(slot(ctx, S_OUT0, (size_t)chunk * 2 + 4, &d_scratch));
    CHK(slot(ctx, S_OUT1, 1, &d_out));
    const float* term = nullptr;
    CHK(library_terms(lib, ctx->stream, &term));
    launch_total_cost(d_grid, gw, gh, lib->d_lab4, term, lib->n, d_scratch, chunk, d_out,
                      ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, 1));
    return sync(ctx);
}

int mh_penalties(mh_context* ctx, const mh_library* lib, const int32_t* grid, int32_t gw,
                 int32_t gh, float weight, const int32_t* queries, int32_t q, float* out) {
    if (!ctx || !lib || !grid || q < 0 || (q > 0 && (!queries || !out)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_penalties: bad arguments");
    CHK(check_geometry(gw, gh));
    const int32_t cells = gw * gh;
    CHK(check_grid_values(grid, cells, lib->n));
    for (int32_t i = 0; i < q; ++i) {
        const int32_t m = queries[3 * i], x = queries[3 * i + 1], y = queries[3 * i + 2];
        if (m < 0 || m >= lib->n || x < 0 || x >= gw || y < 0 || y >= gh)
            return fail(MH_ERR_INVALID_ARGUMENT, "penalty query %d out of range", i);
    }
    if (q == 0) return MH_OK;
    CHK(set_device(ctx));
    int32_t *d_grid, *d_q;
    float* d_out;
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(upload(ctx, S_IN0, queries, (size_t)q * 3, &d_q));
    CHK(slot(ctx, S_OUT0, (size_t)q, &d_out));
    launch_penalties(d_grid, gw, gh, lib->d_lab4, weight, d_q, q, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, (size_t)q));
    return sync(ctx);
}

int mh_swap_deltas(mh_context* ctx, const mh_library* lib, const int32_t* grid, int32_t gw,
                   int32_t gh, const int32_t* pairs, int32_t p, float* out) {
    if (!ctx || !lib || !grid || p < 0 || (p > 0 && (!pairs || !out)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_swap_deltas: bad arguments");
    CHK(check_geometry(gw, gh));
    const int32_t cells = gw * gh;
    CHK(check_grid_values(grid, cells, lib->n));
    for (int32_t i = 0; i < p; ++i) {
        const int32_t* q = pairs + 4 * i;
        if (q[0] < 0 || q[0] >= gw || q[1] < 0 || q[1] >= gh || q[2] < 0 || q[2] >= gw || q[3] < 0 ||
            q[3] >= gh)
            return fail(MH_ERR_INVALID_ARGUMENT, "swap pair %d out of range", i);
    }
    if (p == 0) return MH_OK;
    CHK(set_device(ctx));
    int32_t *d_grid, *d_pairs;
    float* d_out;
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(upload(ctx, S_IN0, pairs, (size_t)p * 4, &d_pairs));
    CHK(slot(ctx, S_OUT0, (size_t)p, &d_out));
    const float* term = nullptr;
    CHK(library_terms(lib, ctx->stream, &term));
    launch_swap_deltas(d_grid, gw, gh, lib->d_lab4, term, lib->n, d_pairs, p, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, (size_t)p));
    return sync(ctx);
}


int mh_anneal(mh_context* ctx, const mh_library* lib, int32_t* grid, int32_t gw, int32_t gh,
              const mh_anneal_params* p, mh_anneal_result* result, float* trace_delta,
              uint8_t* trace_accept) {
    if (!ctx || !lib || !p) return fail(MH_ERR_INVALID_ARGUMENT, "mh_anneal: NULL argument");
    if (p->max_iterations < 0) return fail(MH_ERR_INVALID_ARGUMENT, "max_iterations < 0");
    if (gw == 0 || gh == 0) {  // empty grid: OptimizationResult::default()
        if (result) memset(result, 0, sizeof(*result));
        return MH_OK;
    }
    CHK(check_geometry(gw, gh));
    if (!grid) return fail(MH_ERR_INVALID_ARGUMENT, "grid is NULL");
    const int32_t cells = gw * gh;
    CHK(check_grid_values(grid, cells, lib->n));
    CHK(set_device(ctx));
    int32_t* d_grid;
    AnnealOut* d_out;
    float* d_td = nullptr;
    uint8_t* d_ta = nullptr;
    const size_t iters = (size_t)p->max_iterations;
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(slot(ctx, S_ANNEAL_OUT, 1, &d_out));
    if (trace_delta) CHK(slot(ctx, S_OUT0, iters, &d_td));
    if (trace_accept) CHK(slot(ctx, S_OUT1, iters, &d_ta));
    CHK(run_anneal(ctx, lib, d_grid, gw, gh, p, d_td, d_ta, d_out, nullptr, false));
    AnnealOut o;
    CHK(download(ctx, grid, d_grid, (size_t)cells));
    CHK(download(ctx, &o, d_out, 1));
    CHK(download(ctx, trace_delta, d_td, iters));
    CHK(download(ctx, trace_accept, d_ta, iters));
    CHK(sync(ctx));
    if (result) to_result(o, result);
    return MH_OK;
}

// ----------------------------------------------- a13-a16: adjust/composite
int mh_average_rgb(mh_context* ctx, const uint8_t* rgb, int32_t w, int32_t h, uint8_t out[3]) {
    if (!ctx || !rgb || !out || w <= 0 || h <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_average_rgb: bad arguments");
    CHK(set_device(ctx));
    uint8_t *d_in, *d_out;
    CHK(upload(ctx, S_IN0, rgb, (size_t)w * h * 3, &d_in));
    CHK(slot(ctx, S_OUT0, 3, &d_out));
    launch_avg_rgb(d_in, (int64_t)w * 3, 1, 0, 0, w, h, 1, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, 3));
    return sync(ctx);
}

int mh_optimal_adjustments(mh_context* ctx, const uint8_t* tile_rgb, const uint8_t* target_rgb,
                           int32_t n, float strength, float* out) {
    if (!ctx || n < 0 || (n > 0 && (!tile_rgb || !target_rgb || !out)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_optimal_adjustments: bad arguments");
    if (n == 0) return MH_OK;
    CHK(set_device(ctx));
    uint8_t *d_t, *d_g;
    float4* d_out;
    CHK(upload(ctx, S_IN0, tile_rgb, (size_t)n * 3, &d_t));
    CHK(upload(ctx, S_IN1, target_rgb, (size_t)n * 3, &d_g));
    CHK(slot(ctx, S_OUT0, (size_t)n, &d_out));
    launch_optimal_adjust(d_t, d_g, n, strength, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, reinterpret_cast<float4*>(out), d_out, (size_t)n));
    return sync(ctx);
}

int mh_adjust_pixels(mh_context* ctx, const float adj[4], const uint8_t* in, int64_t npx,
                     uint8_t* out) {
    if (!ctx || !adj || npx < 0 || (npx > 0 && (!in || !out)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_adjust_pixels: bad arguments");
    if (npx == 0) return MH_OK;
    CHK(set_device(ctx));
    uint8_t *d_in, *d_out;
    CHK(upload(ctx, S_IN0, in, (size_t)npx * 3, &d_in));
    CHK(slot(ctx, S_OUT0, (size_t)npx * 3, &d_out));
    launch_adjust_pixels(make_float4(adj[0], adj[1], adj[2], adj[3]), d_in, npx, d_out, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, out, d_out, (size_t)npx * 3));
    return sync(ctx);
}

int mh_resize(mh_context* ctx, const uint8_t* src, int32_t sw, int32_t sh, uint8_t* dst, int32_t dw,
              int32_t dh) {
    if (!ctx || !src || !dst || sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_resize: bad arguments");
    CHK(set_device(ctx));
    ResizeTaps h, v;
    CHK(get_taps(ctx, sw, dw, &h));
    CHK(get_taps(ctx, sh, dh, &v));
    uint8_t *d_src, *d_tmp, *d_dst;
    CHK(upload(ctx, S_IN0, src, (size_t)sw * sh * 3, &d_src));
    CHK(slot(ctx, S_OUT1, (size_t)sh * dw * 3, &d_tmp));
    CHK(slot(ctx, S_OUT0, (size_t)dw * dh * 3, &d_dst));
    launch_resize(d_src, 1, sw, sh, dw, dh, h, v, d_tmp, d_dst, ctx->stream);
    HIPCHK(hipGetLastError());
    CHK(download(ctx, dst, d_dst, (size_t)dw * dh * 3));
    return sync(ctx);
}

int mh_composite(mh_context* ctx, const mh_library* lib, const uint8_t* target, int32_t width,
                 int32_t height, const int32_t* grid, int32_t gw, int32_t gh, float strength,
                 uint8_t* out) {
    if (!ctx || !lib || !target || !grid || !out)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_composite: NULL argument");
    if (!lib->d_pixels) return fail(MH_ERR_INVALID_ARGUMENT, "library holds no pixels");
    CHK(check_geometry(gw, gh));
    const int32_t cells = gw * gh;
    CHK(check_grid_values(grid, cells, lib->n));
    const int32_t tw = width / gw, th = height / gh;
    const size_t out_bytes = (size_t)gw * tw * gh * th * 3;
    if (out_bytes == 0) return MH_OK;
    CHK(set_device(ctx));
    uint8_t *d_t, *d_out;
    int32_t* d_grid;
    CHK(upload(ctx, S_IN0, target, (size_t)width * height * 3, &d_t));
    CHK(upload(ctx, S_GRID, grid, (size_t)cells, &d_grid));
    CHK(slot(ctx, S_OUT0, out_bytes, &d_out));
    CHK(run_composite(ctx, lib, d_t, width, height, d_grid, gw, gh, strength, d_out, nullptr, nullptr,
                      false));
    CHK(download(ctx, out, d_out, out_bytes));
    return sync(ctx);
}

// -------------------------------------------------------------- whole job
}  // extern "C"

namespace {

int check_job_params(mh_context* ctx, const mh_library* lib, const uint8_t* d_target, int32_t width,
                     int32_t height, const mh_job_params* p, const uint8_t* d_image_out) {
    if (!ctx || !lib || !p || !d_target) return fail(MH_ERR_INVALID_ARGUMENT, "mh_generate: NULL argument");
    CHK(check_geometry(p->grid_w, p->grid_h));
    if (p->max_usage < 0) return fail(MH_ERR_INVALID_ARGUMENT, "max_usage must be >= 0");
    if (p->optimization_iterations < 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "optimization_iterations < 0");
    if (!isfinite(p->weight)) return fail(MH_ERR_INVALID_ARGUMENT, "weight must be finite");
    CHK(check_score_mode(p->score_mode));
    const int32_t tw = width / p->grid_w, th = height / p->grid_h;
    if (tw <= 0 || th <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "image %d x %d smaller than grid %d x %d", width, height,
                    p->grid_w, p->grid_h);
    if (d_image_out && !lib->d_pixels)
        return fail(MH_ERR_INVALID_ARGUMENT, "compositing needs a library with pixels");
    return MH_OK;
}

// Device working memory of one job (all device pointers).  temps is only
// read when `own_temps` is set (the allocation-free form fills it itself).
struct JobBufs {
    float* cell_lab = nullptr;
    CandRec* cand = nullptr;
    uint32_t* usage = nullptr;
    int32_t* stats = nullptr;
    AnnealOut* aout = nullptr;
    uint8_t* tavg = nullptr;
    float4* adj = nullptr;
    int32_t* grid = nullptr;
    float* temps = nullptr;
};

// T_0 = T0, T_{k+1} = T_k * decay in f32 (src/optimizer.rs:54, :102): the
// sequential recurrence, one lane (the host table get_temps builds holds the
// same values).
__global__ __launch_bounds__(64) void temps_kernel(float t0, float decay, int64_t len,
                                                  float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    float v = t0;
    for (int64_t i = 0; i < len; ++i) {
        out[i] = v;
        v = v * decay;
    }
}

// generate_mosaic minus file I/O on the context stream into `b`.  noalloc:
// no device allocation anywhere (no context scratch, term matrix, cached
// temperature table or tile build; no phase events).
int enqueue_generate(mh_context* ctx, const mh_library* lib, const uint8_t* d_target, int32_t width,
                     int32_t height, const mh_job_params* p, const JobBufs& b,
                     uint8_t* d_image_out, bool noalloc) {
    const int32_t gw = p->grid_w, gh = p->grid_h, cells = gw * gh;
    const int32_t tw = width / gw, th = height / gh;
    const uint32_t timing = ctx->timing;
    if (noalloc) ctx->timing = 0;
    struct Restore {
        mh_context* c;
        uint32_t t;
        ~Restore() { c->timing = t; }
    } restore{ctx, timing};
    // a1: target cell Lab means (src/main.rs:536-541); the same pixel pass
    // gives the cells' average RGB for the colour adjustment (a13)
    float strength = p->color_strength;
    strength = strength < 0.0f ? 0.0f : (strength > 1.0f ? 1.0f : strength);  // src/main.rs:162
    uint8_t* tavg = d_image_out && strength > 0.0f ? b.tavg : nullptr;
    phase_begin(ctx, MH_PHASE_CELL_LAB);
    launch_cell_lab(d_target, (int64_t)width * 3, gw, (int64_t)th * width * 3, (int64_t)tw * 3, tw,
                    th, cells, ctx->d_lut, b.cell_lab, ctx->stream, tavg);
    phase_end(ctx, MH_PHASE_CELL_LAB);
    // a3-a9: candidates + placement chain (raster order, empty grid, zero usage)
    CHK(run_place(ctx, lib, b.cell_lab, gw, gh, p->max_usage, p->weight, p->score_mode, true,
                  nullptr, cells, b.grid, b.usage, true, nullptr, nullptr, nullptr, b.stats,
                  b.cand));
    // a10-a12: SA when enabled and weight > 0 (src/main.rs:608)
    const bool sa = p->enable_optimization && p->weight > 0.0f;
    ctx->last_ran_anneal = sa ? 1 : 0;
    ctx->last_iters = p->optimization_iterations;
    ctx->last_stats = b.stats;
    ctx->last_aout = b.aout;
    if (sa) {
        mh_anneal_params ap;
        ap.max_iterations = p->optimization_iterations;
        ap.initial_temperature = p->initial_temperature;
        ap.temperature_decay = p->temperature_decay;
        ap.seed = p->seed;
        ap.greedy = 0;
        if (noalloc && ap.max_iterations > 0)
            hipLaunchKernelGGL(temps_kernel, dim3(1), dim3(64), 0, ctx->stream,
                               ap.initial_temperature, ap.temperature_decay, ap.max_iterations,
                               b.temps);
        CHK(run_anneal(ctx, lib, b.grid, gw, gh, &ap, nullptr, nullptr, b.aout,
                       noalloc ? b.temps : nullptr, noalloc));
    }
    // a13-a16: composite
    if (d_image_out)
        CHK(run_composite(ctx, lib, d_target, width, height, b.grid, gw, gh, strength, d_image_out,
                          tavg, b.adj, noalloc));
    ctx->last_out_w = gw * tw;
    ctx->last_out_h = gh * th;
    HIPCHK(hipGetLastError());
    return MH_OK;
}

// ------------------------------------------------ allocation-free job
// Working memory of mh_generate_device_noalloc: one array per GPU, held from
// the first load of this library until the process ends.  Host code never
// uses the symbol's own name as an address: mh_context_create looks up each
// device's copy with hipGetSymbolAddress (with that device current) and
// every launch takes that address.
static __device__ __align__(256) uint8_t scratch[MH_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int32_t, uint8_t*> g_scratch;  // device -> its copy of `scratch`

int scratch_address(int32_t device, uint8_t** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it == g_scratch.end()) {
        uint8_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(MH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        it = g_scratch.emplace(device, ptr).first;
    }
    *out = it->second;
    return MH_OK;
}

// The job's working memory carved from the scratch array (256-byte aligned
// pieces); bytes needed in *need.
void carve(uint8_t* base, int32_t cells, int32_t n, bool adjust, bool temps, int64_t iters,
           bool own_grid, JobBufs* b, size_t* need) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t* at = base ? base + o : nullptr;
        o = (o + bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t c = (size_t)cells;
    b->cell_lab = reinterpret_cast<float*>(take(c * 3 * sizeof(float)));
    b->cand = reinterpret_cast<CandRec*>(take(c * kCandStride * sizeof(CandRec)));
    b->usage = reinterpret_cast<uint32_t*>(take((size_t)(n > 0 ? n : 1) * sizeof(uint32_t)));
    b->stats = reinterpret_cast<int32_t*>(take(8 * sizeof(int32_t)));
    b->aout = reinterpret_cast<AnnealOut*>(take(sizeof(AnnealOut)));
    b->tavg = adjust ? take(c * 3) : nullptr;
    b->adj = adjust ? reinterpret_cast<float4*>(take(c * sizeof(float4))) : nullptr;
    b->temps = temps ? reinterpret_cast<float*>(take((size_t)iters * sizeof(float))) : nullptr;
    b->grid = own_grid ? reinterpret_cast<int32_t*>(take(c * sizeof(int32_t))) : nullptr;
    *need = o;
}

}  // namespace

namespace mh {
namespace host {
int noalloc_scratch_init(mh_context* ctx) { return scratch_address(ctx->device, &ctx->noalloc_scratch); }
}  // namespace host
}  // namespace mh

extern "C" {

int mh_generate_device(mh_context* ctx, const mh_library* lib, const uint8_t* d_target,
                       int32_t width, int32_t height, const mh_job_params* p,
                       int32_t* d_grid_out, uint8_t* d_image_out) {
    CHK(check_job_params(ctx, lib, d_target, width, height, p, d_image_out));
    CHK(set_device(ctx));
    const int32_t cells = p->grid_w * p->grid_h;
    JobBufs b;
    b.grid = d_grid_out;
    CHK(slot(ctx, S_CELL_LAB, (size_t)cells * 3, &b.cell_lab));
    if (!b.grid) CHK(slot(ctx, S_GRID, (size_t)cells, &b.grid));
    CHK(slot(ctx, S_USAGE, (size_t)(lib->n > 0 ? lib->n : 1), &b.usage));
    CHK(slot(ctx, S_STATS, 8, &b.stats));
    CHK(slot(ctx, S_ANNEAL_OUT, 1, &b.aout));
    CHK(slot(ctx, S_CAND, (size_t)cells * kCandStride, &b.cand));
    const float strength = p->color_strength;
    if (d_image_out && strength > 0.0f) {
        CHK(slot(ctx, S_TARGET_AVG, (size_t)cells * 3, &b.tavg));
        CHK(slot(ctx, S_ADJ, (size_t)cells, &b.adj));
    }
    return enqueue_generate(ctx, lib, d_target, width, height, p, b, d_image_out, false);
}

int mh_generate_device_noalloc(mh_context* ctx, const mh_library* lib, const uint8_t* d_target,
                               int32_t width, int32_t height, const mh_job_params* p,
                               int32_t* d_grid_out, uint8_t* d_image_out) {
    CHK(check_job_params(ctx, lib, d_target, width, height, p, d_image_out));
    CHK(set_device(ctx));
    if (!ctx->noalloc_scratch)
        return fail(MH_ERR_HIP, "the scratch array of device %d was not found at set-up",
                    ctx->device);
    const int32_t cells = p->grid_w * p->grid_h;
    const bool adjust = d_image_out && p->color_strength > 0.0f;
    const bool sa = p->enable_optimization && p->weight > 0.0f && p->optimization_iterations > 0;
    JobBufs b;
    size_t need = 0;
    carve(nullptr, cells, lib->n, adjust, sa, p->optimization_iterations, !d_grid_out, &b, &need);
    if (need > (size_t)MH_NOALLOC_SCRATCH_BYTES)
        return fail(MH_ERR_UNSUPPORTED,
                    "job of %d cells, %d materials and %lld annealing iterations needs %zu bytes "
                    "of working memory; the allocation-free entry point holds %llu",
                    cells, lib->n, (long long)p->optimization_iterations, need,
                    (unsigned long long)MH_NOALLOC_SCRATCH_BYTES);
    if (d_image_out) {  // refuse before anything is enqueued
        const uint8_t *tiles, *tile_avg;
        CHK(library_tiles_cached(lib, width / p->grid_w, height / p->grid_h, &tiles, &tile_avg));
    }
    carve(ctx->noalloc_scratch, cells, lib->n, adjust, sa, p->optimization_iterations,
          !d_grid_out, &b, &need);
    if (d_grid_out) b.grid = d_grid_out;
    return enqueue_generate(ctx, lib, d_target, width, height, p, b, d_image_out, true);
}

int mh_generate_device_result(mh_context* ctx, mh_job_result* r) {
    if (!ctx || !r) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    CHK(set_device(ctx));
    int32_t st[6];
    AnnealOut o;
    memset(&o, 0, sizeof(o));
    if (!ctx->last_stats) return fail(MH_ERR_INVALID_ARGUMENT, "no job has run on this context");
    CHK(download(ctx, st, ctx->last_stats, 6));
    if (ctx->last_ran_anneal) CHK(download(ctx, &o, ctx->last_aout, 1));
    CHK(sync(ctx));
    memset(r, 0, sizeof(*r));
    r->place.placed = st[0];
    r->place.main_path = st[1];
    r->place.fallback_ban = st[2];
    r->place.fallback_near = st[3];
    r->place.resets = st[4];
    r->place.empty = st[5];
    if (ctx->last_ran_anneal) to_result(o, &r->anneal);
    r->out_width = ctx->last_out_w;
    r->out_height = ctx->last_out_h;
    return MH_OK;
}

int mh_generate(mh_context* ctx, const mh_library* lib, const uint8_t* target, int32_t width,
                int32_t height, const mh_job_params* p, int32_t* grid_out, uint8_t* image_out,
                mh_job_result* result) {
    if (!ctx || !lib || !p || !target) return fail(MH_ERR_INVALID_ARGUMENT, "mh_generate: NULL argument");
    CHK(check_geometry(p->grid_w, p->grid_h));
    if (width <= 0 || height <= 0) return fail(MH_ERR_INVALID_ARGUMENT, "bad image size");
    CHK(set_device(ctx));
    const int32_t cells = p->grid_w * p->grid_h;
    const int32_t tw = width / p->grid_w, th = height / p->grid_h;
    const size_t out_bytes = (size_t)p->grid_w * tw * p->grid_h * th * 3;
    uint8_t *d_t, *d_img = nullptr;
    int32_t* d_grid;
    CHK(upload(ctx, S_IN0, target, (size_t)width * height * 3, &d_t));
    CHK(slot(ctx, S_IN1, (size_t)cells, &d_grid));
    if (image_out && out_bytes) CHK(slot(ctx, S_OUT3, out_bytes, &d_img));
    CHK(mh_generate_device(ctx, lib, d_t, width, height, p, d_grid, d_img));
    CHK(download(ctx, grid_out, d_grid, (size_t)cells));
    if (d_img) CHK(download(ctx, image_out, d_img, out_bytes));
    mh_job_result r;
    CHK(mh_generate_device_result(ctx, &r));
    if (result) *result = r;
    return MH_OK;
}

}  // extern "C"
