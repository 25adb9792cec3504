// mh_batch.hip -- many mosaic jobs per fused launch (mh_batch_*) and the
// continuous-batching request loop (mh_server_*).
//
// A batch runs every job's whole pipeline with one launch per stage:
//   cell Lab means -> K nearest -> placement chains (a wave per job) ->
//   annealing chains (a wave per job) -> cell averages + adjustments ->
//   composite.
// Wide stages find their job through a segment table (block -> job by
// binary search over per-job block offsets); chain stages take a job array.
// Every job's result is what mh_generate_device gives for it alone: the
// kernels are the same and no state crosses jobs.
//
// The server keeps libraries resident (Lab table, term matrix, resized
// tiles), admits queued jobs while their device footprint fits a memory
// budget, and launches what has been admitted as one batch on a free stream
// slot; batches on different slots overlap on the device.  A scheduler
// thread does admission, launches and completion (event queries), so jobs
// that arrive while others run join the next batch as slots and memory free.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>

#include "mh_internal.hpp"

using namespace mh;
using namespace mh::host;

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// A job after validation, with its library caches resolved and the layout of
// its working memory (offsets into one block of `bytes`).
struct JobSpec {
    mh_job job;
    int32_t gw = 0, gh = 0, cells = 0, tw = 0, th = 0, n = 0;
    int64_t pitch = 0;
    bool sa = false, adjust = false;
    int32_t anneal_kind = 0;  // the annealing variant (mh::anneal_kind)
    bool place_hbm = false;   // placement state in HBM (library beyond LDS)
    bool place_staged = false;  // the LDS chain's staged candidate rows fit too
    float strength = 0.0f;
    const float* term = nullptr;
    const float* temps = nullptr;
    const uint8_t* tiles = nullptr;
    const uint8_t* tile_avg = nullptr;
    size_t o_cell_lab = 0, o_cand = 0, o_usage = 0, o_tavg = 0, o_adj = 0;
    size_t o_grid = 0, o_image = 0;  // SIZE_MAX: the caller's buffer
    size_t bytes = 0;
};

// Host-only checks (no GPU work) and the working-memory layout.
int check_job(const mh_context* ctx, const mh_job& j, JobSpec* s) {
    if (!j.lib || !j.d_target) return fail(MH_ERR_INVALID_ARGUMENT, "job: NULL library or target");
    const mh_job_params& p = j.params;
    CHK(check_geometry(p.grid_w, p.grid_h));
    if (p.max_usage < 0) return fail(MH_ERR_INVALID_ARGUMENT, "max_usage must be >= 0");
    if (p.optimization_iterations < 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "optimization_iterations < 0");
    if (!isfinite(p.weight)) return fail(MH_ERR_INVALID_ARGUMENT, "weight must be finite");
    CHK(check_score_mode(p.score_mode));
    if (j.width <= 0 || j.height <= 0) return fail(MH_ERR_INVALID_ARGUMENT, "bad image size");
    const mh_library* lib = j.lib;
    if (lib->ctx != ctx) return fail(MH_ERR_INVALID_ARGUMENT, "library belongs to another context");
    s->job = j;
    s->gw = p.grid_w;
    s->gh = p.grid_h;
    s->cells = p.grid_w * p.grid_h;
    s->tw = j.width / p.grid_w;
    s->th = j.height / p.grid_h;
    s->n = lib->n;
    if (s->tw <= 0 || s->th <= 0)
        return fail(MH_ERR_INVALID_ARGUMENT, "image %d x %d smaller than grid %d x %d", j.width,
                    j.height, p.grid_w, p.grid_h);
    s->pitch = j.target_pitch ? j.target_pitch : (int64_t)j.width * 3;
    if (s->pitch < (int64_t)j.width * 3)
        return fail(MH_ERR_INVALID_ARGUMENT, "target pitch %lld < width*3", (long long)s->pitch);
    if (!lib->d_pixels) return fail(MH_ERR_INVALID_ARGUMENT, "compositing needs a library with pixels");
    s->sa = p.enable_optimization && p.weight > 0.0f;
    float k = p.color_strength;
    s->strength = k < 0.0f ? 0.0f : (k > 1.0f ? 1.0f : k);  // src/main.rs:162
    s->adjust = s->strength > 0.0f;
    // placement state beyond LDS runs the chain with grid and counts in HBM
    s->place_hbm = place_raster_lds_bytes(s->gw, s->n, false) > ctx->max_lds;
    s->place_staged = place_raster_lds_bytes(s->gw, s->n, true) <= ctx->max_lds;
    s->anneal_kind = s->sa ? anneal_kind(s->cells, s->n, s->gw, s->gh, ctx->max_lds) : 0;
    if (s->anneal_kind < 0)
        return fail(MH_ERR_UNSUPPORTED, "annealing a %d x %d grid: a side of 65,536 or more", s->gw,
                    s->gh);
    // working memory
    const size_t cells = (size_t)s->cells;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    s->o_cell_lab = take(cells * 3 * sizeof(float));
    s->o_cand = take(cells * kCandStride * sizeof(CandRec));
    s->o_usage = take((size_t)(s->n > 0 ? s->n : 1) * sizeof(uint32_t));
    s->o_tavg = s->adjust ? take(cells * 3) : 0;
    s->o_adj = s->adjust ? take(cells * sizeof(float4)) : 0;
    s->o_grid = j.d_grid_out ? SIZE_MAX : take(cells * sizeof(int32_t));
    s->o_image = j.d_image_out ? SIZE_MAX
                               : take((size_t)s->gw * s->tw * s->gh * s->th * 3);
    s->bytes = o;
    return MH_OK;
}

// Library-resident inputs of a job: term matrix, temperature table, resized
// tiles (each built once, on `stream`).
int prepare_job(mh_context* ctx, JobSpec* s, hipStream_t stream) {
    const mh_job_params& p = s->job.params;
    if (s->sa) {
        CHK(library_terms(s->job.lib, stream, &s->term));
        CHK(get_temps(ctx, p.initial_temperature, p.temperature_decay, p.optimization_iterations,
                      &s->temps));
    }
    CHK(library_tiles(s->job.lib, s->tw, s->th, stream, &s->tiles, &s->tile_avg));
    return MH_OK;
}

// One batch's descriptor tables and per-job results, laid out in one block.
struct BatchPlan {
    int32_t n_jobs = 0;
    size_t o_lab = 0, o_lab_st = 0, o_topk = 0, o_topk_st = 0, o_place = 0, o_anneal = 0;
    size_t o_comp = 0, o_adj_st = 0, o_comp_st = 0;
    size_t o_stats = 0, o_aout = 0, bytes = 0;
    int32_t lab_blocks = 0, lab_cpb = 4;
    int32_t n_topk = 0, topk_blocks = 0, topk_run = 1;
    size_t place_lds = 0;
    bool place_staged = false;  // the LDS chains stage their candidate rows
    int32_t place_score = kScoreSq;  // kScoreSq / kScoreDist / kScoreMixed over the jobs
    // placement entries: raster chains with state in LDS first, then the
    // chains with state in HBM, whose grids start empty (filled per launch)
    int32_t n_place_lds = 0;
    std::vector<std::pair<int32_t*, int32_t>> hbm_grids;
    // annealing entries grouped by variant (anneal_kind: the grid in LDS with
    // 16- or 32-bit material ids, or in HBM), one launch per group with its
    // own LDS size; group g's entries start at anneal_start[g]
    int32_t n_anneal = 0;
    int32_t anneal_start[4] = {0, 0, 0, 0};
    size_t anneal_lds[3] = {0, 0, 0};
    int32_t adj_blocks = 0;
    int32_t n_comp = 0, comp_blocks = 0;
    std::vector<int32_t> anneal_job;  // job index of each anneal entry
    // Launch shapes a tuner may set (build_plan keeps them): query cells per
    // top-K wave (0: topk_run_length's choice) and composite threads per block.
    // 64 threads (256 pixels of a row per block) by default: a block holds part
    // of one row, so wider blocks leave more lanes idle at the ends of the
    // rows (on the served stream's mix of row widths, 93% of lanes busy with
    // 256-pixel blocks against 77% with 1,024).
    int32_t tune_topk_run = 0;
    int32_t comp_threads = 64;
};

constexpr int32_t kStatsInts = 8;

size_t plan_bytes(int32_t n, BatchPlan* p) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    const size_t nn = (size_t)(n > 0 ? n : 1);
    p->o_lab = take(nn * sizeof(CellSeg));
    p->o_lab_st = take((nn + 1) * sizeof(int32_t));
    p->o_topk = take(nn * sizeof(TopkSeg));
    p->o_topk_st = take((nn + 1) * sizeof(int32_t));
    p->o_place = take(nn * sizeof(PlaceJob));
    p->o_anneal = take(nn * sizeof(AnnealJob));
    p->o_comp = take(nn * sizeof(CompSeg));
    p->o_adj_st = take((nn + 1) * sizeof(int32_t));
    p->o_comp_st = take((nn + 1) * sizeof(int32_t));
    p->o_stats = take(nn * kStatsInts * sizeof(int32_t));
    p->o_aout = take(nn * sizeof(AnnealOut));
    p->bytes = o;
    return o;
}

int32_t topk_blocks_of(int32_t cells, int32_t run) { return mh::topk_blocks(cells, run); }

template <typename T>
T* at(uint8_t* base, size_t off) {
    return reinterpret_cast<T*>(base + off);
}

// Fill the host image `h` of a batch block whose device copy lives at `d`;
// job i's working memory is at job_mem[i].  stage_rows: the raster chains
// stage their candidate rows in LDS when every job's fit (52 KB more per
// chain).  A batch's one or two chain launches gain from it (the chain time
// is the batch's critical path); the server keeps several batches' chains
// and wide stages on the GPU at once, where the larger LDS footprint costs
// more in chains and blocks resident per CU than the staging saves.
int build_plan(const std::vector<JobSpec>& specs, const std::vector<uint8_t*>& job_mem, uint8_t* h,
               uint8_t* d, bool stage_rows, BatchPlan* p) {
    const int32_t n = (int32_t)specs.size();
    p->n_jobs = n;
    plan_bytes(n, p);
    memset(h, 0, p->bytes);
    CellSeg* lab = at<CellSeg>(h, p->o_lab);
    int32_t* lab_st = at<int32_t>(h, p->o_lab_st);
    TopkSeg* topk = at<TopkSeg>(h, p->o_topk);
    int32_t* topk_st = at<int32_t>(h, p->o_topk_st);
    PlaceJob* place = at<PlaceJob>(h, p->o_place);
    AnnealJob* anneal = at<AnnealJob>(h, p->o_anneal);
    CompSeg* comp = at<CompSeg>(h, p->o_comp);
    int32_t* adj_st = at<int32_t>(h, p->o_adj_st);
    int32_t* comp_st = at<int32_t>(h, p->o_comp_st);
    int32_t* d_stats = at<int32_t>(d, p->o_stats);
    AnnealOut* d_aout = at<AnnealOut>(d, p->o_aout);

    int64_t lab_blocks = 0, topk_blocks = 0, adj_blocks = 0, comp_blocks = 0;
    p->n_topk = p->n_anneal = p->n_comp = 0;
    int64_t topk_queries = 0, all_cells = 0;
    for (const JobSpec& s : specs) {
        if (s.n > 0) topk_queries += s.cells;
        all_cells += s.cells;
    }
    p->lab_cpb = cell_lab_cells_per_block(all_cells);
    p->topk_run = p->tune_topk_run > 0 ? p->tune_topk_run : topk_run_length(topk_queries);
    p->place_lds = 0;
    p->place_staged = stage_rows;  // until a job's staged rows do not fit
    size_t lds_plain = 0, lds_staged = 0;
    p->place_score = kScoreSq;
    p->hbm_grids.clear();
    p->n_place_lds = 0;
    std::vector<PlaceJob> hbm_place;
    for (size_t& l : p->anneal_lds) l = 0;
    p->anneal_job.clear();
    std::vector<AnnealJob> by_kind[3];
    std::vector<int32_t> job_of_kind[3];
    for (int32_t i = 0; i < n; ++i) {
        const JobSpec& s = specs[i];
        const mh_job& j = s.job;
        const mh_library* lib = j.lib;
        uint8_t* m = job_mem[i];
        float* cell_lab = at<float>(m, s.o_cell_lab);
        CandRec* cand = at<CandRec>(m, s.o_cand);
        int32_t* grid = s.o_grid == SIZE_MAX ? j.d_grid_out : at<int32_t>(m, s.o_grid);
        uint8_t* image = s.o_image == SIZE_MAX ? j.d_image_out : at<uint8_t>(m, s.o_image);
        // a1: cell Lab means of the target
        lab_st[i] = (int32_t)lab_blocks;
        // (with the colour adjustment, the same pixel pass gives the cells'
        // average RGB, a13)
        uint8_t* tavg = s.adjust ? at<uint8_t>(m, s.o_tavg) : nullptr;
        lab[i] = CellSeg{j.d_target, s.pitch, (int64_t)s.th * s.pitch, (int64_t)s.tw * 3, s.gw, s.tw,
                         s.th, s.cells, cell_lab, tavg};
        lab_blocks += cell_lab_blocks(s.cells, p->lab_cpb);
        // a3: K nearest (libraries with materials only)
        if (s.n > 0) {
            topk_st[p->n_topk] = (int32_t)topk_blocks;
            topk[p->n_topk] = TopkSeg{lib->d_L, lib->d_A, lib->d_B, cell_lab, nullptr, nullptr,
                                      cand, s.n, s.cells, s.n < kMaxCand ? s.n : kMaxCand,
                                      kCandStride, library_blocks(lib)};
            ++p->n_topk;
            topk_blocks += topk_blocks_of(s.cells, p->topk_run);
        }
        // a4-a9: placement chain
        PlaceJob pj{};
        pj.lab4 = lib->d_lab4;
        pj.n = s.n;
        pj.cand = cand;
        pj.order = nullptr;
        pj.n_order = s.cells;
        pj.grid = grid;
        pj.usage = at<uint32_t>(m, s.o_usage);
        pj.zero_usage = 1;
        pj.grid_w = s.gw;
        pj.grid_h = s.gh;
        pj.max_usage = j.params.max_usage;
        pj.weight = j.params.weight;
        pj.score_sqrt = j.params.score_mode == MH_SCORE_DISTANCE ? 1 : 0;
        pj.stats = d_stats + (size_t)i * kStatsInts;
        if (s.place_hbm) {
            hbm_place.push_back(pj);
            p->hbm_grids.emplace_back(grid, s.cells);
        } else {
            place[p->n_place_lds++] = pj;
            lds_plain = std::max(lds_plain, place_raster_lds_bytes(s.gw, s.n, false));
            lds_staged = std::max(lds_staged, place_raster_lds_bytes(s.gw, s.n, true));
            p->place_staged = p->place_staged && s.place_staged;
        }
        const int32_t kind = pj.score_sqrt ? kScoreDist : kScoreSq;
        p->place_score = i == 0 || p->place_score == kind ? kind : kScoreMixed;
        // a10-a12: annealing chain
        if (s.sa) {
            AnnealJob aj{};
            aj.lab4 = lib->d_lab4;
            aj.term = s.term;
            aj.n = s.n;
            aj.grid = grid;
            aj.grid_w = s.gw;
            aj.grid_h = s.gh;
            aj.max_iter = j.params.optimization_iterations;
            aj.temps = s.temps;
            aj.seed = j.params.seed;
            aj.greedy = 0;
            aj.result = d_aout + i;
            const int32_t kind = s.anneal_kind;
            by_kind[kind].push_back(aj);
            job_of_kind[kind].push_back(i);
            const size_t al = anneal_lds_bytes(s.cells, kind);
            if (al > p->anneal_lds[kind]) p->anneal_lds[kind] = al;
        }
        // a13-a16: adjustments, composite
        float4* adj = s.adjust ? at<float4>(m, s.o_adj) : nullptr;
        adj_st[p->n_comp] = (int32_t)adj_blocks;
        comp_st[p->n_comp] = (int32_t)comp_blocks;
        comp[p->n_comp] = CompSeg{grid, s.gw, s.gh, s.tw, s.th, s.tiles, s.tile_avg, tavg, s.strength,
                                  s.adjust ? 1 : 0, adj, image, composite_narrow(s.n, s.tw, s.th)};
        ++p->n_comp;
        adj_blocks += s.adjust ? cell_adjust_blocks(s.cells) : 0;
        comp_blocks += composite_blocks(s.gw * s.tw, s.gh * s.th, p->comp_threads);
    }
    for (size_t h = 0; h < hbm_place.size(); ++h) place[p->n_place_lds + h] = hbm_place[h];
    p->place_lds = p->place_staged ? lds_staged : lds_plain;
    p->n_anneal = 0;
    for (int kind = 0; kind < 3; ++kind) {
        p->anneal_start[kind] = p->n_anneal;
        for (size_t a = 0; a < by_kind[kind].size(); ++a) {
            anneal[p->n_anneal++] = by_kind[kind][a];
            p->anneal_job.push_back(job_of_kind[kind][a]);
        }
    }
    p->anneal_start[3] = p->n_anneal;
    if (lab_blocks > INT32_MAX || topk_blocks > INT32_MAX || comp_blocks > INT32_MAX)
        return fail(MH_ERR_UNSUPPORTED, "batch too large for one launch");
    lab_st[n] = (int32_t)lab_blocks;
    topk_st[p->n_topk] = (int32_t)topk_blocks;
    adj_st[p->n_comp] = (int32_t)adj_blocks;
    comp_st[p->n_comp] = (int32_t)comp_blocks;
    p->lab_blocks = (int32_t)lab_blocks;
    p->topk_blocks = (int32_t)topk_blocks;
    p->adj_blocks = (int32_t)adj_blocks;
    p->comp_blocks = (int32_t)comp_blocks;
    return MH_OK;
}

// Optional HIP events bracketing each stage of a launch sequence (phases
// whose bit is set in `mask`; events ev[phase][0] before, [1] after).
struct PhaseEvents {
    uint32_t mask = 0;
    hipEvent_t (*ev)[2] = nullptr;
    void mark(int phase, int which, hipStream_t s) const {
        if ((mask >> phase) & 1u) hipEventRecord(ev[phase][which], s);
    }
};

// The stages up to the candidate lists (cell Lab means, nearest materials).
void enqueue_candidates(const BatchPlan& p, uint8_t* d, const float* lut, hipStream_t s,
                        const PhaseEvents& pe = PhaseEvents()) {
    pe.mark(MH_PHASE_CELL_LAB, 0, s);
    launch_cell_lab_batch(at<CellSeg>(d, p.o_lab), at<int32_t>(d, p.o_lab_st), p.n_jobs, p.lab_blocks,
                          p.lab_cpb, lut, s);
    pe.mark(MH_PHASE_CELL_LAB, 1, s);
    pe.mark(MH_PHASE_TOPK, 0, s);
    launch_topk_batch(at<TopkSeg>(d, p.o_topk), at<int32_t>(d, p.o_topk_st), p.n_topk, p.topk_blocks,
                      p.topk_run, s);
    pe.mark(MH_PHASE_TOPK, 1, s);
}

// The chain stages (placement, annealing).
void enqueue_chains(const BatchPlan& p, uint8_t* d, hipStream_t s,
                    const PhaseEvents& pe = PhaseEvents()) {
    pe.mark(MH_PHASE_PLACE, 0, s);
    const PlaceJob* pj = at<PlaceJob>(d, p.o_place);
    if (p.n_place_lds > 0)
        launch_place_batch(pj, p.n_place_lds, true, p.place_staged, p.place_score, p.place_lds,
                           s);
    if (p.n_jobs > p.n_place_lds) {
        for (const auto& g : p.hbm_grids) fill_i32(g.first, g.second, kEmpty, s);
        launch_place_batch(pj + p.n_place_lds, p.n_jobs - p.n_place_lds, false, false,
                           p.place_score, 0, s);
    }
    pe.mark(MH_PHASE_PLACE, 1, s);
    pe.mark(MH_PHASE_ANNEAL, 0, s);
    const AnnealJob* aj = at<AnnealJob>(d, p.o_anneal);
    for (int kind = 0; kind < 3; ++kind) {
        const int32_t count = p.anneal_start[kind + 1] - p.anneal_start[kind];
        if (count > 0)
            launch_anneal_batch(aj + p.anneal_start[kind], count, kind, p.anneal_lds[kind], s);
    }
    pe.mark(MH_PHASE_ANNEAL, 1, s);
}

// The output stages (colour adjustment, composite).
void enqueue_outputs(const BatchPlan& p, uint8_t* d, hipStream_t s,
                     const PhaseEvents& pe = PhaseEvents()) {
    pe.mark(MH_PHASE_ADJUST, 0, s);
    if (p.adj_blocks > 0)  // the target averages came from the cell Lab pass
        launch_cell_adjust_batch(at<CompSeg>(d, p.o_comp), at<int32_t>(d, p.o_adj_st), p.n_comp,
                                 p.adj_blocks, s);
    pe.mark(MH_PHASE_ADJUST, 1, s);
    pe.mark(MH_PHASE_COMPOSITE, 0, s);
    launch_composite_batch(at<CompSeg>(d, p.o_comp), at<int32_t>(d, p.o_comp_st), p.n_comp,
                           p.comp_blocks, p.comp_threads, s);
    pe.mark(MH_PHASE_COMPOSITE, 1, s);
}

void enqueue_plan(const BatchPlan& p, uint8_t* d, const float* lut, hipStream_t s,
                  const PhaseEvents& pe = PhaseEvents()) {
    enqueue_candidates(p, d, lut, s, pe);
    enqueue_chains(p, d, s, pe);
    enqueue_outputs(p, d, s, pe);
}

void fill_result(const JobSpec& s, const int32_t* st, const AnnealOut* ao, mh_job_result* r) {
    memset(r, 0, sizeof(*r));
    r->place.placed = st[0];
    r->place.main_path = st[1];
    r->place.fallback_ban = st[2];
    r->place.fallback_near = st[3];
    r->place.resets = st[4];
    r->place.empty = st[5];
    if (s.sa) to_result(*ao, &r->anneal);
    r->out_width = s.gw * s.tw;
    r->out_height = s.gh * s.th;
}

}  // namespace

// ================================================================= batch
// A batch runs as two groups of jobs on two streams: the jobs whose grids
// have at least half the cells of the largest one ("long": their placement
// chains set the batch time) on the context stream, the others on a side
// stream.  The short jobs' wide stages and chains then overlap the long
// chains instead of running before and after them.  With one group (similar
// sizes, or a single job) the side stream is unused.
struct mh_batch {
    mh_context* ctx = nullptr;
    std::vector<JobSpec> specs;
    std::vector<uint8_t*> job_mem;
    uint8_t* d_jobs = nullptr;   // all jobs' working memory
    uint8_t* d_plan = nullptr;   // both groups' descriptor tables + results
    std::vector<uint8_t> h_plan;
    int32_t n_groups = 0;
    std::vector<int32_t> members[2];  // job indices of each group
    BatchPlan plan[2];
    size_t plan_off[2] = {0, 0};      // each group's offset in d_plan
    std::vector<int32_t> group_of, index_in;  // job -> (group, position)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Group g's descriptor tables, built into the host image h_plan.
int build_group(mh_batch* b, int g) {
    std::vector<JobSpec> sp;
    std::vector<uint8_t*> mem;
    for (int32_t i : b->members[g]) {
        sp.push_back(b->specs[i]);
        mem.push_back(b->job_mem[i]);
    }
    return build_plan(sp, mem, b->h_plan.data() + b->plan_off[g], b->d_plan + b->plan_off[g],
                      true, &b->plan[g]);
}

void enqueue_batch(mh_batch* b) {
    mh_context* ctx = b->ctx;
    if (b->n_groups == 1) {
        enqueue_plan(b->plan[0], b->d_plan + b->plan_off[0], ctx->d_lut, ctx->stream);
        return;
    }
    // the long jobs' candidates first on the whole device, then their chains
    // start while the short jobs run beside them
    enqueue_candidates(b->plan[0], b->d_plan + b->plan_off[0], ctx->d_lut, ctx->stream);
    hipEventRecord(b->fork, ctx->stream);
    hipStreamWaitEvent(b->side, b->fork, 0);
    enqueue_chains(b->plan[0], b->d_plan + b->plan_off[0], ctx->stream);
    enqueue_outputs(b->plan[0], b->d_plan + b->plan_off[0], ctx->stream);
    enqueue_plan(b->plan[1], b->d_plan + b->plan_off[1], ctx->d_lut, b->side);
    hipEventRecord(b->join, b->side);
    hipStreamWaitEvent(ctx->stream, b->join, 0);
}

}  // namespace

namespace mh {
namespace host {

void batch_release(mh_batch* b) {
    mh_context* ctx = b->ctx;
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (b->side) hipStreamSynchronize(b->side);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_jobs) hipFree(b->d_jobs);
    if (b->d_plan) hipFree(b->d_plan);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->join) hipEventDestroy(b->join);
    if (b->side) hipStreamDestroy(b->side);
    b->exec = nullptr;
    b->graph = nullptr;
    b->d_jobs = b->d_plan = nullptr;
    b->fork = b->join = nullptr;
    b->side = nullptr;
    b->ctx = nullptr;
}

}  // namespace host
}  // namespace mh

extern "C" {

int64_t mh_job_footprint(mh_context* ctx, const mh_job* job) {
    if (!ctx || !job) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    JobSpec s;
    CHK(check_job(ctx, *job, &s));
    return (int64_t)s.bytes;
}

void mh_batch_destroy(mh_batch* b) {
    if (!b) return;
    if (mh_context* ctx = b->ctx) {
        {
            std::lock_guard<std::mutex> g(ctx->dep_mu);
            for (size_t i = 0; i < ctx->batches.size(); ++i)
                if (ctx->batches[i] == b) {
                    ctx->batches[i] = ctx->batches.back();
                    ctx->batches.pop_back();
                    break;
                }
        }
        batch_release(b);
    }
    delete b;
}

int mh_batch_create(mh_context* ctx, const mh_job* jobs, int32_t n_jobs, mh_batch** out) {
    if (!ctx || !out || n_jobs <= 0 || !jobs)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_batch_create: bad arguments");
    *out = nullptr;
    CHK(set_device(ctx));
    mh_batch* b = new mh_batch();
    b->ctx = ctx;
    {
        std::lock_guard<std::mutex> g(ctx->dep_mu);
        ctx->batches.push_back(b);
    }
    b->specs.resize((size_t)n_jobs);
    size_t total = 0;
    std::vector<size_t> offs((size_t)n_jobs);
    int rc = MH_OK;
    for (int32_t i = 0; i < n_jobs && !rc; ++i) {
        rc = check_job(ctx, jobs[i], &b->specs[i]);
        if (rc) {
            std::string m = mh_last_error();
            rc = fail(rc, "job %d: %s", i, m.c_str());
            break;
        }
        rc = prepare_job(ctx, &b->specs[i], ctx->stream);
        offs[i] = total;
        total = align_up(total + b->specs[i].bytes);
    }
    if (!rc) {  // groups: long chains on the context stream, the rest beside them
        int32_t max_cells = 0;
        for (const JobSpec& sp : b->specs) max_cells = std::max(max_cells, sp.cells);
        b->group_of.resize((size_t)n_jobs);
        b->index_in.resize((size_t)n_jobs);
        for (int32_t i = 0; i < n_jobs; ++i) {
            const int g = 2 * (int64_t)b->specs[i].cells >= max_cells ? 0 : 1;
            b->group_of[i] = g;
            b->index_in[i] = (int32_t)b->members[g].size();
            b->members[g].push_back(i);
        }
        if (b->members[1].empty()) {
            b->n_groups = 1;
        } else {
            b->n_groups = 2;
            hipError_t e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
            if (e != hipSuccess) rc = fail(MH_ERR_HIP, "batch stream: %s", hipGetErrorString(e));
        }
    }
    hipError_t e = hipSuccess;
    size_t plan_total = 0;
    if (!rc) {
        for (int g = 0; g < b->n_groups; ++g) {
            b->plan_off[g] = plan_total;
            plan_total = align_up(plan_total + plan_bytes((int32_t)b->members[g].size(), &b->plan[g]));
        }
        e = hipMalloc(&b->d_jobs, total ? total : kAlign);
        if (e == hipSuccess) e = hipMalloc(&b->d_plan, plan_total);
        if (e != hipSuccess)
            rc = fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP, "batch memory: %s",
                      hipGetErrorString(e));
    }
    if (!rc) {
        b->job_mem.resize((size_t)n_jobs);
        for (int32_t i = 0; i < n_jobs; ++i) b->job_mem[i] = b->d_jobs + offs[i];
        b->h_plan.assign(plan_total, 0);
        for (int g = 0; g < b->n_groups && !rc; ++g) rc = build_group(b, g);
    }
    if (!rc) {
        e = hipMemcpy(b->d_plan, b->h_plan.data(), plan_total, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(MH_ERR_HIP, "batch tables: %s", hipGetErrorString(e));
    }
    if (rc) {
        mh_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MH_OK;
}

extern "C++" {
namespace {

// Average launch time (ms) of `reps` launches after one untimed launch, all
// on stream s.
template <typename F>
int time_launch(hipStream_t s, hipEvent_t t0, hipEvent_t t1, int32_t reps, F&& launch, float* ms) {
    launch();
    HIPCHK(hipEventRecord(t0, s));
    for (int32_t r = 0; r < reps; ++r) launch();
    HIPCHK(hipEventRecord(t1, s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(t1));
    float total = 0.0f;
    HIPCHK(hipEventElapsedTime(&total, t0, t1));
    *ms = total / (float)reps;
    return MH_OK;
}

// Rebuild group g's tables with its current launch shapes and upload the
// descriptor part (the per-job results that follow it keep their values).
int refresh_group(mh_batch* b, int g) {
    CHK(build_group(b, g));
    HIPCHK(hipMemcpy(b->d_plan + b->plan_off[g], b->h_plan.data() + b->plan_off[g],
                     b->plan[g].o_stats, hipMemcpyHostToDevice));
    return MH_OK;
}

// A captured graph holds the launch shapes it was captured with.
int drop_graph(mh_batch* b) {
    if (b->exec) {
        HIPCHK(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        HIPCHK(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    return MH_OK;
}

void report_tuning(const mh_batch* b, mh_batch_tuning* t) {
    memset(t, 0, sizeof(*t));
    t->n_groups = b->n_groups;
    for (int g = 0; g < b->n_groups; ++g) {
        t->topk_cells_per_wave[g] = b->plan[g].n_topk > 0 ? b->plan[g].topk_run : 0;
        t->composite_threads[g] = b->plan[g].comp_threads;
    }
}

struct EventPair {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    ~EventPair() {
        if (t0) hipEventDestroy(t0);
        if (t1) hipEventDestroy(t1);
    }
};

}  // namespace
}  // extern "C++"

int mh_batch_tune(mh_batch* b, int32_t repeats, mh_batch_tuning* tuning) {
    if (!b) return fail(MH_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (repeats < 1 || repeats > 100)
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_batch_tune: repeats %d outside [1, 100]", repeats);
    mh_context* ctx = b->ctx;
    CHK(set_device(ctx));
    const hipStream_t s = ctx->stream;
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.t0));
    HIPCHK(hipEventCreate(&ev.t1));
    // the timed stages read cell Labs and grids: one run of the batch first
    enqueue_batch(b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    CHK(drop_graph(b));  // captured with the old shapes
    mh_batch_tuning t{};
    t.n_groups = b->n_groups;
    static const int32_t kRuns[] = {1, 2, 4, 8};       // topk_run accepts any run >= 1
    static const int32_t kThreads[] = {64, 128, 256};  // composite_batch_kernel's bounds
    for (int g = 0; g < b->n_groups; ++g) {
        BatchPlan& p = b->plan[g];
        uint8_t* d = b->d_plan + b->plan_off[g];
        if (p.n_topk > 0) {
            float best = 0.0f;
            int32_t pick = 0;
            for (int32_t run : kRuns) {
                p.tune_topk_run = run;
                CHK(refresh_group(b, g));
                float ms = 0.0f;
                CHK(time_launch(s, ev.t0, ev.t1, repeats, [&] {
                    launch_topk_batch(at<TopkSeg>(d, p.o_topk), at<int32_t>(d, p.o_topk_st), p.n_topk,
                                      p.topk_blocks, p.topk_run, s);
                }, &ms));
                if (!pick || ms < best) best = ms, pick = run;
            }
            p.tune_topk_run = pick;
            t.topk_ms[g] = best;
        }
        float best = 0.0f;
        int32_t pick = 0;
        for (int32_t threads : kThreads) {
            p.comp_threads = threads;
            CHK(refresh_group(b, g));
            float ms = 0.0f;
            CHK(time_launch(s, ev.t0, ev.t1, repeats, [&] {
                launch_composite_batch(at<CompSeg>(d, p.o_comp), at<int32_t>(d, p.o_comp_st), p.n_comp,
                                       p.comp_blocks, p.comp_threads, s);
            }, &ms));
            if (!pick || ms < best) best = ms, pick = threads;
        }
        p.comp_threads = pick;
        t.composite_ms[g] = best;
        CHK(refresh_group(b, g));
        t.topk_cells_per_wave[g] = p.n_topk > 0 ? p.topk_run : 0;
        t.composite_threads[g] = p.comp_threads;
    }
    if (tuning) *tuning = t;
    return MH_OK;
}

int mh_batch_set_tuning(mh_batch* b, const mh_batch_tuning* tuning) {
    if (!b || !tuning) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int g = 0; g < 2; ++g) {
        const int32_t run = tuning->topk_cells_per_wave[g], thr = tuning->composite_threads[g];
        if (run < 0 || run > 8)
            return fail(MH_ERR_INVALID_ARGUMENT, "topk_cells_per_wave %d outside [0, 8]", run);
        if (thr != 0 && thr != 64 && thr != 128 && thr != 256)
            return fail(MH_ERR_INVALID_ARGUMENT, "composite_threads %d not 0, 64, 128 or 256", thr);
    }
    mh_context* ctx = b->ctx;
    CHK(set_device(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // the tables may be in use
    CHK(drop_graph(b));
    for (int g = 0; g < b->n_groups; ++g) {
        b->plan[g].tune_topk_run = tuning->topk_cells_per_wave[g];
        if (tuning->composite_threads[g]) b->plan[g].comp_threads = tuning->composite_threads[g];
        CHK(refresh_group(b, g));
    }
    return MH_OK;
}

int mh_batch_get_tuning(mh_batch* b, mh_batch_tuning* tuning) {
    if (!b || !tuning) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    report_tuning(b, tuning);
    return MH_OK;
}

int mh_batch_set_graph(mh_batch* b, int32_t enabled) {
    if (!b) return fail(MH_ERR_INVALID_ARGUMENT, "batch is NULL");
    b->use_graph = enabled != 0;
    return MH_OK;
}

int mh_batch_run(mh_batch* b) {
    if (!b) return fail(MH_ERR_INVALID_ARGUMENT, "batch is NULL");
    mh_context* ctx = b->ctx;
    CHK(set_device(ctx));
    if (!b->use_graph) {
        enqueue_batch(b);
        HIPCHK(hipGetLastError());
        return MH_OK;
    }
    if (!b->exec) {
        // the side stream joins the capture through the fork event
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_batch(b);
        hipError_t le = hipGetLastError();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (le != hipSuccess || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return fail(MH_ERR_HIP, "batch capture: %s",
                        hipGetErrorString(le != hipSuccess ? le : e));
        }
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, ctx->stream));
    return MH_OK;
}

int mh_batch_results(mh_batch* b, mh_job_result* results) {
    if (!b || !results) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    mh_context* ctx = b->ctx;
    CHK(set_device(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // the side stream joined it
    for (int g = 0; g < b->n_groups; ++g) {
        const BatchPlan& p = b->plan[g];
        const uint8_t* d = b->d_plan + b->plan_off[g];
        std::vector<int32_t> st((size_t)p.n_jobs * kStatsInts);
        std::vector<AnnealOut> ao((size_t)p.n_jobs);
        HIPCHK(hipMemcpy(st.data(), d + p.o_stats, st.size() * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ao.data(), d + p.o_aout, ao.size() * sizeof(AnnealOut),
                         hipMemcpyDeviceToHost));
        for (int32_t k = 0; k < p.n_jobs; ++k) {
            const int32_t i = b->members[g][k];
            fill_result(b->specs[i], &st[(size_t)k * kStatsInts], &ao[k], &results[i]);
        }
    }
    return MH_OK;
}

int mh_batch_job_outputs(mh_batch* b, int32_t i, int32_t** d_grid, uint8_t** d_image) {
    if (!b || i < 0 || i >= (int32_t)b->specs.size())
        return fail(MH_ERR_INVALID_ARGUMENT, "bad job index");
    const JobSpec& s = b->specs[i];
    if (d_grid)
        *d_grid = s.o_grid == SIZE_MAX ? s.job.d_grid_out : at<int32_t>(b->job_mem[i], s.o_grid);
    if (d_image)
        *d_image = s.o_image == SIZE_MAX ? s.job.d_image_out : at<uint8_t>(b->job_mem[i], s.o_image);
    return MH_OK;
}

}  // extern "C"

// ================================================================ server
namespace {

// First-fit sub-allocator over one device allocation (offsets only).
struct Arena {
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes

    void reset(size_t bytes) {
        size = bytes;
        used = peak = 0;
        free_.clear();
        free_[0] = bytes;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[*off + bytes] = rest;
            used += bytes;
            if (used > peak) peak = used;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes ? bytes : 1);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// A server's device working memory: chunks allocated as admitted jobs need
// them, up to the budget, each sub-allocated first-fit; a job lives in one
// chunk.  When no job holds memory and a job still does not fit, the chunks
// are given back so that one chunk of the needed size can be made.
struct ChunkArena {
    static constexpr size_t kChunk = (size_t)1 << 30;
    struct Chunk {
        uint8_t* base = nullptr;
        Arena arena;
    };
    size_t budget = 0, reserved = 0, used = 0, peak = 0;
    std::vector<Chunk> chunks;

    void release_all() {
        for (Chunk& c : chunks) hipFree(c.base);
        chunks.clear();
        reserved = 0;
    }
    // Device memory for `bytes` (chunk index in *chunk), or nullptr when it
    // does not fit now.
    uint8_t* alloc(size_t bytes, int32_t* chunk) {
        bytes = align_up(bytes ? bytes : 1);
        for (int attempt = 0; attempt < 2; ++attempt) {
            for (size_t c = 0; c < chunks.size(); ++c) {
                size_t off = 0;
                if (chunks[c].arena.alloc(bytes, &off)) {
                    *chunk = (int32_t)c;
                    return take(chunks[c].base + off, bytes);
                }
            }
            if (bytes <= budget - reserved) {
                const size_t size = std::min(budget - reserved, std::max(kChunk, bytes));
                Chunk c;
                if (hipMalloc(&c.base, size) == hipSuccess) {
                    c.arena.reset(size);
                    chunks.push_back(c);
                    reserved += size;
                    size_t off = 0;
                    chunks.back().arena.alloc(bytes, &off);
                    *chunk = (int32_t)chunks.size() - 1;
                    return take(chunks.back().base + off, bytes);
                }
                (void)hipGetLastError();  // an allocation failure is not sticky
            }
            if (used != 0) return nullptr;  // wait for running jobs to give memory back
            release_all();                  // nothing held: start from one fitting chunk
        }
        return nullptr;
    }
    uint8_t* take(uint8_t* p, size_t bytes) {
        used += bytes;
        if (used > peak) peak = used;
        return p;
    }
    void release(int32_t chunk, const uint8_t* p, size_t bytes) {
        Chunk& c = chunks[(size_t)chunk];
        c.arena.release((size_t)(p - c.base), bytes);
        used -= align_up(bytes ? bytes : 1);
    }
};

enum JobState { J_QUEUED = 0, J_RUNNING = 1, J_DONE = 2, J_FAILED = 3 };

// Chain streams of a server with n slots: one per slot up to four, then one
// per two slots, at most eight (each has an output stream beside it).
inline int32_t chain_streams(int32_t n_slots) {
    return n_slots <= 4 ? n_slots : std::min(std::max(n_slots / 2, 4), 8);
}
constexpr int kWideStreams = 2;       // wide streams of a server (batches alternate)

struct JobRec {
    JobSpec spec;
    int state = J_QUEUED;
    int err = MH_OK;
    std::string msg;
    mh_job_result result{};
    double t_submit = 0, t_launch = 0, t_done = 0;
    uint8_t* mem = nullptr;  // working memory (in chunk `chunk` of the server's arena)
    int32_t chunk = 0;
    bool holds_mem = false;
    int32_t passed_over = 0;  // batches formed without it while it was the oldest queued
};

// One batch in flight: its descriptor tables, result staging and events.
// Its wide stages run on the server's wide stream, its chains on `lane`.
struct StreamSlot {
    int32_t lane = 0;             // chain stream of the running batch
    int32_t wide = 0;             // wide stream of the running batch
    hipEvent_t uploaded = nullptr; // descriptor tables on the device (upload stream)
    hipEvent_t ready = nullptr;    // candidate lists written (wide stream)
    hipEvent_t chained = nullptr;  // grids placed and annealed (chain stream)
    hipEvent_t done = nullptr;     // images written, results copied back (output stream)
    uint8_t* d_plan = nullptr;    // descriptor tables + results (device)
    uint8_t* h_plan = nullptr;    // pinned staging of the tables
    uint8_t* h_res = nullptr;     // pinned copy of stats + anneal results
    BatchPlan plan;
    std::vector<int64_t> tickets;
    bool busy = false;
    // per-phase timing of the batch on this slot (events made at creation)
    hipEvent_t ev[MH_PHASES][2] = {};
    uint32_t timed = 0;   // phases bracketed in the running batch
    int64_t cells = 0;    // cells of the running batch
};

}  // namespace

// Batches go through three kinds of stream.  The wide stages up to the
// candidate lists (cell Labs, K nearest), which fill the device, run on two
// wide streams, consecutive batches alternating (so one batch's launch tail
// overlaps the next batch's start); each batch's chains (placement,
// annealing: one wave per job, latency-bound) run on one of a few chain
// streams after an event from the wide stream; its output stages
// (adjustment, composite: wide again) run on that chain stream's output
// stream after an event from the chain stream, so a batch's outputs never
// wait for another chain stream's chains.  While one batch's chains run,
// other batches' wide stages use the rest of the device.  Every stream
// is made with a full CU mask, which gives it a hardware queue of its own
// (streams without one share a few queues, and a wide launch would wait
// behind chains in a shared in-order queue).
struct mh_server {
    mh_context* ctx = nullptr;
    int32_t max_batch = 128;
    ChunkArena arena;
    hipStream_t wide[kWideStreams] = {};      // wide stages, batches alternating
    hipStream_t up = nullptr;                 // descriptor uploads
    int64_t launched = 0;                     // batches launched (picks the wide stream)
    std::vector<hipStream_t> lanes, outs;  // chain streams and their output streams
    std::vector<int32_t> lane_load;  // batches in flight per chain stream
    std::vector<StreamSlot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::deque<JobRec> jobs;  // by ticket
    std::thread worker;
    std::atomic<bool> stop{false};
    int64_t batches = 0, batched_jobs = 0, completed = 0;
    int32_t max_seen_batch = 0;
    // per-phase device timing (mh_server_set_timing_phases / mh_server_timing)
    uint32_t timing = 0;
    double phase_ms[MH_PHASES] = {};
    int64_t phase_n[MH_PHASES] = {}, phase_cells[MH_PHASES] = {};
    int fatal = MH_OK;
    std::string fatal_msg;
};

namespace {

// Batch formation over a window of the queue, its next
// kWindowBatches * max_batch jobs.  The batch's anchor is the oldest queued
// job -- except at the end of a backlog, when the whole queue fits the
// window: then it is the job with the most cells (the oldest of equals),
// so the last batches hold the shortest chains, unless the oldest job has
// already been passed over kWindowBatches times.  The rest of the batch are
// the window's jobs nearest to the anchor in cell count (by ratio), in queue
// order.  A placement launch lasts as long as its longest chain (one wave
// walks a job's cells in order), so batches of like-sized jobs keep each
// chain stream from waiting on one long job per batch.  Which jobs share a
// batch changes no job's result.  With at most max_batch jobs queued this is
// first-come first-served.
constexpr size_t kWindowBatches = 8;

std::vector<size_t> pick_batch(const mh_server& sv) {
    const size_t b = (size_t)sv.max_batch;
    const size_t wn = std::min(sv.queue.size(), b * kWindowBatches);
    std::vector<size_t> pos;
    if (wn <= b) {
        for (size_t i = 0; i < wn; ++i) pos.push_back(i);
        return pos;
    }
    auto cells = [&](size_t i) { return sv.jobs[(size_t)sv.queue[i]].spec.cells; };
    size_t a = 0;
    if (sv.queue.size() == wn &&
        sv.jobs[(size_t)sv.queue[0]].passed_over < (int32_t)kWindowBatches)
        for (size_t i = 1; i < wn; ++i)
            if (cells(i) > cells(a)) a = i;
    auto lcells = [&](size_t i) { return std::log((double)std::max<int32_t>(cells(i), 1)); };
    const double c0 = lcells(a);
    std::vector<std::pair<double, size_t>> d;
    d.reserve(wn - 1);
    for (size_t i = 0; i < wn; ++i)
        if (i != a) d.push_back({std::fabs(lcells(i) - c0), i});
    std::nth_element(d.begin(), d.begin() + (ptrdiff_t)(b - 1), d.end());
    pos.push_back(a);
    for (size_t k = 0; k + 1 < b; ++k) pos.push_back(d[k].second);
    std::sort(pos.begin(), pos.end());
    return pos;
}

// Drops queue positions `taken` (ascending, all inside the queue).
void remove_positions(std::deque<int64_t>* q, const std::vector<size_t>& taken) {
    const size_t wn = taken.back() + 1;
    std::vector<int64_t> keep;
    keep.reserve(wn - taken.size());
    size_t j = 0;
    for (size_t i = 0; i < wn; ++i) {
        if (j < taken.size() && taken[j] == i)
            ++j;
        else
            keep.push_back((*q)[i]);
    }
    q->erase(q->begin(), q->begin() + (ptrdiff_t)wn);
    q->insert(q->begin(), keep.begin(), keep.end());
}

// Launch the admitted jobs `tickets` as one batch on slot `sl`.  Called by
// the worker without the lock held (the records it reads are not touched by
// other threads while RUNNING).
int launch_batch(mh_server* sv, StreamSlot& sl, const std::vector<JobRec*>& recs) {
    mh_context* ctx = sv->ctx;
    hipStream_t w = sv->wide[sl.wide], c = sv->lanes[(size_t)sl.lane];
    std::vector<JobSpec> specs;
    std::vector<uint8_t*> mem;
    specs.reserve(recs.size());
    mem.reserve(recs.size());
    for (JobRec* r : recs) {
        CHK(prepare_job(ctx, &r->spec, w));
        specs.push_back(r->spec);
        mem.push_back(r->mem);
    }
    CHK(build_plan(specs, mem, sl.h_plan, sl.d_plan, false, &sl.plan));
    // the tables go up on their own stream as soon as the batch is admitted;
    // the wide stream, which usually still has earlier batches to run, waits
    // for them only if the copy has not landed by then
    HIPCHK(hipMemcpyAsync(sl.d_plan, sl.h_plan, sl.plan.o_stats, hipMemcpyHostToDevice, sv->up));
    HIPCHK(hipEventRecord(sl.uploaded, sv->up));
    HIPCHK(hipStreamWaitEvent(w, sl.uploaded, 0));
    PhaseEvents pe;
    pe.mask = sl.timed;
    pe.ev = sl.ev;
    enqueue_candidates(sl.plan, sl.d_plan, ctx->d_lut, w, pe);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(sl.ready, w));
    HIPCHK(hipStreamWaitEvent(c, sl.ready, 0));
    enqueue_chains(sl.plan, sl.d_plan, c, pe);
    HIPCHK(hipGetLastError());
    hipStream_t o = sv->outs[(size_t)sl.lane];
    HIPCHK(hipEventRecord(sl.chained, c));
    HIPCHK(hipStreamWaitEvent(o, sl.chained, 0));
    enqueue_outputs(sl.plan, sl.d_plan, o, pe);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sl.h_res, sl.d_plan + sl.plan.o_stats, sl.plan.bytes - sl.plan.o_stats,
                          hipMemcpyDeviceToHost, o));
    HIPCHK(hipEventRecord(sl.done, o));
    return MH_OK;
}

// The chain stream with the fewest batches in flight.
int32_t pick_lane(const mh_server* sv) {
    int32_t best = 0;
    for (int32_t i = 1; i < (int32_t)sv->lane_load.size(); ++i)
        if (sv->lane_load[(size_t)i] < sv->lane_load[(size_t)best]) best = i;
    return best;
}

// A stream with a hardware queue of its own (full CU mask), or a plain
// non-blocking stream where the runtime has no CU-masked streams.
hipError_t make_stream(int device, hipStream_t* s) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess && prop.multiProcessorCount > 0) {
        std::vector<uint32_t> mask((size_t)(prop.multiProcessorCount + 31) / 32, 0xffffffffu);
        if (prop.multiProcessorCount % 32)
            mask.back() = (1u << (prop.multiProcessorCount % 32)) - 1u;
        e = hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data());
        if (e == hipSuccess) return e;
        (void)hipGetLastError();
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

void worker_loop(mh_server* sv) {
    hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        bool progressed = false;
        // 1. completions
        for (StreamSlot& sl : sv->slots) {
            if (!sl.busy) continue;
            lk.unlock();
            const hipError_t q = hipEventQuery(sl.done);
            lk.lock();
            if (q == hipErrorNotReady) continue;
            const double t = now_ms();
            const int32_t* st = reinterpret_cast<const int32_t*>(sl.h_res);
            const AnnealOut* ao =
                reinterpret_cast<const AnnealOut*>(sl.h_res + (sl.plan.o_aout - sl.plan.o_stats));
            if (q == hipSuccess)
                for (int ph = 0; ph < MH_PHASES; ++ph) {
                    float ms = 0.0f;
                    if (((sl.timed >> ph) & 1u) &&
                        hipEventElapsedTime(&ms, sl.ev[ph][0], sl.ev[ph][1]) == hipSuccess) {
                        sv->phase_ms[ph] += ms;
                        ++sv->phase_n[ph];
                        sv->phase_cells[ph] += sl.cells;
                    }
                }
            for (size_t i = 0; i < sl.tickets.size(); ++i) {
                JobRec& r = sv->jobs[(size_t)sl.tickets[i]];
                if (q == hipSuccess) {
                    fill_result(r.spec, st + i * kStatsInts, ao + i, &r.result);
                    r.state = J_DONE;
                } else {
                    r.state = J_FAILED;
                    r.err = MH_ERR_HIP;
                    r.msg = std::string("batch failed: ") + hipGetErrorString(q);
                }
                r.t_done = t;
                if (r.holds_mem) sv->arena.release(r.chunk, r.mem, r.spec.bytes);
                r.holds_mem = false;
                ++sv->completed;
            }
            sl.tickets.clear();
            sl.busy = false;
            --sv->lane_load[(size_t)sl.lane];
            progressed = true;
            sv->cv_done.notify_all();
        }
        // 2. admission + launch on a free slot
        StreamSlot* free_slot = nullptr;
        for (StreamSlot& sl : sv->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (free_slot && !sv->queue.empty()) {
            std::vector<int64_t> take;
            std::vector<JobRec*> recs;
            std::vector<size_t> taken;  // queue positions, ascending
            size_t first = 0;           // the batch's first job (queue position)
            for (size_t i : pick_batch(*sv)) {
                const int64_t t = sv->queue[i];
                JobRec& r = sv->jobs[(size_t)t];
                uint8_t* mem = sv->arena.alloc(r.spec.bytes, &r.chunk);
                if (!mem) {  // the rest waits for memory
                    first = i;
                    break;
                }
                r.mem = mem;
                r.holds_mem = true;
                r.state = J_RUNNING;
                r.t_launch = now_ms();
                take.push_back(t);
                recs.push_back(&r);  // deque elements keep their address across push_back
                taken.push_back(i);
            }
            if (take.empty()) {
                // the batch's first job waits for memory -- unless nothing
                // runs that could give any back, when the device itself has
                // no room for it
                bool busy = false;
                for (const StreamSlot& sl : sv->slots) busy |= sl.busy;
                if (!busy) {
                    JobRec& r = sv->jobs[(size_t)sv->queue[first]];
                    r.state = J_FAILED;
                    r.err = MH_ERR_OUT_OF_MEMORY;
                    r.msg = "no device memory for the job's working set";
                    r.t_done = now_ms();
                    ++sv->completed;
                    sv->queue.erase(sv->queue.begin() + (ptrdiff_t)first);
                    sv->cv_done.notify_all();
                    progressed = true;
                }
            } else {
                if (taken.front() != 0) ++sv->jobs[(size_t)sv->queue.front()].passed_over;
                remove_positions(&sv->queue, taken);
            }
            if (!take.empty()) {
                free_slot->busy = true;
                free_slot->lane = pick_lane(sv);
                free_slot->wide = (int32_t)(sv->launched++ % kWideStreams);
                ++sv->lane_load[(size_t)free_slot->lane];
                free_slot->tickets = take;
                free_slot->timed = sv->timing;
                free_slot->cells = 0;
                for (JobRec* r : recs) free_slot->cells += r->spec.cells;
                ++sv->batches;
                sv->batched_jobs += (int64_t)take.size();
                if ((int32_t)take.size() > sv->max_seen_batch) sv->max_seen_batch = (int32_t)take.size();
                lk.unlock();
                const int rc = launch_batch(sv, *free_slot, recs);
                const std::string m = rc ? mh_last_error() : "";
                lk.lock();
                if (rc) {
                    // nothing of this batch may still run: wait for its streams
                    lk.unlock();
                    hipStreamSynchronize(sv->up);
                    hipStreamSynchronize(sv->wide[free_slot->wide]);
                    hipStreamSynchronize(sv->lanes[(size_t)free_slot->lane]);
                    hipStreamSynchronize(sv->outs[(size_t)free_slot->lane]);
                    lk.lock();
                    for (int64_t t : take) {
                        JobRec& r = sv->jobs[(size_t)t];
                        r.state = J_FAILED;
                        r.err = rc;
                        r.msg = m;
                        r.t_done = now_ms();
                        if (r.holds_mem) sv->arena.release(r.chunk, r.mem, r.spec.bytes);
                        r.holds_mem = false;
                        ++sv->completed;
                    }
                    free_slot->tickets.clear();
                    free_slot->busy = false;
                    --sv->lane_load[(size_t)free_slot->lane];
                    sv->cv_done.notify_all();
                }
                progressed = true;
            }
        }
        bool any_busy = false;
        for (StreamSlot& sl : sv->slots) any_busy |= sl.busy;
        if (sv->stop && sv->queue.empty() && !any_busy) break;
        if (progressed) continue;
        if (any_busy) {
            // poll completions without holding the lock
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            sv->cv_work.wait(lk, [&] { return sv->stop.load() || !sv->queue.empty(); });
        }
    }
}

void server_free(mh_server* sv) {
    hipSetDevice(sv->ctx->device);
    if (sv->up) hipStreamSynchronize(sv->up);
    for (hipStream_t w : sv->wide)
        if (w) hipStreamSynchronize(w);
    for (hipStream_t c : sv->lanes)
        if (c) hipStreamSynchronize(c);
    for (hipStream_t o : sv->outs)
        if (o) hipStreamSynchronize(o);
    for (StreamSlot& sl : sv->slots) {
        if (sl.uploaded) hipEventDestroy(sl.uploaded);
        if (sl.ready) hipEventDestroy(sl.ready);
        if (sl.chained) hipEventDestroy(sl.chained);
        if (sl.done) hipEventDestroy(sl.done);
        for (auto& pair : sl.ev)
            for (hipEvent_t e : pair)
                if (e) hipEventDestroy(e);
        if (sl.d_plan) hipFree(sl.d_plan);
        if (sl.h_plan) hipHostFree(sl.h_plan);
        if (sl.h_res) hipHostFree(sl.h_res);
    }
    for (hipStream_t c : sv->lanes)
        if (c) hipStreamDestroy(c);
    for (hipStream_t w : sv->wide)
        if (w) hipStreamDestroy(w);
    if (sv->up) hipStreamDestroy(sv->up);
    for (hipStream_t o : sv->outs)
        if (o) hipStreamDestroy(o);
    sv->arena.release_all();
    delete sv;
}

}  // namespace

extern "C" {

int mh_server_create(mh_context* ctx, const mh_server_params* p, mh_server** out) {
    if (!ctx || !out) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    const int64_t budget = p && p->memory_budget > 0 ? p->memory_budget : (int64_t)64 << 30;
    const int32_t max_batch = p && p->max_batch > 0 ? p->max_batch : 128;
    const int32_t n_slots = p && p->streams > 0 ? p->streams : 8;
    if (n_slots > 16 || max_batch > 65536)
        return fail(MH_ERR_INVALID_ARGUMENT, "streams must be <= 16 and max_batch <= 65536");
    CHK(set_device(ctx));
    mh_server* sv = new mh_server();
    sv->ctx = ctx;
    sv->max_batch = max_batch;
    sv->slots.resize((size_t)n_slots);
    BatchPlan tmp;
    const size_t plan = plan_bytes(max_batch, &tmp);
    hipError_t e = hipSuccess;
    for (hipStream_t& w : sv->wide)
        if (e == hipSuccess) e = make_stream(ctx->device, &w);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sv->up, hipStreamNonBlocking);
    sv->lanes.assign((size_t)chain_streams(n_slots), nullptr);
    sv->outs.assign(sv->lanes.size(), nullptr);
    sv->lane_load.assign(sv->lanes.size(), 0);
    for (size_t i = 0; i < sv->lanes.size(); ++i) {
        if (e == hipSuccess) e = make_stream(ctx->device, &sv->lanes[i]);
        if (e == hipSuccess) e = make_stream(ctx->device, &sv->outs[i]);
    }
    for (StreamSlot& sl : sv->slots) {
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.chained, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        for (auto& pair : sl.ev)
            for (hipEvent_t& ev : pair)
                if (e == hipSuccess) e = hipEventCreate(&ev);
        if (e == hipSuccess) e = hipMalloc(&sl.d_plan, plan);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_plan, plan, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_res, plan - tmp.o_stats, hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        server_free(sv);
        return fail(e == hipErrorOutOfMemory ? MH_ERR_OUT_OF_MEMORY : MH_ERR_HIP,
                    "server resources (budget %lld B): %s", (long long)budget, hipGetErrorString(e));
    }
    sv->arena.budget = (size_t)budget;
    sv->worker = std::thread(worker_loop, sv);
    *out = sv;
    return MH_OK;
}

int mh_server_submit(mh_server* sv, const mh_job* job, int64_t* ticket) {
    if (!sv || !job || !ticket) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    JobSpec s;
    CHK(check_job(sv->ctx, *job, &s));
    if (align_up(s.bytes) > sv->arena.budget)
        return fail(MH_ERR_OUT_OF_MEMORY, "job needs %zu B, more than the server budget %zu B", s.bytes,
                    sv->arena.budget);
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->stop) return fail(MH_ERR_INVALID_ARGUMENT, "server is shutting down");
    JobRec r;
    r.spec = s;
    r.t_submit = now_ms();
    sv->jobs.push_back(std::move(r));
    *ticket = (int64_t)sv->jobs.size() - 1;
    sv->queue.push_back(*ticket);
    sv->cv_work.notify_one();
    return MH_OK;
}

int mh_server_submit_many(mh_server* sv, const mh_job* jobs, int32_t n, int64_t* tickets) {
    if (!sv || n < 0 || (n > 0 && (!jobs || !tickets)))
        return fail(MH_ERR_INVALID_ARGUMENT, "mh_server_submit_many: bad arguments");
    // every job is checked before any is queued
    std::vector<JobSpec> specs((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        if (const int rc = check_job(sv->ctx, jobs[i], &specs[(size_t)i]))
            return fail(rc, "job %d: %s", i, mh_last_error());
        if (align_up(specs[(size_t)i].bytes) > sv->arena.budget)
            return fail(MH_ERR_OUT_OF_MEMORY, "job %d needs %zu B, more than the server budget %zu B",
                        i, specs[(size_t)i].bytes, sv->arena.budget);
    }
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->stop) return fail(MH_ERR_INVALID_ARGUMENT, "server is shutting down");
    const double t = now_ms();
    for (int32_t i = 0; i < n; ++i) {
        JobRec r;
        r.spec = specs[(size_t)i];
        r.t_submit = t;
        sv->jobs.push_back(std::move(r));
        tickets[i] = (int64_t)sv->jobs.size() - 1;
        sv->queue.push_back(tickets[i]);
    }
    sv->cv_work.notify_one();
    return MH_OK;
}

int mh_server_wait(mh_server* sv, int64_t ticket, mh_job_result* result, double* latency_ms) {
    if (!sv) return fail(MH_ERR_INVALID_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (ticket < 0 || ticket >= (int64_t)sv->jobs.size())
        return fail(MH_ERR_NOT_FOUND, "no job with ticket %lld", (long long)ticket);
    sv->cv_done.wait(lk, [&] { return sv->jobs[(size_t)ticket].state >= J_DONE; });
    const JobRec& r = sv->jobs[(size_t)ticket];
    if (latency_ms) *latency_ms = r.t_done - r.t_submit;
    if (r.state == J_FAILED) return fail(r.err, "%s", r.msg.c_str());
    if (result) *result = r.result;
    return MH_OK;
}

int mh_server_poll(mh_server* sv, int64_t ticket, int32_t* done) {
    if (!sv || !done) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> g(sv->mu);
    if (ticket < 0 || ticket >= (int64_t)sv->jobs.size())
        return fail(MH_ERR_NOT_FOUND, "no job with ticket %lld", (long long)ticket);
    *done = sv->jobs[(size_t)ticket].state >= J_DONE ? 1 : 0;
    return MH_OK;
}

int mh_server_drain(mh_server* sv) {
    if (!sv) return fail(MH_ERR_INVALID_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->completed == (int64_t)sv->jobs.size(); });
    return MH_OK;
}

int mh_server_job_times(mh_server* sv, int64_t ticket, double times_ms[3]) {
    if (!sv || !times_ms) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> g(sv->mu);
    if (ticket < 0 || ticket >= (int64_t)sv->jobs.size())
        return fail(MH_ERR_NOT_FOUND, "no job with ticket %lld", (long long)ticket);
    const JobRec& r = sv->jobs[(size_t)ticket];
    times_ms[0] = r.t_submit;
    times_ms[1] = r.t_launch;
    times_ms[2] = r.t_done;
    return MH_OK;
}

int mh_server_get_stats(mh_server* sv, mh_server_stats* st) {
    if (!sv || !st) return fail(MH_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> g(sv->mu);
    st->submitted = (int64_t)sv->jobs.size();
    st->completed = sv->completed;
    st->batches = sv->batches;
    st->max_batch_jobs = sv->max_seen_batch;
    st->memory_in_use = (int64_t)sv->arena.used;
    st->memory_peak = (int64_t)sv->arena.peak;
    st->memory_budget = (int64_t)sv->arena.budget;
    return MH_OK;
}

int mh_server_set_timing_phases(mh_server* sv, uint32_t phase_mask) {
    if (!sv) return fail(MH_ERR_INVALID_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> g(sv->mu);
    sv->timing = phase_mask & ((1u << MH_PHASES) - 1u);
    return MH_OK;
}

int mh_server_timing(mh_server* sv, double ms[MH_PHASES], int64_t launches[MH_PHASES],
                     int64_t cells[MH_PHASES]) {
    if (!sv) return fail(MH_ERR_INVALID_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> g(sv->mu);
    for (int ph = 0; ph < MH_PHASES; ++ph) {
        if (ms) ms[ph] = sv->phase_ms[ph];
        if (launches) launches[ph] = sv->phase_n[ph];
        if (cells) cells[ph] = sv->phase_cells[ph];
        sv->phase_ms[ph] = 0.0;
        sv->phase_n[ph] = sv->phase_cells[ph] = 0;
    }
    return MH_OK;
}

void mh_server_destroy(mh_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    server_free(sv);
}

}  // extern "C"
