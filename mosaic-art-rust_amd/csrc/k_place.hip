// k_place.hip -- greedy placement chain (generate_mosaic's cell loop,
// src/main.rs:529-545, driving find_and_use_best_tile_with_position
// :335-418 and fallback_tile_selection :420-475, with UsageTracker
// src/lib.rs:57-89, can_place_at_position :309-333 and
// AdjacencyPenaltyCalculator::calculate_penalty src/adjacency.rs:63-88).
//
// The chain is sequential across cells (global usage counts, a global
// usage reset, and the just-placed neighbours), so one job is one wave64
// walking its cells in order.  Per cell the wave evaluates the K <= 100
// candidates in parallel (2 per lane): usage < max, neighbour ban, weighted
// penalty (similarities recomputed from Labs -- bit-identical to the stored
// matrix), score = sqdist + penalty; the reference's first strict-< winner
// is the minimum of (score, rank).  Many jobs run as many independent waves.
//
// The GUI generator's variant (src/gui/app_full.rs:1396-1563) scores
// sqrt(sqdist) + penalty with everything else unchanged; with score_sqrt set
// the candidate distances are replaced by their correctly rounded square
// roots as they are loaded.  sqrt is monotone, so the candidates stay in
// ascending distance order and the contender bound below holds unchanged.
// The raster chain converts a cell's records one step before they are
// scored (when their loads have landed anyway), in a variant compiled for
// it, so the squared-distance chain carries no extra work.
//
// Two variants:
//  * place_raster_kernel -- what generate_mosaic runs: raster order from an
//    empty grid, so only Up and Left can be filled.  The left neighbour is
//    the previous winner (registers), the up neighbour comes from a per-column
//    LDS row (material + Lab of the last placement in that column), usage
//    counts live in LDS, candidate records are read 24 cells at a time into
//    registers (from an LDS double buffer that the next group's rows are
//    copied into meanwhile, where LDS holds it; else straight from global
//    memory), the min is a DPP reduction, and the winner's id / Lab / count
//    are read from its lane.  Winners wait one per lane in a register that
//    is stored to the grid, coalesced, once per group of 24 cells.
//    Latency-bound: a few hundred cycles per cell.
//  * place_chain_kernel -- any cell order over a pre-filled grid and usage
//    (find_and_use_best_tile_with_position at arbitrary cells); grid and usage
//    in LDS, all four neighbours read per cell.  Its HBM form (launched with
//    no LDS) keeps grid and counts in the job's own buffers, for libraries
//    whose counts do not fit LDS; it also runs the raster order then.
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

// ------------------------------------------------------------ general chain
// State access of the general chain: grid and usage counts in LDS, or (kHbm,
// for libraries whose counts do not fit LDS) the job's own buffers in HBM,
// read and written with agent-scope atomics so no lane sees a stale cache
// line of a word another lane has just written.
template <bool kHbm>
struct ChainState {
    int32_t* grid;
    uint32_t* usage;
    __device__ __forceinline__ int32_t cell(int32_t i) const {
        if constexpr (kHbm)
            return __hip_atomic_load((MH_GAS int32_t*)grid + i, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        else
            return grid[i];
    }
    __device__ __forceinline__ void set_cell(int32_t i, int32_t v) const {
        if constexpr (kHbm)
            __hip_atomic_store((MH_GAS int32_t*)grid + i, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            grid[i] = v;
    }
    __device__ __forceinline__ uint32_t count(int32_t m) const {
        if constexpr (kHbm)
            return __hip_atomic_load((MH_GAS uint32_t*)usage + m, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        else
            return usage[m];
    }
    __device__ __forceinline__ void set_count(int32_t m, uint32_t v) const {
        if constexpr (kHbm)
            __hip_atomic_store((MH_GAS uint32_t*)usage + m, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            usage[m] = v;
    }
    // every lane's writes before every lane's later reads
    __device__ __forceinline__ void sync() const {
        if constexpr (kHbm) __threadfence();
        __syncthreads();
    }
};

template <bool kHbm>
__global__ __launch_bounds__(64) void place_chain_kernel(PlaceJob single,
                                                         const PlaceJob* __restrict__ jobs) {
    extern __shared__ int32_t sm[];
    const PlaceJob job = uniform_copy(jobs ? jobs[blockIdx.x] : single);
    const int lane = threadIdx.x;
    const int32_t gw = job.grid_w, gh = job.grid_h, n = job.n;
    const int32_t cells = gw * gh;
    ChainState<kHbm> S;
    if constexpr (kHbm) {
        S.grid = job.grid;
        S.usage = job.usage;
        if (job.zero_usage)
            for (int32_t i = lane; i < n; i += 64) S.set_count(i, 0u);
    } else {
        S.grid = sm;
        S.usage = reinterpret_cast<uint32_t*>(sm + cells);
        for (int32_t i = lane; i < cells; i += 64) S.grid[i] = gload(job.grid, i);
        for (int32_t i = lane; i < n; i += 64)
            S.usage[i] = job.zero_usage ? 0u : gload(job.usage, i);
    }
    S.sync();

    const int32_t K = n < kMaxCand ? n : kMaxCand;
    const uint32_t max_usage = (uint32_t)job.max_usage;
    const float weight = job.weight;
    const bool use_pen = weight > 0.0f;
    int32_t st_main = 0, st_ban = 0, st_near = 0, st_resets = 0, st_empty = 0;
    const int32_t steps = job.order ? job.n_order : cells;

    for (int32_t c = 0; c < steps; ++c) {
        const int32_t cell = job.order ? gload(job.order, c) : c;
        if (n == 0) {
            ++st_empty;
            continue;
        }
        const int32_t x = cell % gw, y = cell / gw;
        // neighbours in get_adjacent_positions order: Up, Down, Left, Right
        const int32_t nb_up = y > 0 ? S.cell(cell - gw) : kEmpty;
        const int32_t nb_down = y < gh - 1 ? S.cell(cell + gw) : kEmpty;
        const int32_t nb_left = x > 0 ? S.cell(cell - 1) : kEmpty;
        const int32_t nb_right = x < gw - 1 ? S.cell(cell + 1) : kEmpty;
        float4 lu = make_float4(0, 0, 0, 0), ld = lu, ll = lu, lr = lu;
        if (use_pen) {
            if (nb_up != kEmpty) lu = gload4(job.lab4, nb_up);
            if (nb_down != kEmpty) ld = gload4(job.lab4, nb_down);
            if (nb_left != kEmpty) ll = gload4(job.lab4, nb_left);
            if (nb_right != kEmpty) lr = gload4(job.lab4, nb_right);
        }
        const uint64_t* crow = reinterpret_cast<const uint64_t*>(job.cand) + (int64_t)cell * kCandStride;

        uint64_t best = ~0ull;
        bool placeable[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int32_t r = lane + 64 * s;
            placeable[s] = false;
            if (r < K) {
                const uint64_t rec = gload(crow, r);
                const int32_t m = (int32_t)(uint32_t)rec;
                float dm = __uint_as_float((uint32_t)(rec >> 32));
                if (job.score_sqrt) dm = sqrtf(dm);
                const bool pl = m != nb_up && m != nb_down && m != nb_left && m != nb_right;
                placeable[s] = pl;
                if (pl && S.count(m) < max_usage) {
                    float pen = 0.0f;
                    if (use_pen) {
                        const float4 lm = gload4(job.lab4, m);
                        float p = 0.0f;
                        if (nb_up != kEmpty) p += adj_term(lm, lu);
                        if (nb_down != kEmpty) p += adj_term(lm, ld);
                        if (nb_left != kEmpty) p += adj_term(lm, ll);
                        if (nb_right != kEmpty) p += adj_term(lm, lr);
                        pen = p * weight;
                    }
                    const float score = dm + pen;
                    const uint64_t key = ((uint64_t)__float_as_uint(score) << 32) | (uint32_t)r;
                    best = key < best ? key : best;
                }
            }
        }
        best = wave_min_u64(best);
        int32_t rank, stage;
        if (best != ~0ull) {
            rank = (int32_t)(best & 0xffffffffu);
            stage = 0;
            ++st_main;
        } else {
            // fallback_tile_selection: reset every usage count, then the first
            // candidate passing the neighbour ban, else nearest_one.
            for (int32_t i = lane; i < n; i += 64) S.set_count(i, 0u);
            ++st_resets;
            const uint64_t b0 = wave_ballot(placeable[0]);
            const uint64_t b1 = wave_ballot(placeable[1]);
            if (b0) {
                rank = __ffsll((unsigned long long)b0) - 1;
                stage = 1;
                ++st_ban;
            } else if (b1) {
                rank = 64 + __ffsll((unsigned long long)b1) - 1;
                stage = 1;
                ++st_ban;
            } else {
                rank = 0;
                stage = 2;
                ++st_near;
            }
        }
        const uint64_t wrec = gload(crow, rank);
        const int32_t w = (int32_t)(uint32_t)wrec;
        const float wd_raw = __uint_as_float((uint32_t)(wrec >> 32));
        const float wd = job.score_sqrt ? sqrtf(wd_raw) : wd_raw;
        S.sync();
        if (lane == 0) {
            S.set_count(w, S.count(w) + 1u);
            S.set_cell(cell, w);
            if (job.stage) gstore(job.stage, cell, (uint8_t)stage);
            if (job.dist) gstore(job.dist, cell, wd);
            if (job.score)
                gstore(job.score, cell, stage == 0 ? __uint_as_float((uint32_t)(best >> 32)) : wd);
        }
        S.sync();
    }

    if constexpr (!kHbm) {
        for (int32_t i = lane; i < cells; i += 64) gstore(job.grid, i, S.grid[i]);
        for (int32_t i = lane; i < n; i += 64) gstore(job.usage, i, S.usage[i]);
    }
    if (lane == 0 && job.stats) {
        gstore(job.stats, 0, st_main + st_ban + st_near);
        gstore(job.stats, 1, st_main);
        gstore(job.stats, 2, st_ban);
        gstore(job.stats, 3, st_near);
        gstore(job.stats, 4, st_resets);
        gstore(job.stats, 5, st_empty);
    }
}

// ------------------------------------------------------------- raster chain
struct CandSet {
    int32_t m0, m1;   // material of rank lane / lane+64 (-1: no candidate)
    float d0, d1;     // sqdist, or sqrt(sqdist) for the GUI score (+inf: no candidate)
};

// Loads one cell's candidate records (ranks >= K hold the -1 sentinel
// written by the candidate kernel).  Every load is issued unconditionally so
// the compiler's memory wait counts stay exact across the unrolled group.
// Cells past the grid read the last row: the chain only looks at them as the
// next cells of the last one, whose results are never used.
__device__ __forceinline__ void load_cands(CandSet& s, const PlaceJob& job, int32_t cell,
                                           int32_t cells, int lane) {
    const int32_t row = cell < cells ? cell : cells - 1;
    const uint64_t* rec = reinterpret_cast<const uint64_t*>(job.cand) + (int64_t)row * kCandStride;
    const uint64_t r0 = gload(rec, lane), r1 = gload(rec, lane + 64);
    s.m0 = (int32_t)(uint32_t)r0;
    s.d0 = __uint_as_float((uint32_t)(r0 >> 32));
    s.m1 = (int32_t)(uint32_t)r1;
    s.d1 = __uint_as_float((uint32_t)(r1 >> 32));
}

// Candidate rows staged in LDS: the raster chain copies the rows of the next
// group of cells into one half of a double buffer while it works through the
// current group, with one global-to-LDS copy of 16 bytes per lane (1 KB, one
// cell's 128 records) per cell, so a group starts with LDS reads instead of a
// wait for global memory.
constexpr int32_t kRowBytes = kCandStride * 8;
static_assert(kRowBytes == 64 * 16, "one 16-byte copy per lane moves a row");

// Starts the copy of cell `cell`'s row (cells past the grid copy the last
// row) into `dst` (LDS, 16-byte aligned, uniform).
__device__ __forceinline__ void stage_row(uint8_t* dst, const PlaceJob& job, int32_t cell,
                                          int32_t cells, int lane) {
    const int32_t row = cell < cells ? cell : cells - 1;
    const uint8_t* src = reinterpret_cast<const uint8_t*>(job.cand) + (int64_t)row * kRowBytes +
                         16 * lane;
    __builtin_amdgcn_global_load_lds((const MH_GAS void*)src,
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// One cell's candidate records from a staged row.
__device__ __forceinline__ void read_cands(CandSet& s, const uint8_t* row, int lane) {
    const uint64_t* rec = reinterpret_cast<const uint64_t*>(row);
    const uint64_t r0 = rec[lane], r1 = rec[lane + 64];
    s.m0 = (int32_t)(uint32_t)r0;
    s.d0 = __uint_as_float((uint32_t)(r0 >> 32));
    s.m1 = (int32_t)(uint32_t)r1;
    s.d1 = __uint_as_float((uint32_t)(r1 >> 32));
}

// GUI score: the Euclidean distance replaces the squared one.
__device__ __forceinline__ void to_distance(CandSet& s) {
    s.d0 = sqrtf(s.d0);
    s.d1 = sqrtf(s.d1);
}

constexpr int32_t kNoMat = -2;    // "no previous winner" (never a material or the -1 sentinel)
// The count a -1 sentinel candidate reads (the guard word before the usage
// counts): above every usage limit, and one more use does not wrap it.
constexpr uint32_t kGuard = 0x7fffffffu;

// Byte offset of the raster chain's staged candidate rows in its LDS: after
// the lane scratch (64 words), the column row (gw words), the guard word and
// the counts (n words), rounded up to 16 bytes.
__host__ __device__ __forceinline__ size_t place_stage_offset(int32_t grid_w, int32_t n) {
    const size_t words = 64 + (size_t)grid_w + 1 + (size_t)(n > 0 ? n : 1);
    return (words * 4 + 15) & ~(size_t)15;
}

struct RasterState {
    int32_t x;         // current cell's column
    float pen2;        // 2 * weight (the penalty bound), -inf for weight <= 0
    int32_t left_m;    // previous winner (kEmpty at the start of a row)
    int32_t up_m;      // material above the current cell (kEmpty in row 0)
    uint64_t ob0;      // rank 0-63 candidates of the current cell that are usable
    // read one step ahead for the next cell: usage counts of its rank 0-63
    // candidates and its up neighbour, both as of two commits ago
    uint32_t pre_u;
    int32_t pre_up;
    int32_t lane_zero;  // 0 in every lane, unknown to the compiler (see raster_step)
    int32_t prev_w;    // winner of the previous cell (kNoMat before the first)
    int32_t st_ban, st_near, st_resets;
};

// Lane 0's LDS write targets, with every other lane sent to its own scratch
// word: a target is base + index * mul bytes, where lane 0 holds the array
// and mul = 4 and the other lanes hold their scratch word and mul = 0, so one
// 24-bit multiply-add forms each address (indices are below 2^24).
struct CommitPorts {
    uint32_t mul;                  // 4 in lane 0, else 0
    uint32_t* usage;               // usage_s (lane 0) or the lane's scratch word
    int32_t* row;                  // row_mat (lane 0) or the lane's scratch word
};

// v with lane `slot` replaced by the uniform value x.  Written out because a
// per-lane select (lane == slot ? x : v) has its lane masks hoisted out of
// the unrolled group, one 64-bit mask per slot, more than the scalar
// registers hold.
template <int kSlot>
__device__ __forceinline__ int32_t write_lane(int32_t v, int32_t x) {
    // (x is uniform; the read-first-lane only pins it to a scalar register)
    asm("v_writelane_b32 %0, %1, %2"
        : "+v"(v)
        : "s"(__builtin_amdgcn_readfirstlane(x)), "n"(kSlot));
    return v;
}

// f(Slot<0>{}), ..., f(Slot<N - 1>{}) in order, each with its index as a
// compile-time constant.
template <int I>
struct Slot {
    static constexpr int value = I;
};
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(Slot<I>{});
        static_for<I + 1, N>(f);
    }
}

// Index of the lowest set bit of a uniform mask, -1 when it is empty: the
// scalar find-first-one as is (the generic form adds a separate zero test).
__device__ __forceinline__ int32_t first_lane(uint64_t m) {
    // (m is uniform; the read-first-lanes only pin it to scalar registers)
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(m >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)m);
    const uint64_t ms = ((uint64_t)hi << 32) | lo;
    int32_t r;
    asm("s_ff1_i32_b64 %0, %1" : "=s"(r) : "s"(ms));
    // an asm result counts as per-lane; re-reading it marks it uniform
    return __builtin_amdgcn_readfirstlane(r);
}

template <typename T>
__device__ __forceinline__ T* port_at(T* base, int32_t index, uint32_t mul) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + __umul24((uint32_t)index, mul));
}

// One cell of the raster chain, in grids at least three columns wide (the
// up neighbour of the cell after next was committed at least two cells
// before it, so it can be read a step early).
//
// Speculation: the winner is usually the first usable candidate f.  Its
// material w is read first and the next cell's usable mask is built from it
// (w bans that candidate as the left neighbour and adds one use), so the
// next decision does not wait for this cell's contender check.  The check
// then confirms the speculation: candidates are in ascending (sqdist, idx)
// order and the penalty is at most fl(2w_pen) (two terms <= 1, times the
// weight), so score_f <= fl(d_f + 2 w_pen) =: bound and a candidate with
// d > bound scores strictly worse than f.  When f is the only usable rank
// 0-63 candidate within the bound and rank 64 (the nearest of ranks 64-127)
// is absent or beyond it, f wins.  Otherwise the exact full evaluation runs
// (usage counts of ranks 64-127 read from LDS, penalties for the contenders
// with the neighbours' Labs, a DPP minimum, or the usage-reset fallback) and
// the next cell's mask is rebuilt from the true winner.
//
// The cell after next has its usage counts and up neighbour read from LDS in
// this step (so the reads have a whole step to land); when they are used one
// step later they are corrected in registers for the two winners committed
// in between (+1 each, or a base of 0 after a usage reset).  Lane 0
// commits one use of the winner (LDS atomic) and the column's latest
// material; the winner itself is written into lane kSlot of a register that
// is stored to the grid once per group of cells (a vector store inside the
// group would make the next use of a group-loaded record wait for it).
template <bool kOutputs, int kSlot>
__device__ __forceinline__ void raster_step(const CandSet& s, const CandSet& sn,
                                            const CandSet& snn, int32_t c,
                                            RasterState& st, const PlaceJob& job,
                                            uint32_t* usage_s, const int32_t* row_mat,
                                            const CommitPorts& port, int32_t& wins, int32_t n,
                                            int lane) {
    const uint32_t kNone = 0xffffffffu;
    const int32_t gw = job.grid_w;
    const int32_t x = st.x;
    const bool wrap_n = x + 1 == gw;
    const int32_t xn = wrap_n ? 0 : x + 1;
    const int32_t x2 = xn + 1 == gw ? 0 : xn + 1;
    const uint32_t max_usage = (uint32_t)job.max_usage;
    // Reads for the cell after next (they see every commit up to the previous
    // cell).  A -1 sentinel reads the guard word before the counts, which is
    // above every usage limit, so the lane is never usable.
    uint32_t nnu0 = usage_s[snn.m0];
    // The up neighbour: row_mat starts all kEmpty, so row 0 needs no test.  It
    // is the same in every lane; a lane-dependent address keeps it in a
    // vector register, so the read is waited for only where the next step
    // compares against it (a uniform value would be moved to a scalar
    // register right after the read, stalling on it here).
    const int32_t nnup = row_mat[x2 + st.lane_zero];

    // The next cell's usable mask, built from lane masks.  Its counts before
    // this cell's commit are `base`; this cell's winner w adds one use to the
    // lanes holding w and bans them as the left neighbour unless the next
    // cell starts a row:
    //   usable = valid & up-ban & base < max & (lanes != w | (row start & base + 1 < max))
    // Everything but the w compare is known before w.
    const uint32_t base = st.pre_u + (sn.m0 == st.prev_w ? 1u : 0u);
    const int32_t up_next = st.pre_up;
    const uint64_t lt0 = lanes_ult(base, max_usage);
    // keep_w: base + 1 < max for max >= 1 (with max == 0, lt0 is empty and
    // masks it) ...
    // (sentinel lanes read the guard count, so lt0 excludes them)
    const uint64_t m_base = lanes_ne(sn.m0, up_next) & lt0;
    // ... and lt1 only where the next cell starts a row (base < 0 never holds)
    const uint64_t keep_w = lanes_ult(base, wrap_n ? max_usage - 1u : 0u);

    // speculative winner and the next cell's usable mask; with no usable
    // rank 0-63 candidate f is lane 63, and the check below fails
    const uint64_t ob0 = st.ob0;
    const int32_t f = first_lane(ob0) & 63;
    int32_t w = __builtin_amdgcn_readlane(s.m0, f);
    uint64_t nob = m_base & (~lanes_eq(sn.m0, w) | keep_w);

    // confirm it: every other usable rank 0-63 candidate and rank 64 (the
    // nearest of ranks 64-127) lie beyond the bound.  With a weight <= 0 the
    // bound is -inf and f wins whenever it exists; sentinel ranks carry an
    // infinite distance.  With no usable rank 0-63 candidate the bound is
    // +inf, which every rank 64-127 lane meets, so the check fails.
    const uint64_t rest = ob0 & (ob0 - 1ull);  // usable ranks after f
    const float pen = ob0 ? st.pen2 : __builtin_inff();
    const float bound = readlane_f(s.d0, f) + pen;
    uint64_t viol = (rest & lanes_fle(s.d0, bound)) | lanes_fle(s.d1, bound);
    if (kOutputs && job.score) viol = 1;

    int32_t rank = f, stage = 0;
    uint32_t win_score = 0;
    if (__builtin_expect(viol != 0, 0)) {
        // every cell gets a tile, so a filled up neighbour means row > 0
        const int32_t up_m = st.up_m;
        const int32_t left_m = x > 0 ? st.left_m : kEmpty;
        const bool pl0 = s.m0 >= 0 && s.m0 != up_m && s.m0 != left_m;
        const bool ok0 = lane_in(ob0);
        const uint32_t u1 = usage_s[s.m1];  // a -1 sentinel reads the guard word
        const bool pl1 = s.m1 >= 0 && s.m1 != up_m && s.m1 != left_m;
        const bool ok1 = pl1 && u1 < max_usage;
        const uint64_t ob1 = wave_ballot(ok1);
        if (ob0 | ob1) {
            const int32_t f2 = ob0 ? __ffsll((unsigned long long)ob0) - 1
                                   : 64 + __ffsll((unsigned long long)ob1) - 1;
            const float d_f = f2 < 64 ? readlane_f(s.d0, f2) : readlane_f(s.d1, f2 & 63);
            rank = f2;
            if (job.weight > 0.0f) {
                const float bound = d_f + 2.0f * job.weight;
                const bool ct0 = ok0 && s.d0 <= bound;
                const bool ct1 = ok1 && s.d1 <= bound;
                const uint64_t cb0 = wave_ballot(ct0), cb1 = wave_ballot(ct1);
                if (__popcll(cb0) + __popcll(cb1) > 1 || (kOutputs && job.score)) {
                    // calculate_penalty for the contenders: Up then Left (Down
                    // and Right are still empty in raster order)
                    const float4 up_lab =
                        up_m >= 0 ? gload4(job.lab4, up_m) : make_float4(0, 0, 0, 0);
                    const float4 left_lab =
                        x > 0 ? gload4(job.lab4, left_m) : make_float4(0, 0, 0, 0);
                    // the contenders' Labs (other lanes read material 0,
                    // whose value they do not use)
                    const float4 lab0 = gload4(job.lab4, ct0 ? s.m0 : 0);
                    const float4 lab1 = gload4(job.lab4, ct1 ? s.m1 : 0);
                    uint32_t k0 = kNone, k1 = kNone;
                    if (cb0) {
                        float p = 0.0f;
                        if (up_m >= 0) p += adj_term(lab0, up_lab);
                        if (x > 0) p += adj_term(lab0, left_lab);
                        const float sc = s.d0 + p * job.weight;
                        k0 = ct0 ? __float_as_uint(sc) : kNone;
                    }
                    if (cb1) {
                        float p = 0.0f;
                        if (up_m >= 0) p += adj_term(lab1, up_lab);
                        if (x > 0) p += adj_term(lab1, left_lab);
                        const float sc = s.d1 + p * job.weight;
                        k1 = ct1 ? __float_as_uint(sc) : kNone;
                    }
                    const uint32_t kmin = wave_min_u32(k0 < k1 ? k0 : k1);
                    const uint64_t b0 = wave_ballot(k0 == kmin);
                    rank = b0 ? __ffsll((unsigned long long)b0) - 1
                              : 64 + __ffsll((unsigned long long)wave_ballot(k1 == kmin)) - 1;
                    win_score = kmin;
                } else {
                    win_score = __float_as_uint(d_f);  // sole contender: no penalty needed
                }
            } else {
                win_score = __float_as_uint(d_f);  // penalty 0: scores are the distances
            }
        } else {
            // fallback_tile_selection: reset every count, first placeable, else nearest_one
            for (int32_t i = lane; i < n; i += 64) usage_s[i] = 0u;
            ++st.st_resets;
            // the counts read ahead for the cell after next predate the reset
            // (sentinel lanes keep the guard)
            nnu0 = nnu0 == kGuard ? kGuard : 0u;
            const uint64_t b0 = wave_ballot(pl0);
            const uint64_t b1 = wave_ballot(pl1);
            if (b0) {
                rank = __ffsll((unsigned long long)b0) - 1;
                stage = 1;
                ++st.st_ban;
            } else if (b1) {
                rank = 64 + __ffsll((unsigned long long)b1) - 1;
                stage = 1;
                ++st.st_ban;
            } else {
                rank = 0;
                stage = 2;
                ++st.st_near;
            }
        }
        w = rank < 64 ? __builtin_amdgcn_readlane(s.m0, rank & 63)
                      : __builtin_amdgcn_readlane(s.m1, rank & 63);
        const uint32_t h = sn.m0 == w ? 1u : 0u;
        const uint32_t nu = stage == 0 ? base + h : h;  // a reset zeroed every count
        nob = wave_ballot(sn.m0 >= 0 && sn.m0 != up_next && (xn == 0 || sn.m0 != w) &&
                          nu < max_usage);
        // Wait here for this path's neighbour-Lab gathers (vmcnt 0; expcnt and
        // lgkmcnt untouched).  Left pending, they would make every later step
        // wait for all memory reads in flight, since the fast path cannot tell
        // them from the group's candidate loads.
        __builtin_amdgcn_s_waitcnt(0x0f70);
    }
    // Lane 0 commits; every other lane writes its own scratch word instead,
    // so the commit is straight-line code with no branch on the lane (the
    // step stays one block for the scheduler).  A per-lane address also
    // keeps the add a plain ds_add.
    atomicAdd(port_at(port.usage, w, port.mul), 1u);
    *port_at(port.row, x, port.mul) = w;
    wins = write_lane<kSlot>(wins, w);  // the winner waits in lane kSlot
    if (kOutputs) {
        const float dw = rank < 64 ? readlane_f(s.d0, rank & 63) : readlane_f(s.d1, rank & 63);
        if (lane == 0) {
            if (job.stage) gstore(job.stage, c, (uint8_t)stage);
            if (job.dist) gstore(job.dist, c, dw);
            if (job.score) gstore(job.score, c, stage == 0 ? __uint_as_float(win_score) : dw);
        }
    }
    __builtin_amdgcn_wave_barrier();
    st.pre_u = nnu0;
    st.pre_up = nnup;
    st.prev_w = w;
    st.up_m = up_next;
    st.left_m = w;
    st.ob0 = nob;
    st.x = xn;
}

// Raster order in grids one or two columns wide: the up neighbour is the
// winner one (one column) or two (two columns) cells back and the left
// neighbour the previous winner in the second column.  One cell at a time
// with every usable candidate scored -- the same decisions as raster_step,
// without its read-ahead of the up neighbour, which needs three columns.
struct NarrowStats {
    int32_t ban, near, resets;
};

template <bool kOutputs>
__device__ __forceinline__ NarrowStats raster_narrow(const PlaceJob& job, bool sq,
                                                     uint32_t* usage_s, int lane) {
    NarrowStats st{0, 0, 0};
    const int32_t gw = job.grid_w, cells = gw * job.grid_h;
    const uint32_t max_usage = (uint32_t)job.max_usage;
    const float weight = job.weight;
    int32_t w1 = kEmpty, w2 = kEmpty;  // winners one and two cells back
    for (int32_t c = 0; c < cells; ++c) {
        CandSet s;
        load_cands(s, job, c, cells, lane);
        if (sq) to_distance(s);
        const int32_t up_m = c >= gw ? (gw == 1 ? w1 : w2) : kEmpty;
        const int32_t left_m = gw == 2 && (c & 1) ? w1 : kEmpty;
        const bool pl0 = s.m0 >= 0 && s.m0 != up_m && s.m0 != left_m;
        const bool pl1 = s.m1 >= 0 && s.m1 != up_m && s.m1 != left_m;
        const bool ok0 = pl0 && usage_s[s.m0 >= 0 ? s.m0 : 0] < max_usage;
        const bool ok1 = pl1 && usage_s[s.m1 >= 0 ? s.m1 : 0] < max_usage;
        float sc0 = s.d0, sc1 = s.d1;
        if (weight > 0.0f) {
            // calculate_penalty: Up then Left (Down and Right are still empty)
            const float4 lab0 = gload4(job.lab4, s.m0 >= 0 ? s.m0 : 0);
            const float4 lab1 = gload4(job.lab4, s.m1 >= 0 ? s.m1 : 0);
            float p0 = 0.0f, p1 = 0.0f;
            if (up_m >= 0) {
                const float4 u = gload4(job.lab4, up_m);
                p0 += adj_term(lab0, u);
                p1 += adj_term(lab1, u);
            }
            if (left_m >= 0) {
                const float4 l = gload4(job.lab4, left_m);
                p0 += adj_term(lab0, l);
                p1 += adj_term(lab1, l);
            }
            sc0 = s.d0 + p0 * weight;
            sc1 = s.d1 + p1 * weight;
        }
        // first strict-< winner: the minimum of (score, rank)
        const uint64_t k0 = ok0 ? ((uint64_t)__float_as_uint(sc0) << 32) | (uint32_t)lane : ~0ull;
        const uint64_t k1 =
            ok1 ? ((uint64_t)__float_as_uint(sc1) << 32) | (uint32_t)(lane + 64) : ~0ull;
        const uint64_t best = wave_min_u64(k0 < k1 ? k0 : k1);
        int32_t rank, stage = 0;
        if (best != ~0ull) {
            rank = (int32_t)(uint32_t)best;
        } else {
            // fallback_tile_selection: reset every count, first placeable, else nearest_one
            for (int32_t i = lane; i < job.n; i += 64) usage_s[i] = 0u;
            ++st.resets;
            const uint64_t b0 = wave_ballot(pl0);
            const uint64_t b1 = wave_ballot(pl1);
            if (b0) {
                rank = __ffsll((unsigned long long)b0) - 1;
                stage = 1;
                ++st.ban;
            } else if (b1) {
                rank = 64 + __ffsll((unsigned long long)b1) - 1;
                stage = 1;
                ++st.ban;
            } else {
                rank = 0;
                stage = 2;
                ++st.near;
            }
        }
        const int32_t w = rank < 64 ? __builtin_amdgcn_readlane(s.m0, rank & 63)
                                    : __builtin_amdgcn_readlane(s.m1, rank & 63);
        const float dw = rank < 64 ? readlane_f(s.d0, rank & 63) : readlane_f(s.d1, rank & 63);
        __syncthreads();
        if (lane == 0) {
            usage_s[w] += 1u;
            gstore(job.grid, c, w);
            if (kOutputs) {
                if (job.stage) gstore(job.stage, c, (uint8_t)stage);
                if (job.dist) gstore(job.dist, c, dw);
                if (job.score)
                    gstore(job.score, c, stage == 0 ? __uint_as_float((uint32_t)(best >> 32)) : dw);
            }
        }
        __syncthreads();
        w2 = w1;
        w1 = w;
    }
    return st;
}

// The look-ahead chain over a grid at least three columns wide (kSq: the
// GUI score, whose distances are square-rooted as the records are used;
// kStaged: the candidate rows come through `stage` in LDS).  The staged and
// unstaged chains are compiled apart: merged, the compiler's wait counts
// for the records would also wait for the staging copies in flight.
template <bool kOutputs, bool kSq, bool kStaged>
__device__ __forceinline__ void raster_wide(const PlaceJob& job, uint32_t* usage_s,
                                            int32_t* row_mat, int32_t* lane_scratch,
                                            uint8_t* stage, RasterState& st, int lane) {
    const int32_t gw = job.grid_w, n = job.n;
    const int32_t cells = gw * job.grid_h;
    CommitPorts port;
    port.mul = lane == 0 ? 4u : 0u;
    port.usage = lane == 0 ? usage_s : reinterpret_cast<uint32_t*>(lane_scratch);
    port.row = lane == 0 ? row_mat : lane_scratch;
    // Candidate records for kGroup cells (and the two after them) are read
    // at the top of each group, from the LDS half the previous group filled
    // (stage_row) while the next group's rows are requested into the other
    // half, or (unstaged: LDS too small) loaded from global memory with
    // one memory latency per group.  The records are consumed in the same
    // loop iteration, so the compiler's wait counts stay exact (a load
    // consumed after the loop's back edge would make it wait for every load
    // in flight).  The previous group's winners are stored just after the
    // reads.  24 cells per group (140 VGPRs, none spilled); 12 cells cost 6%
    // more chain time.
    constexpr int kGroup = 24;
    constexpr int32_t kHalf = (kGroup + 2) * kRowBytes;
    static_assert(kGroup <= 64, "a group's winners wait in one lane each");
    CandSet q[kGroup + 2];
    st.x = 0;
    st.pen2 = job.weight > 0.0f ? 2.0f * job.weight : -__builtin_inff();
    st.left_m = kEmpty;
    st.up_m = kEmpty;
    st.lane_zero = job.n < 0 ? lane : 0;  // n > 0 here
    st.prev_w = kNoMat;
    {
        CandSet h0, h1;
        load_cands(h0, job, 0, cells, lane);
        load_cands(h1, job, 1, cells, lane);
        const uint32_t u0 = usage_s[h0.m0 >= 0 ? h0.m0 : 0];
        st.ob0 = wave_ballot(h0.m0 >= 0 && u0 < (uint32_t)job.max_usage);
        st.pre_u = usage_s[h1.m0];  // a sentinel reads the guard
        st.pre_up = kEmpty;  // cell 1 is in row 0
    }
    // Winners of the group starting at g0 (at most kGroup cells) to the
    // grid: cell g0 + i's winner waits in lane i of `wins`.
    int32_t wins = kEmpty;
    auto flush = [&](int32_t g0) {
        if (lane < kGroup && g0 + lane < cells) gstore(job.grid, g0 + lane, wins);
    };
    // rows of the group at cell g0 into half h
    auto stage_group = [&](int32_t g0, int32_t h) {
#pragma unroll
        for (int i = 0; i < kGroup + 2; ++i)
            stage_row(stage + h * kHalf + i * kRowBytes, job, g0 + i, cells, lane);
    };
    // the rows of the group at cell g0: from half h (waiting for their
    // copies first), or straight from global memory when nothing is staged
    auto read_group = [&](int32_t g0, int32_t h) {
        if constexpr (kStaged) {
            __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): the half's copies have landed
#pragma unroll
            for (int i = 0; i < kGroup + 2; ++i)
                read_cands(q[i], stage + h * kHalf + i * kRowBytes, lane);
        } else {
#pragma unroll
            for (int i = 0; i < kGroup + 2; ++i) load_cands(q[i], job, g0 + i, cells, lane);
        }
    };
    int32_t c = 0, half = 0;
    if (kStaged) stage_group(0, 0);
    for (; c + kGroup <= cells; c += kGroup) {
        read_group(c, half);
        if (c > 0) flush(c - kGroup);
        half ^= 1;
        if (kStaged && c + kGroup < cells) stage_group(c + kGroup, half);
        if (kSq) to_distance(q[0]);
        static_for<0, kGroup>([&](auto slot) __attribute__((always_inline)) {
            constexpr int i = decltype(slot)::value;
            raster_step<kOutputs, i>(q[i], q[i + 1], q[i + 2], c + i, st, job, usage_s,
                                     row_mat, port, wins, n, lane);
            // the step waited for the records after q[i + 1]'s, so these are in
            if (kSq && i + 1 < kGroup) to_distance(q[i + 1]);
        });
    }
    int32_t last = c - kGroup;  // start of the last group run (negative: none)
    if (c < cells) {
        // the last, partial group (its rows were staged by the previous
        // group or before the loop), with the steps past the grid skipped
        read_group(c, half);
        if (c > 0) flush(c - kGroup);
        const int32_t rem = cells - c;
        if (kSq) to_distance(q[0]);
        static_for<0, kGroup>([&](auto slot) __attribute__((always_inline)) {
            constexpr int i = decltype(slot)::value;
            if (i < rem) {
                raster_step<kOutputs, i>(q[i], q[i + 1], q[i + 2], c + i, st, job,
                                         usage_s, row_mat, port, wins, n, lane);
                if (kSq && i + 1 < kGroup) to_distance(q[i + 1]);
            }
        });
        last = c;
    }
    if (last >= 0) flush(last);
    if (kStaged) __builtin_amdgcn_s_waitcnt(0x0f70);  // nothing staged is left in flight
}

// kScore: kScoreSq / kScoreDist fix the score for every job of the launch,
// kScoreMixed reads each job's score_sqrt.
template <bool kOutputs, int kScore>
__global__ __launch_bounds__(64) void place_raster_kernel(PlaceJob single,
                                                          const PlaceJob* __restrict__ jobs,
                                                          int32_t staged) {
    extern __shared__ __align__(16) uint32_t smr[];
    const PlaceJob job = uniform_copy(jobs ? jobs[blockIdx.x] : single);
    const bool sq = kScore == kScoreMixed ? job.score_sqrt != 0 : kScore == kScoreDist;
    const int lane = threadIdx.x;
    const int32_t gw = job.grid_w, n = job.n;
    const int32_t cells = gw * job.grid_h;
    int32_t* lane_scratch = reinterpret_cast<int32_t*>(smr) + lane;  // 64 words, one per lane
    int32_t* row_mat = reinterpret_cast<int32_t*>(smr) + 64;
    // one guard word between the column row and the counts: a -1 sentinel
    // candidate reads it as its count
    uint32_t* usage_s = reinterpret_cast<uint32_t*>(row_mat + gw + 1);
    // the staged candidate rows after the counts, 16-byte aligned (when the
    // launch holds them)
    uint8_t* stage =
        staged ? reinterpret_cast<uint8_t*>(smr) + place_stage_offset(gw, n) : nullptr;

    if (n == 0) {  // src/main.rs:342-345: no tiles, every cell stays empty
        for (int32_t i = lane; i < cells; i += 64) gstore(job.grid, i, kEmpty);
        if (lane == 0 && job.stats) {
            for (int i = 0; i < 5; ++i) gstore(job.stats, i, 0);
            gstore(job.stats, 5, cells);
        }
        return;
    }
    for (int32_t i = lane; i < n; i += 64) usage_s[i] = job.zero_usage ? 0u : gload(job.usage, i);
    for (int32_t i = lane; i < gw; i += 64) row_mat[i] = kEmpty;
    if (lane == 0) usage_s[-1] = kGuard;
    __syncthreads();

    RasterState st;
    st.st_ban = st.st_near = st.st_resets = 0;
    if (gw < 3) {
        const NarrowStats ns = raster_narrow<kOutputs>(job, sq, usage_s, lane);
        st.st_ban = ns.ban;
        st.st_near = ns.near;
        st.st_resets = ns.resets;
    } else {
        // (kScoreMixed picks a compiled chain per job: a run-time test per
        // step costs the chain scalar registers it does not have)
        constexpr bool kDist = kScore == kScoreDist;
        const bool dist = kScore == kScoreMixed && job.score_sqrt;
        if (stage && dist)
            raster_wide<kOutputs, true, true>(job, usage_s, row_mat, lane_scratch, stage, st, lane);
        else if (stage)
            raster_wide<kOutputs, kDist, true>(job, usage_s, row_mat, lane_scratch, stage, st, lane);
        else if (dist)
            raster_wide<kOutputs, true, false>(job, usage_s, row_mat, lane_scratch, stage, st, lane);
        else
            raster_wide<kOutputs, kDist, false>(job, usage_s, row_mat, lane_scratch, stage, st, lane);
    }
    for (int32_t i = lane; i < n; i += 64) gstore(job.usage, i, usage_s[i]);
    if (lane == 0 && job.stats) {
        gstore(job.stats, 0, cells);
        gstore(job.stats, 1, cells - st.st_ban - st.st_near);
        gstore(job.stats, 2, st.st_ban);
        gstore(job.stats, 3, st.st_near);
        gstore(job.stats, 4, st.st_resets);
        gstore(job.stats, 5, 0);
    }
}

size_t place_lds_bytes(int32_t cells, int32_t n) {
    return sizeof(int32_t) * ((size_t)cells + (size_t)n);
}

size_t place_raster_lds_bytes(int32_t grid_w, int32_t n, bool staged) {
    // lane scratch, the column row, the guard word, the counts, then two
    // halves of staged candidate rows (24 + 2 cells each)
    if (staged) return place_stage_offset(grid_w, n) + 2 * (size_t)(24 + 2) * kRowBytes;
    return sizeof(int32_t) * (64 + (size_t)grid_w + 1) +
           sizeof(uint32_t) * (size_t)(n > 0 ? n : 1);
}

void launch_place_single(const PlaceJob& job, bool raster, bool staged, size_t lds_bytes,
                         hipStream_t stream) {
    const bool outputs = job.stage || job.dist || job.score;
    const PlaceJob* none = nullptr;
    if (raster && outputs && job.score_sqrt)
        hipLaunchKernelGGL((place_raster_kernel<true, kScoreDist>), dim3(1), dim3(64), lds_bytes,
                           stream, job, none, (int32_t)staged);
    else if (raster && outputs)
        hipLaunchKernelGGL((place_raster_kernel<true, kScoreSq>), dim3(1), dim3(64), lds_bytes,
                           stream, job, none, (int32_t)staged);
    else if (raster && job.score_sqrt)
        hipLaunchKernelGGL((place_raster_kernel<false, kScoreDist>), dim3(1), dim3(64), lds_bytes,
                           stream, job, none, (int32_t)staged);
    else if (raster)
        hipLaunchKernelGGL((place_raster_kernel<false, kScoreSq>), dim3(1), dim3(64), lds_bytes,
                           stream, job, none, (int32_t)staged);
    else if (lds_bytes > 0)
        hipLaunchKernelGGL(place_chain_kernel<false>, dim3(1), dim3(64), lds_bytes, stream, job,
                           (const PlaceJob*)nullptr);
    else
        hipLaunchKernelGGL(place_chain_kernel<true>, dim3(1), dim3(64), 0, stream, job,
                           (const PlaceJob*)nullptr);
}

void launch_place_batch(const PlaceJob* d_jobs, int32_t n_jobs, bool raster, bool staged,
                        int32_t score_kind, size_t lds_bytes, hipStream_t stream) {
    if (n_jobs <= 0) return;
    PlaceJob unused{};
    if (raster && score_kind == kScoreMixed)
        hipLaunchKernelGGL((place_raster_kernel<false, kScoreMixed>), dim3(n_jobs), dim3(64),
                           lds_bytes, stream, unused, d_jobs, (int32_t)staged);
    else if (raster && score_kind == kScoreDist)
        hipLaunchKernelGGL((place_raster_kernel<false, kScoreDist>), dim3(n_jobs), dim3(64),
                           lds_bytes, stream, unused, d_jobs, (int32_t)staged);
    else if (raster)
        hipLaunchKernelGGL((place_raster_kernel<false, kScoreSq>), dim3(n_jobs), dim3(64),
                           lds_bytes, stream, unused, d_jobs, (int32_t)staged);
    else if (lds_bytes > 0)
        hipLaunchKernelGGL(place_chain_kernel<false>, dim3(n_jobs), dim3(64), lds_bytes, stream,
                           unused, d_jobs);
    else
        hipLaunchKernelGGL(place_chain_kernel<true>, dim3(n_jobs), dim3(64), 0, stream, unused,
                           d_jobs);
}

hipError_t place_configure(int max_lds) {
    const void* kernels[] = {
        reinterpret_cast<const void*>(&place_chain_kernel<false>),
        reinterpret_cast<const void*>(&place_raster_kernel<true, kScoreSq>),
        reinterpret_cast<const void*>(&place_raster_kernel<true, kScoreDist>),
        reinterpret_cast<const void*>(&place_raster_kernel<false, kScoreSq>),
        reinterpret_cast<const void*>(&place_raster_kernel<false, kScoreDist>),
        reinterpret_cast<const void*>(&place_raster_kernel<false, kScoreMixed>),
    };
    for (const void* k : kernels) {
        const hipError_t e = raise_dynamic_lds(k, max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mh
