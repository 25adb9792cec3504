// k_anneal.hip -- adjacency cost, swap deltas and simulated annealing
// (AdjacencyPenaltyCalculator::calculate_total_cost / calculate_swap_delta /
// calculate_penalty, src/adjacency.rs:63-214; MosaicOptimizer::optimize and
// optimize_greedy, src/optimizer.rs:41-198).
//
// Annealing keeps the reference's one-proposal-at-a-time semantics exactly
// while evaluating 64 proposals per round on one wave:
//   * proposal `it` draws x1,y1,x2,y2,u from a counter RNG keyed by (seed,it),
//     so any proposal can be evaluated out of order;
//   * skipped proposals (same cell / empty cell) do not cool, and emptiness
//     never changes under swaps, so proposal j's temperature index is the
//     number of non-skipped proposals before it (a ballot prefix count) and
//     T comes from a table built by the sequential f32 recurrence T *= decay;
//   * every lane computes its delta on the round-start grid; accepted lanes
//     stamp the two cells they would write (LDS atomicMax of
//     round<<6 | 63-lane, i.e. the lowest writer wins); a lane whose read set
//     (both cells and their 4-neighbours) holds a stamp from a lower lane is a
//     conflict.  Lanes before the first conflict read exactly what the serial
//     chain would have read, so they commit; the rest are re-drawn next round.
//   * the running cost is accumulated in proposal order (f32, as the
//     reference's `current_cost += delta`).
// Result: grid, costs and counts are bit-identical to the serial chain.
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

// -------------------------------------------------------------- helpers
template <typename G>
__device__ __forceinline__ int32_t gget(const G* g, int32_t i) {
    return (int32_t)g[i];
}
template <typename G>
__device__ __forceinline__ void gput(G* g, int32_t i, int32_t v) {
    g[i] = (G)v;
}

// A grid kept in HBM (annealing of grids too large for LDS).  Every access
// is an agent-scope atomic, so a read never sees a stale cache line of a
// cell another lane of the wave has just written.
struct HbmGrid {
    int32_t* p;
};
__device__ __forceinline__ int32_t gget(const HbmGrid* g, int32_t i) {
    return __hip_atomic_load((MH_GAS int32_t*)g->p + i, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gput(HbmGrid* g, int32_t i, int32_t v) {
    __hip_atomic_store((MH_GAS int32_t*)g->p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One adjacency term 1/(sim(a,b)+1): a gather from the library's term
// matrix when it is resident, else computed from the Lab table (same bits).
// The source is a type rather than a run-time test inside every read, so a
// caller's gathers are straight-line code that the compiler issues together
// (behind a test per read, each gather was waited for on its own).
struct TermMatrix {
    const float* term;
    int32_t n;
    __device__ __forceinline__ float operator()(int32_t a, int32_t b) const {
        return gload(term, (int64_t)a * n + b);
    }
};
struct TermLab {
    const float4* lab4;
    __device__ __forceinline__ float operator()(int32_t a, int32_t b) const {
        return adj_term(gload4(lab4, a), gload4(lab4, b));
    }
};

// calculate_swap_delta (src/adjacency.rs:133-214) on a grid in any address
// space.  Term order: neighbours of p1 (Up, Down, Left, Right, skipping p2),
// then of p2 (skipping p1), then the mutual term when adjacent.
//
// Branch-free form: the eight neighbour slots are read first, then every
// term of both rows (m1 and m2) is requested at once -- an absent slot reads
// a valid dummy entry -- so the gathers cost one memory latency instead of
// one each.  Absent terms are added as +0.0: the sums start at +0 and only
// grow, and x + (+0) == x bit for bit, so this equals skipping them.
//
// The neighbour slots (Up, Down, Left, Right of p1, then of p2) are read by
// swap_slots, which callers may issue together with their own reads of p1
// and p2; a slot outside the grid reads its own cell, always in bounds, and
// swap_delta_core masks it out.
template <typename G>
__device__ __forceinline__ void swap_slots(const G* grid, int32_t gw, int32_t gh, int32_t x1,
                                           int32_t y1, int32_t x2, int32_t y2, int32_t nb[8]) {
    const int32_t p1 = y1 * gw + x1, p2 = y2 * gw + x2;
    nb[0] = gget(grid, y1 > 0 ? p1 - gw : p1);
    nb[1] = gget(grid, y1 < gh - 1 ? p1 + gw : p1);
    nb[2] = gget(grid, x1 > 0 ? p1 - 1 : p1);
    nb[3] = gget(grid, x1 < gw - 1 ? p1 + 1 : p1);
    nb[4] = gget(grid, y2 > 0 ? p2 - gw : p2);
    nb[5] = gget(grid, y2 < gh - 1 ? p2 + gw : p2);
    nb[6] = gget(grid, x2 > 0 ? p2 - 1 : p2);
    nb[7] = gget(grid, x2 < gw - 1 ? p2 + 1 : p2);
}

// swap_delta_core in two parts, so that a caller can do other work while
// the gathers are in flight: swap_terms requests every term (m1 and m2 must
// be material ids), swap_sum adds them in the reference's order.
struct SwapTerms {
    float r1[8], r2[8];  // T(m1, a_i), T(m2, a_i)
    float tm;            // T(m1, m2)
    bool ok[8];          // slot i holds a neighbour that is not the partner
    bool is_adj;
};

template <typename TS>
__device__ __forceinline__ void swap_terms(int32_t gw, int32_t gh, const TS& T, int32_t x1,
                                           int32_t y1, int32_t x2, int32_t y2, int32_t m1,
                                           int32_t m2, const int32_t nb[8], SwapTerms& st) {
    const int32_t p1 = y1 * gw + x1, p2 = y2 * gw + x2;
    const bool in[8] = {y1 > 0, y1 < gh - 1, x1 > 0, x1 < gw - 1,
                        y2 > 0, y2 < gh - 1, x2 > 0, x2 < gw - 1};
    const int32_t pos[8] = {p1 - gw, p1 + gw, p1 - 1, p1 + 1, p2 - gw, p2 + gw, p2 - 1, p2 + 1};
    int32_t a[8];
    st.is_adj = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int32_t other = i < 4 ? p2 : p1;
        const bool self = in[i] && pos[i] == other;  // the swap partner
        if (i < 4) st.is_adj = st.is_adj || self;
        const int32_t g = in[i] ? nb[i] : kEmpty;
        st.ok[i] = in[i] && !self && g != kEmpty;
        a[i] = st.ok[i] ? g : 0;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        st.r1[i] = T(m1, a[i]);
        st.r2[i] = T(m2, a[i]);
    }
    st.tm = T(m1, m2);
}

__device__ __forceinline__ float swap_sum(const SwapTerms& st) {
    float old_cost = 0.0f, new_cost = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // p1's neighbours: old with m1, new with m2
        old_cost += st.ok[i] ? st.r1[i] : 0.0f;
        new_cost += st.ok[i] ? st.r2[i] : 0.0f;
    }
#pragma unroll
    for (int i = 4; i < 8; ++i) {  // p2's neighbours: old with m2, new with m1
        old_cost += st.ok[i] ? st.r2[i] : 0.0f;
        new_cost += st.ok[i] ? st.r1[i] : 0.0f;
    }
    old_cost += st.is_adj ? st.tm : 0.0f;
    new_cost += st.is_adj ? st.tm : 0.0f;
    return new_cost - old_cost;
}

template <typename TS>
__device__ __forceinline__ float swap_delta_core(int32_t gw, int32_t gh, const TS& T, int32_t x1,
                                                 int32_t y1, int32_t x2, int32_t y2, int32_t m1,
                                                 int32_t m2, const int32_t nb[8]) {
    if (m1 == kEmpty || m2 == kEmpty) return 0.0f;
    if (m1 == m2) return 0.0f;
    SwapTerms st;
    swap_terms(gw, gh, T, x1, y1, x2, y2, m1, m2, nb, st);
    return swap_sum(st);
}

template <typename G, typename TS>
__device__ float swap_delta_dev(const G* grid, int32_t gw, int32_t gh, const TS& T,
                                int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
    int32_t nb[8];
    swap_slots(grid, gw, gh, x1, y1, x2, y2, nb);
    return swap_delta_core(gw, gh, T, x1, y1, x2, y2, gget(grid, y1 * gw + x1),
                           gget(grid, y2 * gw + x2), nb);
}

// calculate_total_cost (src/adjacency.rs:91-130) by one wave: the terms of
// `chunk` cells at a time are computed in parallel into scratch (right term,
// down term per cell), then lane 0 adds them in the reference's order
// (raster, right before down).  An absent term is stored as +0.0: the running
// sum starts at +0 and only grows, and x + (+0) == x bit for bit, so adding
// it leaves the sum exactly as the reference's skipped addition does.
// scratch holds 2*chunk + 2 floats and is 16-byte aligned.
// The result is broadcast to every lane.
template <typename G, typename TS>
__device__ float total_cost_wave(const G* grid, int32_t gw, int32_t gh, const TS& T,
                                 float* scratch, int32_t chunk) {
    const int lane = threadIdx.x;
    const int32_t cells = gw * gh;
    float total = 0.0f;
    for (int32_t base = 0; base < cells; base += chunk) {
        const int32_t len = min(chunk, cells - base);
        for (int32_t i = lane; i < len; i += 64) {
            const int32_t c = base + i;
            const int32_t x = c % gw, y = c / gw;
            const int32_t m = gget(grid, c);
            float tr = 0.0f, td = 0.0f;
            if (m != kEmpty) {
                if (x < gw - 1) {
                    const int32_t r = gget(grid, c + 1);
                    if (r != kEmpty) tr = T(m, r);
                }
                if (y < gh - 1) {
                    const int32_t d = gget(grid, c + gw);
                    if (d != kEmpty) td = T(m, d);
                }
            }
            scratch[2 * i] = tr;
            scratch[2 * i + 1] = td;
        }
        if ((len & 1) != 0 && lane == 0) {  // pad to a whole float4
            scratch[2 * len] = 0.0f;
            scratch[2 * len + 1] = 0.0f;
        }
        __syncthreads();
        if (lane == 0) {
            const float4* v = reinterpret_cast<const float4*>(scratch);
            const int32_t n4 = (2 * len + 3) / 4;
            for (int32_t i = 0; i < n4; ++i) {
                const float4 t = v[i];
                total = total + t.x;
                total = total + t.y;
                total = total + t.z;
                total = total + t.w;
            }
        }
        __syncthreads();
    }
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(total)));
}

// ---------------------------------------------------------- annealing
// Stamp words per job: one per cell plus slack, a multiple of 4 (the words
// double as the cost scratch: 2 floats per cell of a chunk + 2 of padding).
__host__ __device__ __forceinline__ int32_t stamp_words(int32_t cells) {
    return (cells + 8) & ~3;
}
// One word per lane for packing a round's accepted deltas.
constexpr int32_t kPackWords = 64;
// Proposal draws, made ahead in blocks of 128 iterations (two per lane) into
// a ring of 256 slots of three words: x1 | y1 << 16, x2 | y2 << 16, u (both
// sides are below 2^16: a grid in LDS is far smaller, and the HBM variant
// requires it).
constexpr int32_t kRing = 256;
constexpr int32_t kRingWords = 3 * kRing;
// Stamp table of the HBM variant: cell p stamps word p mod kHashStamps.  Two
// cells sharing a word can only make a lane see a conflict that is not there,
// which ends the round early; the commits stay those of the serial chain.
constexpr int32_t kHashStamps = 16384;

// kHbm: the grid stays in HBM (the job's own grid, swapped in place) and the
// stamps are hashed into kHashStamps LDS words, for grids too large for LDS.
template <typename G, bool kHbm>
__global__ __launch_bounds__(64) void anneal_kernel(AnnealJob single,
                                                    const AnnealJob* __restrict__ jobs) {
    extern __shared__ __align__(16) uint32_t sm[];
    const AnnealJob job = uniform_copy(jobs ? jobs[blockIdx.x] : single);
    const int lane = threadIdx.x;
    const int32_t gw = job.grid_w, gh = job.grid_h;
    const int32_t cells = gw * gh;
    const int32_t stamps = kHbm ? kHashStamps : stamp_words(cells);
    const int32_t stamp_cells = kHbm ? kHashStamps : cells;  // words cleared between uses
    uint32_t* stamp = sm;
    auto sidx = [&](int32_t p) { return kHbm ? (p & (kHashStamps - 1)) : p; };
    float* pack_s = reinterpret_cast<float*>(sm + stamps);
    uint32_t* ring_xy1 = sm + stamps + kPackWords;
    uint32_t* ring_xy2 = ring_xy1 + kRing;
    float* ring_u = reinterpret_cast<float*>(ring_xy2 + kRing);
    HbmGrid hbm{job.grid};
    auto grid_s = [&]() {
        if constexpr (kHbm)
            return &hbm;
        else
            return reinterpret_cast<G*>(sm + stamps + kPackWords + kRingWords);
    }();

    if constexpr (!kHbm)
        for (int32_t i = lane; i < cells; i += 64) grid_s[i] = (G)gload(job.grid, i);
    __syncthreads();

    // the stamp words double as the cost scratch (2 terms per cell + padding)
    const TermMatrix TM{job.term, job.n};
    const TermLab TL{job.lab4};
    float* cost_scratch = reinterpret_cast<float*>(stamp);
    const int32_t chunk = (stamps / 2 - 1) & ~1;
    float cost = job.term ? total_cost_wave(grid_s, gw, gh, TM, cost_scratch, chunk)
                          : total_cost_wave(grid_s, gw, gh, TL, cost_scratch, chunk);
    for (int32_t i = lane; i < stamp_cells; i += 64) stamp[i] = 0u;
    __syncthreads();

    const float initial = cost;
    float best = cost;
    int64_t improved = 0, accepted = 0, evaluated = 0, rounds = 0;
    int64_t s = 0, kbase = 0;
    uint32_t tag = 1;
    const int64_t max_iter = job.max_iter;

    // Proposal `it` is drawn once, in a block of 128 made by the whole wave
    // (refill) a round or two before the round that first needs it; a round
    // re-evaluating proposals after a conflict reads them again instead of
    // drawing them again.  A round refills, while its term gathers are in
    // flight, when fewer than 128 proposals past its window start are drawn;
    // a round advances the window by at most 64, so the next window is
    // always drawn.  A block never overwrites a slot of the round's window
    // [s, s + 64): it fills [filled, filled + 128) with s + 64 <= filled <
    // s + 128, inside [s + 64, s + 256).
    int64_t filled = 0;
    auto refill = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t j = filled + 64 * h + lane;
            const uint32_t slot = (uint32_t)j & (kRing - 1);
            uint32_t a = 0, b = 0;
            float uj = 0.0f;
            if (j < max_iter) {
                a = (uint32_t)draw_range(job.seed, (uint64_t)j, 0, gw) |
                    ((uint32_t)draw_range(job.seed, (uint64_t)j, 1, gh) << 16);
                b = (uint32_t)draw_range(job.seed, (uint64_t)j, 2, gw) |
                    ((uint32_t)draw_range(job.seed, (uint64_t)j, 3, gh) << 16);
                if (!job.greedy) uj = draw_unit(job.seed, (uint64_t)j);
            }
            ring_xy1[slot] = a;
            ring_xy2[slot] = b;
            ring_u[slot] = uj;
        }
        filled += 128;
    };
    refill();
    __syncthreads();  // other lanes' draws are read below
    // The serial cost sum of a round's accepted deltas (packed in proposal
    // order into pack_s, see below) runs in the next round, after that
    // round's term gathers are requested, so it overlaps their latency; the
    // last round's runs after the loop.  pend: the count waiting (-1: none).
    int32_t pend = -1;
    auto serial_sum = [&](int32_t n_acc) {
        // Eight deltas at a time come back as two uniform 16-byte LDS reads,
        // so the chain adds VGPRs (no lane reads into SGPRs between the adds).
        const float4* pack4 = reinterpret_cast<const float4*>(pack_s);
        if (job.greedy) {  // optimize_greedy counts every accepted swap
            for (int32_t q = 0; q < n_acc; q += 8) {
                const float4 u0 = pack4[q >> 2], u1 = pack4[(q >> 2) + 1];
                const float v[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                for (int i = 0; i < 8; ++i) cost = cost + v[i];
            }
            improved += n_acc;
        } else {
            uint32_t better_n = 0;
            for (int32_t q = 0; q < n_acc; q += 8) {
                const float4 u0 = pack4[q >> 2], u1 = pack4[(q >> 2) + 1];
                const float v[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    cost = cost + v[i];
                    better_n += cost < best ? 1u : 0u;
                    best = __builtin_fminf(best, cost);
                }
            }
            improved += better_n;
        }
    };
    while (s < max_iter) {
        const int64_t it = s + lane;
        const bool active = it < max_iter;
        const uint32_t slot = (uint32_t)it & (kRing - 1);
        const uint32_t xy1 = ring_xy1[slot], xy2 = ring_xy2[slot];
        // Every ring slot up to `filled` holds in-grid coordinates (zeros past
        // max_iter), so all lanes read both cells and their neighbour slots
        // at once, before the skip test.
        const int32_t x1 = (int32_t)(xy1 & 0xffffu), y1 = (int32_t)(xy1 >> 16);
        const int32_t x2 = (int32_t)(xy2 & 0xffffu), y2 = (int32_t)(xy2 >> 16);
        const int32_t m1 = gget(grid_s, y1 * gw + x1), m2 = gget(grid_s, y2 * gw + x2);
        int32_t nb[8];
        swap_slots(grid_s, gw, gh, x1, y1, x2, y2, nb);
        const bool skip = !active || (x1 == x2 && y1 == y2) || m1 == kEmpty || m2 == kEmpty;
        const uint64_t nonskip = wave_ballot(!skip);
        // every lane requests its swap's terms (a skipped lane those of
        // material 0, which exists when any lane is not skipped), then the
        // temperature and the uniform draw; the draws of later proposals and
        // the previous round's serial sum run while they are in flight
        SwapTerms st;
        if (nonskip) {
            const int32_t g1 = skip ? 0 : m1, g2 = skip ? 0 : m2;
            if (job.term)
                swap_terms(gw, gh, TM, x1, y1, x2, y2, g1, g2, nb, st);
            else
                swap_terms(gw, gh, TL, x1, y1, x2, y2, g1, g2, nb, st);
        }
        float temp = 1.0f, u = 0.0f;
        if (!skip && !job.greedy) {
            temp = gload(job.temps, kbase + lanes_before(nonskip));
            u = ring_u[slot];
        }
        if (filled < s + 128) refill();  // read from the next round on, after its barriers
        if (pend >= 0) serial_sum(pend);
        float delta = 0.0f;
        bool acc = false;
        if (!skip) {
            delta = m1 == m2 ? 0.0f : swap_sum(st);  // swap_delta_core's value
            if (job.greedy) {
                acc = delta < 0.0f;
            } else if (delta < 0.0f) {
                acc = true;
            } else {
                // Metropolis test u < exp(x), x = -delta/T (src/optimizer.rs:77-84).
                // An estimate from the hardware reciprocal and exp2 is within
                // 2^-15 of exp(x), relative, down to 2^-126 (below, it may be 0,
                // and u is either 0 or at least 2^-24), so it decides alone
                // unless u lies within 2^-10 of it (or it is 0 or not finite);
                // those lanes compute x and exp exactly.
                const float e = __builtin_amdgcn_exp2f(-delta * __builtin_amdgcn_rcpf(temp) *
                                                       1.44269504f);
                if (__builtin_fabsf(u - e) > e * (1.0f / 1024.0f))
                    acc = u < e;
                else
                    acc = u < exp_f32(-delta / temp);
            }
        }
        const int32_t p1 = y1 * gw + x1, p2 = y2 * gw + x2;
        const uint32_t mine = (tag << 6) | (uint32_t)(63 - lane);
        if (acc) {
            atomicMax(&stamp[sidx(p1)], mine);
            atomicMax(&stamp[sidx(p2)], mine);
        }
        __syncthreads();
        bool conflict = false;
        if (!skip) {
            // read set: both cells and their in-grid neighbours, as ten fixed
            // slots (an out-of-grid neighbour re-reads its own cell, which is
            // checked anyway), so the reads are issued together
            const int32_t q[10] = {p1,
                                   p2,
                                   y1 > 0 ? p1 - gw : p1,
                                   y1 < gh - 1 ? p1 + gw : p1,
                                   x1 > 0 ? p1 - 1 : p1,
                                   x1 < gw - 1 ? p1 + 1 : p1,
                                   y2 > 0 ? p2 - gw : p2,
                                   y2 < gh - 1 ? p2 + gw : p2,
                                   x2 > 0 ? p2 - 1 : p2,
                                   x2 < gw - 1 ? p2 + 1 : p2};
#pragma unroll
            for (int i = 0; i < 10; ++i) {
                const uint32_t v = stamp[sidx(q[i])];
                conflict |= (v >> 6) == tag && (int)(63 - (v & 63u)) < lane;
            }
        }
        const uint64_t cmask = wave_ballot(conflict);
        int32_t first = cmask ? (__ffsll((unsigned long long)cmask) - 1) : 64;
        const int64_t remain = max_iter - s;
        if ((int64_t)first > remain) first = (int32_t)remain;
        const uint64_t commit = first >= 64 ? ~0ull : ((1ull << first) - 1ull);
        const uint64_t accmask = wave_ballot(acc) & commit;
        __syncthreads();
        if (lane < first) {
            if (acc) {
                // no lower lane of this round wrote p1 or p2 (that would be a
                // conflict), so the round-start reads m1, m2 are still current
                gput(grid_s, p1, m2);
                gput(grid_s, p2, m1);
            }
            if (job.trace_delta) gstore(job.trace_delta, it, skip ? __int_as_float(0x7fc00000) : delta);
            if (job.trace_accept) gstore(job.trace_accept, it, (uint8_t)(acc ? 1 : 0));
        }
        // an HBM grid's swaps are complete at agent scope before the next
        // round's reads (the barriers below order the lanes)
        if constexpr (kHbm) __threadfence();
        // Running cost in proposal order (f32, as the reference's
        // `current_cost += delta`).  The committed accepted deltas are packed
        // into words 0..n-1 of pack_s in proposal order (the other words hold
        // +0.0; pack_s is 16-byte aligned), so the serial sum reads them by a
        // counter instead of walking the mask.  Every lane runs the same
        // uniform loop, so cost / best / improved stay identical across the
        // wave.  Adding a +0.0 past the last delta changes nothing: the cost is
        // never -0.0 and never below best.  best = fminf(best, cost) equals
        // `cost < best ? cost : best` here (no NaN, no -0.0) and keeps the
        // compare off the dependent chain.
        pack_s[lane] = 0.0f;
        if (lane_in(accmask)) pack_s[lanes_before(accmask)] = delta;
        __syncthreads();  // other lanes' writes, not this lane's own stores, are read next round
        pend = __popcll(accmask);
        accepted += __popcll(accmask);
        const int64_t ev = __popcll(nonskip & commit);
        evaluated += ev;
        kbase += ev;
        s += first;
        ++rounds;
        ++tag;
        if (tag >= (1u << 26)) {
            __syncthreads();
            for (int32_t i = lane; i < stamp_cells; i += 64) stamp[i] = 0u;
            tag = 1;
        }
        __syncthreads();
    }

    if (pend >= 0) serial_sum(pend);

    if constexpr (!kHbm)
        for (int32_t i = lane; i < cells; i += 64) gstore(job.grid, i, (int32_t)grid_s[i]);
    if (lane == 0 && job.result) {
        AnnealOut r;
        r.initial_cost = initial;
        r.final_cost = cost;
        r.best_cost = job.greedy ? cost : best;
        r.improved = improved;
        r.accepted = accepted;
        r.iterations = max_iter;
        r.evaluated = evaluated;
        r.rounds = rounds;
        *job.result = r;
    }
}

int32_t anneal_kind(int32_t cells, int32_t n, int32_t gw, int32_t gh, size_t max_lds) {
    if (n < 32768 && anneal_lds_bytes(cells, kAnnealNarrow) <= max_lds) return kAnnealNarrow;
    if (anneal_lds_bytes(cells, kAnnealWide) <= max_lds) return kAnnealWide;
    if (gw < 65536 && gh < 65536 && anneal_lds_bytes(cells, kAnnealHbm) <= max_lds)
        return kAnnealHbm;
    return -1;
}

size_t anneal_lds_bytes(int32_t cells, int32_t kind) {
    if (kind == kAnnealHbm)
        return (size_t)(kHashStamps + kPackWords + kRingWords) * sizeof(uint32_t);
    return (size_t)(stamp_words(cells) + kPackWords + kRingWords) * sizeof(uint32_t) +
           (size_t)cells * (kind == kAnnealNarrow ? sizeof(int16_t) : sizeof(int32_t));
}

namespace {
template <typename... A>
void launch_anneal(int32_t kind, dim3 grid, size_t lds_bytes, hipStream_t stream, A... args) {
    if (kind == kAnnealNarrow)
        hipLaunchKernelGGL((anneal_kernel<int16_t, false>), grid, dim3(64), lds_bytes, stream,
                           args...);
    else if (kind == kAnnealWide)
        hipLaunchKernelGGL((anneal_kernel<int32_t, false>), grid, dim3(64), lds_bytes, stream,
                           args...);
    else
        hipLaunchKernelGGL((anneal_kernel<int32_t, true>), grid, dim3(64), lds_bytes, stream,
                           args...);
}
}  // namespace

void launch_anneal_single(const AnnealJob& job, int32_t kind, size_t lds_bytes, hipStream_t stream) {
    launch_anneal(kind, dim3(1), lds_bytes, stream, job, (const AnnealJob*)nullptr);
}

void launch_anneal_batch(const AnnealJob* d_jobs, int32_t n_jobs, int32_t kind, size_t lds_bytes,
                         hipStream_t stream) {
    if (n_jobs <= 0) return;
    AnnealJob unused{};
    launch_anneal(kind, dim3(n_jobs), lds_bytes, stream, unused, d_jobs);
}

hipError_t anneal_configure(int max_lds) {
    const void* kernels[] = {
        reinterpret_cast<const void*>(&anneal_kernel<int16_t, false>),
        reinterpret_cast<const void*>(&anneal_kernel<int32_t, false>),
        reinterpret_cast<const void*>(&anneal_kernel<int32_t, true>),
    };
    for (const void* k : kernels) {
        const hipError_t e = raise_dynamic_lds(k, max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ------------------------------------------------ cost / delta / penalty
__global__ __launch_bounds__(64) void total_cost_kernel(const int32_t* __restrict__ grid,
                                                        int32_t gw, int32_t gh,
                                                        const float4* __restrict__ lab4,
                                                        const float* __restrict__ term, int32_t n,
                                                        float* __restrict__ scratch, int32_t chunk,
                                                        float* __restrict__ out) {
    const float t = term ? total_cost_wave(grid, gw, gh, TermMatrix{term, n}, scratch, chunk)
                         : total_cost_wave(grid, gw, gh, TermLab{lab4}, scratch, chunk);
    if (threadIdx.x == 0) *out = t;
}

void launch_total_cost(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4,
                       const float* term, int32_t n, float* scratch, int32_t chunk, float* out,
                       hipStream_t stream) {
    hipLaunchKernelGGL(total_cost_kernel, dim3(1), dim3(64), 0, stream, grid, gw, gh, lab4, term, n,
                       scratch, chunk, out);
}

__global__ void swap_delta_kernel(const int32_t* __restrict__ grid, int32_t gw, int32_t gh,
                                  const float4* __restrict__ lab4, const float* __restrict__ term,
                                  int32_t n, const int32_t* __restrict__ pairs, int32_t p,
                                  float* __restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p) return;
    const int32_t* q = pairs + 4 * (int64_t)i;
    out[i] = term ? swap_delta_dev(grid, gw, gh, TermMatrix{term, n}, q[0], q[1], q[2], q[3])
                  : swap_delta_dev(grid, gw, gh, TermLab{lab4}, q[0], q[1], q[2], q[3]);
}

void launch_swap_deltas(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4,
                        const float* term, int32_t n, const int32_t* pairs, int32_t p, float* out,
                        hipStream_t stream) {
    if (p <= 0) return;
    hipLaunchKernelGGL(swap_delta_kernel, dim3((p + 255) / 256), dim3(256), 0, stream, grid, gw, gh,
                       lab4, term, n, pairs, p, out);
}

// calculate_penalty (src/adjacency.rs:63-88): Up, Down, Left, Right.
__global__ void penalty_kernel(const int32_t* __restrict__ grid, int32_t gw, int32_t gh,
                               const float4* __restrict__ lab4, float weight,
                               const int32_t* __restrict__ queries, int32_t nq,
                               float* __restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int32_t m = queries[3 * (int64_t)i], x = queries[3 * (int64_t)i + 1],
                  y = queries[3 * (int64_t)i + 2];
    const float4 lm = lab4[m];
    float p = 0.0f;
    int32_t nb[4], cnt = 0;
    if (y > 0) nb[cnt++] = (y - 1) * gw + x;
    if (y < gh - 1) nb[cnt++] = (y + 1) * gw + x;
    if (x > 0) nb[cnt++] = y * gw + x - 1;
    if (x < gw - 1) nb[cnt++] = y * gw + x + 1;
    for (int32_t j = 0; j < cnt; ++j) {
        const int32_t o = grid[nb[j]];
        if (o != kEmpty) p += adj_term(lm, lab4[o]);
    }
    out[i] = p * weight;
}

void launch_penalties(const int32_t* grid, int32_t gw, int32_t gh, const float4* lab4, float weight,
                      const int32_t* queries, int32_t nq, float* out, hipStream_t stream) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(penalty_kernel, dim3((nq + 255) / 256), dim3(256), 0, stream, grid, gw, gh,
                       lab4, weight, queries, nq, out);
}

}  // namespace mh
