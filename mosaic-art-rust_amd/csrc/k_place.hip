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
// penalty (similarities recomputed from the Lab table -- bit-identical to the
// stored matrix), score = sqdist + penalty, and the reference's first
// strict-< winner becomes a min over (score_bits << 32 | rank).  Usage counts
// and the grid live in LDS.  Latency-bound: one cell is a few hundred cycles.
//
// Many jobs run as many independent waves (one block per job).
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

__global__ __launch_bounds__(64) void place_chain_kernel(PlaceJob single,
                                                         const PlaceJob* __restrict__ jobs) {
    extern __shared__ int32_t sm[];
    const PlaceJob job = jobs ? jobs[blockIdx.x] : single;
    const int lane = threadIdx.x;
    const int32_t gw = job.grid_w, gh = job.grid_h, n = job.n;
    const int32_t cells = gw * gh;
    int32_t* grid_s = sm;
    uint32_t* usage_s = reinterpret_cast<uint32_t*>(sm + cells);

    for (int32_t i = lane; i < cells; i += 64) grid_s[i] = job.grid[i];
    for (int32_t i = lane; i < n; i += 64) usage_s[i] = job.usage[i];
    __syncthreads();

    const int32_t K = n < kMaxCand ? n : kMaxCand;
    const uint32_t max_usage = (uint32_t)job.max_usage;
    const float weight = job.weight;
    const bool use_pen = weight > 0.0f;
    int32_t st_main = 0, st_ban = 0, st_near = 0, st_resets = 0, st_empty = 0;
    const int32_t steps = job.order ? job.n_order : cells;

    for (int32_t c = 0; c < steps; ++c) {
        const int32_t cell = job.order ? job.order[c] : c;
        if (n == 0) {
            ++st_empty;
            continue;
        }
        const int32_t x = cell % gw, y = cell / gw;
        // neighbours in get_adjacent_positions order: Up, Down, Left, Right
        const int32_t nb_up = y > 0 ? grid_s[cell - gw] : kEmpty;
        const int32_t nb_down = y < gh - 1 ? grid_s[cell + gw] : kEmpty;
        const int32_t nb_left = x > 0 ? grid_s[cell - 1] : kEmpty;
        const int32_t nb_right = x < gw - 1 ? grid_s[cell + 1] : kEmpty;
        float4 lu = make_float4(0, 0, 0, 0), ld = lu, ll = lu, lr = lu;
        if (use_pen) {
            if (nb_up != kEmpty) lu = job.lab4[nb_up];
            if (nb_down != kEmpty) ld = job.lab4[nb_down];
            if (nb_left != kEmpty) ll = job.lab4[nb_left];
            if (nb_right != kEmpty) lr = job.lab4[nb_right];
        }
        const int32_t* cidx = job.cand_idx + (int64_t)cell * kCandStride;
        const float* cd = job.cand_d + (int64_t)cell * kCandStride;

        uint64_t best = ~0ull;
        int32_t m_slot[2];
        bool placeable[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int32_t r = lane + 64 * s;
            m_slot[s] = kEmpty;
            placeable[s] = false;
            if (r < K) {
                const int32_t m = cidx[r];
                const float d = cd[r];
                m_slot[s] = m;
                const bool pl = m != nb_up && m != nb_down && m != nb_left && m != nb_right;
                placeable[s] = pl;
                if (pl && usage_s[m] < max_usage) {
                    float pen = 0.0f;
                    if (use_pen) {
                        const float4 lm = job.lab4[m];
                        float p = 0.0f;
                        if (nb_up != kEmpty) p += adj_term(lm, lu);
                        if (nb_down != kEmpty) p += adj_term(lm, ld);
                        if (nb_left != kEmpty) p += adj_term(lm, ll);
                        if (nb_right != kEmpty) p += adj_term(lm, lr);
                        pen = p * weight;
                    }
                    const float score = d + pen;
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
            for (int32_t i = lane; i < n; i += 64) usage_s[i] = 0;
            ++st_resets;
            const uint64_t b0 = __ballot(placeable[0]);
            const uint64_t b1 = __ballot(placeable[1]);
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
        const int32_t w = cidx[rank];
        __syncthreads();
        if (lane == 0) {
            usage_s[w] = usage_s[w] + 1u;
            grid_s[cell] = w;
            if (job.stage) job.stage[cell] = (uint8_t)stage;
            if (job.dist) job.dist[cell] = cd[rank];
            if (job.score) {
                float sc = cd[rank];
                if (stage == 0) sc = __uint_as_float((uint32_t)(best >> 32));
                job.score[cell] = sc;
            }
        }
        __syncthreads();
    }

    for (int32_t i = lane; i < cells; i += 64) job.grid[i] = grid_s[i];
    for (int32_t i = lane; i < n; i += 64) job.usage[i] = usage_s[i];
    if (lane == 0 && job.stats) {
        job.stats[0] = st_main + st_ban + st_near;
        job.stats[1] = st_main;
        job.stats[2] = st_ban;
        job.stats[3] = st_near;
        job.stats[4] = st_resets;
        job.stats[5] = st_empty;
    }
}

size_t place_lds_bytes(int32_t cells, int32_t n) {
    return sizeof(int32_t) * ((size_t)cells + (size_t)n);
}

void launch_place_single(const PlaceJob& job, size_t lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL(place_chain_kernel, dim3(1), dim3(64), lds_bytes, stream, job,
                       (const PlaceJob*)nullptr);
}

void launch_place_batch(const PlaceJob* d_jobs, int32_t n_jobs, size_t lds_bytes,
                        hipStream_t stream) {
    if (n_jobs <= 0) return;
    PlaceJob unused{};
    hipLaunchKernelGGL(place_chain_kernel, dim3(n_jobs), dim3(64), lds_bytes, stream, unused, d_jobs);
}

hipError_t place_configure(int max_lds) {
    return raise_dynamic_lds(reinterpret_cast<const void*>(&place_chain_kernel), max_lds);
}

}  // namespace mh
