// k_topk.hip -- K nearest materials per cell (the kiddo k-d tree query
// nearest_n::<SquaredEuclidean>(lab, min(N,100)) at src/main.rs:348-352).
//
// Exact: each material gets the 64-bit key (float_bits(sqdist) << 32 |
// index).  Squared distances are >= +0, so the bit pattern orders like the
// value and keys are unique; the K smallest keys are the K nearest materials
// ascending with ties broken by index.
//
// The library is laid out in blocks of 64 materials (MatBlocks: median
// splits on the widest Lab axis, so each block is a compact box).  A block
// whose box lies farther than the threshold tau holds no candidate: its
// bound (box_bound) is a lower bound on every member's computed distance,
// so skipping it changes nothing.  This is the k-d tree's pruning, flat:
// one wave tests 64 block boxes at once and scans only the blocks in reach.
//
// One wave (a 64-thread workgroup) answers a run of consecutive query cells.
// Per cell it keeps a candidate buffer in LDS and a threshold tau such that
// at least K keys are <= tau:
//   * tau starts as the largest of this cell's distances to the previous
//     cell's K answers (K materials lie within tau, so the K nearest do too);
//     neighbouring cells have similar Lab means, so tau is tight;
//   * one pass over the blocks within tau appends every member within tau
//     (ballot + popcount compaction, one block of 64 per step);
//   * when the buffer would overflow it is sorted and cut to its K smallest,
//     whose largest becomes the new (tighter) tau;
//   * at the end the buffer is sorted and the first K keys are the answer.
//     Sorts are bitonic networks over 2 or 4 keys per lane held in
//     registers; partners are exchanged by DPP / swizzle / permute lane
//     moves, so no sort step touches LDS.
// The first cell of a run (no previous answers) takes the blocks nearest by
// bound until each lane's two smallest member keys number at least K; the
// K-th smallest of those real distances, rounded up within 2^12 ulps, bounds
// the K-th smallest overall.
// The threshold is compared on the distance bits alone (one compare per
// material): keys at exactly the threshold distance may enter the buffer
// beyond the K-th, and the sorts cut them.
// The batched kernel serves many (library, query set) segments in one
// launch: block b handles run b - starts[j] of the segment j holding b.
#include <type_traits>

#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

constexpr int kBuf = 256;      // candidate buffer (keys) per wave
constexpr int kMaxK = 128;     // largest K served
constexpr int kMaxRun = 8;     // query cells per wave

__device__ __forceinline__ uint64_t make_key(uint32_t dbits, int32_t i) {
    return ((uint64_t)dbits << 32) | (uint32_t)i;
}

// max = ~min(~v), with wave_min_u32's DPP steps; all 64 lanes active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return ~wave_min_u32(~v); }

// The value of lane ^ m (m a power of two below 64) for every lane: DPP
// lane swizzles within rows of 16 for m = 1, 2, 8, the swizzle unit for
// m = 4, 16 (within halves of 32), a lane permute for m = 32.  m is a
// compile-time constant after unrolling, so only one case remains.
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int m) {
    switch (m) {
        case 1: return __builtin_amdgcn_mov_dpp(v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
        case 2: return __builtin_amdgcn_mov_dpp(v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
        case 8: return __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, false);  // row_ror:8
        case 4: return __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1f);      // xor 4
        case 16: return __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1f);    // xor 16
        default: return __shfl_xor(v, 32, 64);
    }
}

__device__ __forceinline__ uint64_t lane_xor64(uint64_t v, int m) {
    return ((uint64_t)lane_xor((uint32_t)(v >> 32), m) << 32) | lane_xor((uint32_t)v, m);
}

// Lanes whose index has bit b set (b a power of two below 64).
__device__ __forceinline__ constexpr uint64_t lane_bit_mask(int b) {
    return b == 1 ? 0xAAAAAAAAAAAAAAAAull
         : b == 2 ? 0xCCCCCCCCCCCCCCCCull
         : b == 4 ? 0xF0F0F0F0F0F0F0F0ull
         : b == 8 ? 0xFF00FF00FF00FF00ull
         : b == 16 ? 0xFFFF0000FFFF0000ull
                   : 0xFFFFFFFF00000000ull;
}

// Ascending bitonic sort of the S*64 keys held in registers, element
// s*64 + lane in key[s].  Pairs 64 or more apart are in one lane
// (compare-exchange in registers); closer pairs are in partner lanes, which
// exchange through the lane swizzles above -- no LDS round trip per stage.
// Whether a lane keeps the smaller key of its pair is a compile-time lane
// mask for each stage.
template <int S>
__device__ __forceinline__ void reg_sort(uint64_t (&key)[S]) {
    constexpr int P = S * 64;
#pragma unroll
    for (int size = 2; size <= P; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 64) {
                const int sd = stride / 64;
#pragma unroll
                for (int t = 0; t < S; ++t) {
                    if ((t & sd) != 0) continue;
                    const bool asc = ((t * 64) & size) == 0;
                    const uint64_t a = key[t], b = key[t | sd];
                    const bool swap = (a > b) == asc;
                    key[t] = swap ? b : a;
                    key[t | sd] = swap ? a : b;
                }
            } else {
#pragma unroll
                for (int t = 0; t < S; ++t) {
                    // this lane is the pair's lower element where its stride
                    // bit is clear; the pair ascends where bit `size` of the
                    // element index is clear; the lower element of an
                    // ascending pair (and the upper of a descending one) keeps
                    // the minimum
                    const uint64_t upper = lane_bit_mask(stride);
                    const uint64_t desc = size < 64 ? lane_bit_mask(size)
                                        : (((t * 64) & size) != 0 ? ~0ull : 0ull);
                    const uint64_t keep_min = ~(upper ^ desc);
                    const uint64_t mine = key[t];
                    const uint64_t other = lane_xor64(mine, stride);
                    const uint64_t other_less = __builtin_amdgcn_uicmpl(other, mine, 36);
                    // take the partner's key where it is the one to keep (keys
                    // are unique but for the padding, where either is the same)
                    const uint64_t take = ~(other_less ^ keep_min);
                    key[t] = __builtin_amdgcn_inverse_ballot_w64(take) ? other : mine;
                }
            }
        }
    }
}

// Sorts buf[0..count) ascending in place (count <= 256); the sorted keys
// are also left in key[] (slot-major, padded with the largest key).
__device__ __forceinline__ void sort_buffer(uint64_t* buf, int32_t count) {
    const int lane = threadIdx.x;
    if (count <= 128) {
        uint64_t key[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) key[t] = t * 64 + lane < count ? buf[t * 64 + lane] : ~0ull;
        reg_sort<2>(key);
#pragma unroll
        for (int t = 0; t < 2; ++t) buf[t * 64 + lane] = key[t];
    } else {
        uint64_t key[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) key[t] = t * 64 + lane < count ? buf[t * 64 + lane] : ~0ull;
        reg_sort<4>(key);
#pragma unroll
        for (int t = 0; t < 4; ++t) buf[t * 64 + lane] = key[t];
    }
    __builtin_amdgcn_wave_barrier();
}

// Keeps the buffer's keys whose distance bits are <= some t with
// k <= #(keys <= t) <= 128 (bisection over t), compacted in key order of
// arrival, and returns their count; returns count unchanged when ties make
// no such t exist.  The k smallest keys are among those kept, and 128 keys
// take the 2-per-lane sort instead of the 4-per-lane one.
__device__ __forceinline__ int32_t cut_buffer(uint64_t* buf, int32_t count, int32_t k) {
    const int lane = threadIdx.x;
    uint64_t key[4];
    uint32_t d[4];
    uint32_t dmax = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bool in = t * 64 + lane < count;
        key[t] = in ? buf[t * 64 + lane] : ~0ull;
        d[t] = (uint32_t)(key[t] >> 32);
        if (in) dmax = d[t] > dmax ? d[t] : dmax;
    }
    dmax = wave_max_u32(dmax);
    auto below = [&](uint32_t x) {
        int32_t c = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) c += __popcll(lanes_ule(d[t], x) & lanes_slt(t * 64 + lane, count));
        return c;
    };
    uint32_t lo = 0, hi = dmax, sel = 0;
    bool found = false;
    while (lo <= hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        const int32_t c = below(mid);
        if (c < k) {
            lo = mid + 1;
        } else if (c > 128) {
            if (mid == 0) break;
            hi = mid - 1;
        } else {
            sel = mid;
            found = true;
            break;
        }
    }
    if (!found) return count;
    int32_t base = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint64_t keep = lanes_ule(d[t], sel) & lanes_slt(t * 64 + lane, count);
        if (lane_in(keep)) buf[base + lanes_before(keep)] = key[t];
        base += __popcll(keep);
    }
    __builtin_amdgcn_wave_barrier();
    return base;
}

// Lower bound, as float bits, on the squared distance sq_distance computes
// from the query to any material in a block's box.  Per axis the smallest
// |fl(q - x)| over x in [lo, hi] is at a face (fl(q - x) is monotone in x),
// or 0 inside: max(fl(lo - q), fl(q - hi), 0), since fl(lo - q) = -fl(q - lo)
// (one is positive only outside the box, on its side).  The squares and the
// sum, in sq_distance's order, are monotone in their operands, so no member
// computes a smaller distance.
__device__ __forceinline__ float face_gap(float q, float lo, float hi) {
    return __builtin_fmaxf(__builtin_fmaxf(lo - q, q - hi), 0.0f);
}
__device__ __forceinline__ uint32_t box_bound(float ql, float qa, float qb, float4 lo, float4 hi) {
    const float tl = face_gap(ql, lo.x, hi.x);
    const float ta = face_gap(qa, lo.y, hi.y);
    const float tb = face_gap(qb, lo.z, hi.z);
    return __float_as_uint(tl * tl + ta * ta + tb * tb);
}

// Answers query cells [q0, q0 + run) of segment s.
__device__ __forceinline__ void topk_run(const TopkSeg& s, int32_t q0, int32_t run,
                                         uint64_t* buf, int32_t* prev) {
    const int lane = threadIdx.x;
    const float* __restrict__ mL = s.mL;
    const float* __restrict__ mA = s.mA;
    const float* __restrict__ mB = s.mB;
    const MatBlocks& B = s.blk;
    const int32_t n = s.n;
    const int32_t nb = B.nb;
    const int32_t k = s.k < n ? s.k : n;
    const int32_t q_end = q0 + run < s.nq ? q0 + run : s.nq;
    constexpr int kSteps = 4;  // blocks per load phase

    // The boxes of the first two chunks of blocks (blocks lane and 64 + lane)
    // are the same for every cell of the run: loaded once, so a cell's first
    // bounds wait for no memory (libraries of up to 8,192 materials need no
    // other box).
    float4 box_c[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int32_t b = c * 64 + lane;
        const int32_t bc = b < nb ? b : nb - 1;
        box_c[c][0] = gload4(B.box, 2 * bc);
        box_c[c][1] = gload4(B.box, 2 * bc + 1);
    }

    for (int32_t q = q0; q < q_end; ++q) {
        const float ql = gload(s.query, 3 * (int64_t)q), qa = gload(s.query, 3 * (int64_t)q + 1),
                    qb = gload(s.query, 3 * (int64_t)q + 2);
        // this lane's bound of block b = c0 + lane (blocks past the last bound
        // to ~0)
        auto bound_of = [&](int32_t b) -> uint32_t {
            const int32_t c0 = b - lane;
            const int32_t bc = b < nb ? b : nb - 1;
            uint32_t lb;
            if (c0 == 0)
                lb = box_bound(ql, qa, qb, box_c[0][0], box_c[0][1]);
            else if (c0 == 64)
                lb = box_bound(ql, qa, qb, box_c[1][0], box_c[1][1]);
            else
                lb = box_bound(ql, qa, qb, gload4(B.box, 2 * bc), gload4(B.box, 2 * bc + 1));
            return b < nb ? lb : 0xffffffffu;
        };
        // member `lane` of block b: distance bits and material (-1: padding)
        auto member = [&](int32_t b, uint32_t& d, int32_t& m) {
            const uint32_t off = 4u * (uint32_t)(b * 64 + lane);  // slots < 2^30
            m = gload_at(B.pidx, off);
            d = __float_as_uint(sq_distance(ql, qa, qb, gload_at(B.pL, off), gload_at(B.pA, off),
                                            gload_at(B.pB, off)));
        };
        // Threshold on the distance bits: at least k materials lie at or
        // below th, so the k nearest all do.  Keys at exactly th beyond the
        // k-th may enter the buffer too; the final sort cuts them.
        uint32_t th = 0xffffffffu;
        if (q > q0) {
            // the previous cell's k answers
            uint32_t m = 0;
            for (int32_t r = lane; r < k; r += 64) {
                const int32_t i = prev[r];
                const uint32_t d = __float_as_uint(
                    sq_distance(ql, qa, qb, gload(mL, i), gload(mA, i), gload(mB, i)));
                m = d > m ? d : m;
            }
            th = wave_max_u32(m);
        } else if (n > kBuf) {
            // The first cell of a run: the blocks nearest by bound, one at a
            // time, until every lane's two smallest distances among their
            // members cover k materials; the k-th smallest of those keys
            // bounds the k-th nearest overall.  Two full blocks give 128
            // keys and only the last block is partial, so three suffice.
            // Each lane keeps its three smallest (bound, block) keys over its
            // blocks lane, lane + 64, ..., so every bound is computed once;
            // the wave minimum of the lanes' smallest is the next nearest
            // block, and the lane that held it moves on to its next.
            uint64_t lk0 = ~0ull, lk1 = ~0ull, lk2 = ~0ull;
            for (int32_t c0 = 0; c0 < nb; c0 += 64) {
                const int32_t b = c0 + lane;
                const uint64_t key = ((uint64_t)bound_of(b) << 32) | (uint32_t)b;
                const bool l0 = key < lk0, l1 = key < lk1, l2 = key < lk2;
                lk2 = l1 ? lk1 : (l2 ? key : lk2);
                lk1 = l0 ? lk0 : (l1 ? key : lk1);
                lk0 = l0 ? key : lk0;
            }
            uint64_t two0 = ~0ull, two1 = ~0ull;
#pragma unroll
            for (int it = 0; it < 3; ++it) {
                const uint64_t best = wave_min_u64(lk0);
                const int32_t blk =
                    __builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(best & 0xffffffffu));
                const bool own = lk0 == best;  // keys are unique (the block in the low half)
                lk0 = own ? lk1 : lk0;
                lk1 = own ? lk2 : lk1;
                lk2 = own ? ~0ull : lk2;
                uint32_t d;
                int32_t m;
                member(blk, d, m);
                const uint64_t key = m >= 0 ? make_key(d, m) : ~0ull;
                const bool lt0 = key < two0, lt1 = key < two1;
                two1 = lt0 ? two0 : (lt1 ? key : two1);
                two0 = lt0 ? key : two0;
                const int32_t real = __popcll(wave_ballot(two0 != ~0ull)) +
                                     __popcll(wave_ballot(two1 != ~0ull));
                if (real >= k) break;
            }
            // th: the k-th smallest of those distances, rounded up to the end
            // of its 2^12-ulp step -- the largest t of the form j * 2^12 with
            // fewer than k keys below it, built bit by bit from the top; then
            // at least k keys lie at or below t + 2^12 - 1.  (Padding keys
            // have all bits set and are never counted.)
            const uint32_t e0 = (uint32_t)(two0 >> 32), e1 = (uint32_t)(two1 >> 32);
            uint32_t t = 0;
#pragma unroll
            for (int bit = 31; bit >= 12; --bit) {
                const uint32_t cand = t | (1u << bit);
                const int32_t below = __popcll(lanes_ult(e0, cand)) + __popcll(lanes_ult(e1, cand));
                if (below < k) t = cand;
            }
            th = t | 0xfffu;
        }
        int32_t count = 0;
        // appends the passing lanes of one block to the buffer, first
        // cutting the buffer to its k smallest when it would overflow
        auto take = [&](uint32_t d, int32_t m, uint64_t mask) {
            if (count + __popcll(mask) > kBuf) {
                sort_buffer(buf, count);
                count = k;
                th = (uint32_t)(buf[k - 1] >> 32);
                mask &= lanes_ule(d, th);
            }
            if (lane_in(mask)) buf[count + lanes_before(mask)] = make_key(d, m);
            count += __popcll(mask);
        };
        // every block whose bound is within th, up to kSteps per load phase
        // (their loads all issued before the first distance is needed); th
        // only tightens, so a chunk's remaining blocks are re-tested after
        // each phase
        for (int32_t c0 = 0; c0 < nb; c0 += 64) {
            const uint32_t lb = bound_of(c0 + lane);
            uint64_t todo = lanes_ule(lb, th) & lanes_slt(c0 + lane, nb);
            while (todo) {
                int32_t bs[kSteps];
                int32_t cnt = 0;
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    bs[u] = c0 + __ffsll((unsigned long long)todo) - 1;  // a repeat once empty
                    if (todo) {
                        ++cnt;
                        todo &= todo - 1;
                    } else {
                        bs[u] = bs[0];
                    }
                }
                uint32_t d[kSteps];
                int32_t m[kSteps];
#pragma unroll
                for (int u = 0; u < kSteps; ++u) member(bs[u], d[u], m[u]);
#pragma unroll
                for (int u = 0; u < kSteps; ++u) {
                    if (u < cnt) {
                        const uint64_t mask = lanes_ule(d[u], th) & lanes_nonneg(m[u]);
                        if (mask) take(d[u], m[u], mask);
                    }
                }
                todo &= lanes_ule(lb, th);
            }
        }
        if (count > 128) count = cut_buffer(buf, count, k);
        sort_buffer(buf, count);
        // answer: the first k keys, ascending
        const int32_t out_stride = s.out_stride;
        for (int32_t r = lane; r < out_stride; r += 64) {
            const int64_t o = (int64_t)q * out_stride + r;
            if (r < k) {
                const uint64_t key = buf[r];
                const int32_t m = (int32_t)(uint32_t)(key & 0xffffffffu);
                prev[r] = m;
                if (s.out_idx) gstore(s.out_idx, o, m);
                if (s.out_d) gstore(s.out_d, o, __uint_as_float((uint32_t)(key >> 32)));
                // the key is {m, d} as one little-endian word: m in the low
                // half (CandRec::m, offset 0), the distance bits in the high
                if (s.out_cand) gstore((uint64_t*)s.out_cand, o, key);
            } else if (s.out_cand) {
                // padded (chain) rows: ranks beyond k, -1 with an infinite distance
                gstore((uint64_t*)s.out_cand, o, (uint64_t)0x7f800000u << 32 | 0xffffffffu);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void topk_kernel(TopkSeg s, int32_t run) {
    __shared__ uint64_t buf[kBuf];
    __shared__ int32_t prev[kMaxK];
    topk_run(s, blockIdx.x * run, run, buf, prev);
}

// Eight waves per SIMD (the hardware's most): the compiler otherwise takes
// 106 SGPRs, which leaves room for seven; at eight it keeps 26 of them in
// VGPR lanes.  Served stream 58.6-58.7k -> 59.0k jobs/s.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void topk_batch_kernel(
    const TopkSeg* __restrict__ segs, const int32_t* __restrict__ starts, int32_t n_segs,
    int32_t run) {
    __shared__ uint64_t buf[kBuf];
    __shared__ int32_t prev[kMaxK];
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    topk_run(segs[j], (blockIdx.x - starts[j]) * run, run, buf, prev);
}

// Query cells per wave: runs reuse the previous answer as a bound, but a
// launch wants enough waves to hide the scan's memory latency (~32 per CU):
// one cell per wave up to 8,192 cells (4,800 cells: 26 us against 30 us with
// runs of 2), then longer runs.
int32_t topk_run_length(int64_t total_queries) {
    int64_t r = total_queries / (256 * 32);
    return (int32_t)(r < 1 ? 1 : (r > kMaxRun ? kMaxRun : r));
}

int32_t topk_blocks(int32_t nq, int32_t run) { return nq > 0 ? (nq + run - 1) / run : 0; }

void launch_topk(const float* mL, const float* mA, const float* mB, int32_t n, const MatBlocks& blk,
                 const float* query, int32_t nq, int32_t k, int32_t* out_idx, float* out_d,
                 CandRec* out_cand, int32_t out_stride, hipStream_t stream) {
    if (k > n) k = n;
    if (nq <= 0 || n <= 0 || k <= 0) return;
    const TopkSeg s{mL, mA, mB, query, out_idx, out_d, out_cand, n, nq, k, out_stride, blk};
    const int32_t run = topk_run_length(nq);
    hipLaunchKernelGGL(topk_kernel, dim3(topk_blocks(nq, run)), dim3(64), 0, stream, s, run);
}

void launch_topk_batch(const TopkSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                       int32_t total_blocks, int32_t run, hipStream_t stream) {
    if (n_segs <= 0 || total_blocks <= 0) return;
    hipLaunchKernelGGL(topk_batch_kernel, dim3(total_blocks), dim3(64), 0, stream, d_segs, d_starts,
                       n_segs, run);
}

hipError_t raise_dynamic_lds(const void* kernel, int max_lds) {
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, kernel);
    if (e != hipSuccess) return e;
    const int dyn = max_lds - (int)attr.sharedSizeBytes;
    if (dyn <= 0) return hipErrorInvalidValue;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
}

hipError_t configure_kernel_limits(int max_lds) {
    hipError_t e = place_configure(max_lds);
    if (e == hipSuccess) e = anneal_configure(max_lds);
    if (e == hipSuccess) e = resize_configure(max_lds);
    return e;
}

}  // namespace mh
