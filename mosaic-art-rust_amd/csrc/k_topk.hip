// k_topk.hip -- K nearest materials per cell (the kiddo k-d tree query
// nearest_n::<SquaredEuclidean>(lab, min(N,100)) at src/main.rs:348-352).
//
// Brute force over all N materials, exact: each material gets the 64-bit
// key (float_bits(sqdist) << 32 | index).  Squared distances are >= +0, so
// the bit pattern orders like the value and keys are unique; the K smallest
// keys are the K nearest materials ascending with ties broken by index.
//
// One wave (a 64-thread workgroup) answers a run of consecutive query cells.
// Per cell it keeps a candidate buffer in LDS and a threshold tau such that
// at least K keys are <= tau:
//   * tau starts as the largest of this cell's distances to the previous
//     cell's K answers (K materials lie within tau, so the K nearest do too);
//     neighbouring cells have similar Lab means, so tau is tight;
//   * one pass over the N materials appends every material within tau
//     (ballot + popcount compaction, 64 materials per step);
//   * when the buffer would overflow it is sorted and cut to its K smallest,
//     whose largest becomes the new (tighter) tau;
//   * at the end the buffer is sorted and the first K keys are the answer.
//     Sorts are bitonic networks over 2 or 4 keys per lane held in
//     registers; partners are exchanged by DPP / swizzle / permute lane
//     moves, so no sort step touches LDS.
// The first cell of a run (no previous answers) gets its tau from one extra
// pass that keeps each lane's two smallest distances: the K-th smallest of
// those 128 real distances bounds the K-th smallest overall.  The threshold
// is compared on the distance bits alone (one compare per material): keys at
// exactly the threshold distance may enter the buffer beyond the K-th, and
// the sorts cut them.  The cost is one distance evaluation and one compare
// per material plus a few register sorts, with no histograms or atomics.
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

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
        v = o > v ? o : v;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

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

// Answers query cells [q0, q0 + run) of segment s.
__device__ __forceinline__ void topk_run(const TopkSeg& s, int32_t q0, int32_t run,
                                         uint64_t* buf, int32_t* prev) {
    const int lane = threadIdx.x;
    const float* __restrict__ mL = s.mL;
    const float* __restrict__ mA = s.mA;
    const float* __restrict__ mB = s.mB;
    const int32_t n = s.n;
    const int32_t k = s.k < n ? s.k : n;
    const int32_t q_end = q0 + run < s.nq ? q0 + run : s.nq;
    const uint64_t lanes_below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // materials in whole blocks of kSteps x 64, then a clamped tail block
    constexpr int kSteps = 4;
    const int32_t n_full = n - n % (64 * kSteps);

    for (int32_t q = q0; q < q_end; ++q) {
        const float ql = gload(s.query, 3 * (int64_t)q), qa = gload(s.query, 3 * (int64_t)q + 1),
                    qb = gload(s.query, 3 * (int64_t)q + 2);
        // Four 64-material steps per load phase: all twelve loads are issued
        // before the first distance is needed (one memory latency per four
        // steps).  Whole blocks load from a uniform base at fixed lane
        // offsets (no per-step address or bounds arithmetic); the tail block
        // clamps its loads in bounds and marks lanes past n.  f(d, i, valid)
        // sees each step's distance bits, the material index as a function
        // (evaluated only where used) and the in-range lane mask.
        auto block = [&](int32_t base, auto tail_c, auto&& f) {
            constexpr bool tail = decltype(tail_c)::value;
            float l[kSteps], a[kSteps], b[kSteps];
#pragma unroll
            for (int u = 0; u < kSteps; ++u) {
                int32_t i = base + 64 * u + lane;
                if (tail) i = i < n ? i : n - 1;
                const uint32_t off = 4u * (uint32_t)i;  // n < 2^30
                l[u] = gload_at(mL, off);
                a[u] = gload_at(mA, off);
                b[u] = gload_at(mB, off);
            }
            uint32_t d[kSteps];
#pragma unroll
            for (int u = 0; u < kSteps; ++u)
                d[u] = __float_as_uint(sq_distance(ql, qa, qb, l[u], a[u], b[u]));
#pragma unroll
            for (int u = 0; u < kSteps; ++u) {
                const int32_t first = base + 64 * u;
                const uint64_t valid = tail ? lanes_slt(first + lane, n) : ~0ull;
                f(d[u], first, valid);
            }
        };
        auto scan = [&](auto&& f) {
            for (int32_t base = 0; base < n_full; base += 64 * kSteps)
                block(base, std::false_type{}, f);
            if (n_full < n) block(n_full, std::true_type{}, f);
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
            // The first cell of a run: one pass keeps every lane's two
            // smallest distances (of distinct materials); the k-th smallest
            // of those 128 bounds the k-th smallest overall.
            uint32_t two0 = 0xffffffffu, two1 = 0xffffffffu;
            scan([&](uint32_t d, int32_t, uint64_t valid) {
                const uint32_t v = ((valid >> lane) & 1ull) ? d : 0xffffffffu;
                const bool lt0 = v < two0, lt1 = v < two1;
                two1 = lt0 ? two0 : (lt1 ? v : two1);
                two0 = lt0 ? v : two0;
            });
            uint64_t two[2] = {(uint64_t)two0 << 32 | (uint32_t)lane,
                               (uint64_t)two1 << 32 | (uint32_t)(lane + 64)};
            reg_sort<2>(two);
            const uint64_t v = k - 1 < 64 ? two[0] : two[1];
            th = (uint32_t)__builtin_amdgcn_readlane((int32_t)(v >> 32), (k - 1) & 63);
        }
        int32_t count = 0;
        // appends the passing lanes of one 64-material step to the buffer,
        // first cutting the buffer to its k smallest when it would overflow
        auto take = [&](uint32_t d, int32_t first, uint64_t mask) {
            if (count + __popcll(mask) > kBuf) {
                sort_buffer(buf, count);
                count = k;
                th = (uint32_t)(buf[k - 1] >> 32);
                mask &= lanes_ule(d, th);
            }
            if ((mask >> lane) & 1ull)
                buf[count + __popcll(mask & lanes_below)] = make_key(d, first + lane);
            count += __popcll(mask);
        };
        scan([&](uint32_t d, int32_t first, uint64_t valid) {
            const uint64_t mask = lanes_ule(d, th) & valid;
            if (mask) take(d, first, mask);
        });
        sort_buffer(buf, count);
        // answer: the first k keys, ascending
        const int32_t out_stride = s.out_stride;
        for (int32_t r = lane; r < out_stride; r += 64) {
            const int64_t o = (int64_t)q * out_stride + r;
            if (r < k) {
                const uint64_t key = buf[r];
                const int32_t m = (int32_t)(uint32_t)(key & 0xffffffffu);
                const float d = __uint_as_float((uint32_t)(key >> 32));
                prev[r] = m;
                gstore(s.out_idx, o, m);
                if (s.out_d) gstore(s.out_d, o, d);
                if (s.out_lab)
                    s.out_lab[o] = make_float4(gload(mL, m), gload(mA, m), gload(mB, m), d);
            } else if (s.out_lab) {
                // padded (chain) rows: ranks beyond k, -1 with an infinite distance
                gstore(s.out_idx, o, (int32_t)-1);
                s.out_lab[o] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0x7f800000));
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

__global__ __launch_bounds__(64) void topk_batch_kernel(const TopkSeg* __restrict__ segs,
                                                        const int32_t* __restrict__ starts,
                                                        int32_t n_segs, int32_t run) {
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

void launch_topk(const float* mL, const float* mA, const float* mB, int32_t n, const float* query,
                 int32_t nq, int32_t k, int32_t* out_idx, float* out_d, float4* out_lab,
                 int32_t out_stride, hipStream_t stream) {
    if (k > n) k = n;
    if (nq <= 0 || n <= 0 || k <= 0) return;
    const TopkSeg s{mL, mA, mB, query, out_idx, out_d, out_lab, n, nq, k, out_stride};
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
