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
//   * tau starts as the largest of this cell's keys to the previous cell's K
//     answers (K materials have a key <= tau, so the K smallest do too);
//     neighbouring cells have similar Lab means, so tau is tight;
//   * one pass over the N materials appends every key <= tau (ballot +
//     mbcnt compaction, 64 materials per step);
//   * when the buffer would overflow it is sorted and cut to its K smallest,
//     whose largest becomes the new (tighter) tau;
//   * at the end the buffer is sorted (bitonic network in LDS) and the first
//     K keys are the answer.
// The first cell of a run starts from tau = +inf.  The cost is one distance
// evaluation per material plus a few small sorts, with no histograms or
// atomics.  The batched kernel serves many (library, query set) segments in
// one launch: block b handles run b - starts[j] of the segment j holding b.
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

constexpr int kBuf = 256;      // candidate buffer (keys) per wave
constexpr int kMaxK = 128;     // largest K served
constexpr int kMaxRun = 8;     // query cells per wave

__device__ __forceinline__ uint64_t make_key(uint32_t dbits, int32_t i) {
    return ((uint64_t)dbits << 32) | (uint32_t)i;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Ascending bitonic sort of buf[0..P) by one wave (P = 128 or 256); entries
// [count, P) are first filled with the largest key.  The network is fully
// unrolled, so every index is a compile-time function of the lane.
template <int P>
__device__ __forceinline__ void wave_sort(uint64_t* buf, int32_t count) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int r = 0; r < P / 64; ++r)
        if (r * 64 + lane >= count) buf[r * 64 + lane] = ~0ull;
    __syncthreads();
#pragma unroll
    for (int size = 2; size <= P; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int h = 0; h < P / 128; ++h) {
                const int t = h * 64 + lane;
                const int i = ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const uint64_t a = buf[i], b = buf[j];
                if ((a > b) == asc) {
                    buf[i] = b;
                    buf[j] = a;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ void sort_buffer(uint64_t* buf, int32_t count) {
    if (count <= 128)
        wave_sort<128>(buf, count);
    else
        wave_sort<256>(buf, count);
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

    for (int32_t q = q0; q < q_end; ++q) {
        const float ql = gload(s.query, 3 * (int64_t)q), qa = gload(s.query, 3 * (int64_t)q + 1),
                    qb = gload(s.query, 3 * (int64_t)q + 2);
        auto dist_bits = [&](int32_t i) {
            return __float_as_uint(sq_distance(ql, qa, qb, gload(mL, i), gload(mA, i), gload(mB, i)));
        };
        // threshold from the previous cell's answers
        uint64_t tau = ~0ull;
        if (q > q0) {
            uint64_t m = 0;
            for (int32_t r = lane; r < k; r += 64) {
                const int32_t i = prev[r];
                const uint64_t kk = make_key(dist_bits(i), i);
                m = kk > m ? kk : m;
            }
            tau = wave_max_u64(m);
        }
        int32_t count = 0;
        // appends the keys <= tau of one 64-material step to the buffer,
        // first cutting the buffer to its k smallest when it would overflow
        auto take = [&](uint64_t key, bool valid) {
            bool pass = valid && key <= tau;
            uint64_t mask = wave_ballot(pass);
            if (mask == 0) return;
            if (count + __popcll(mask) > kBuf) {
                sort_buffer(buf, count);
                count = k;
                tau = buf[k - 1];
                pass = pass && key <= tau;
                mask = wave_ballot(pass);
            }
            if (pass) buf[count + __popcll(mask & lanes_below)] = key;
            count += __popcll(mask);
        };
        // two steps per iteration; loads are clamped in bounds and issued
        // together, the keys of lanes past n are masked out
        for (int32_t base = 0; base < n; base += 128) {
            const int32_t i0 = base + lane, i1 = i0 + 64;
            const uint32_t d0 = dist_bits(i0 < n ? i0 : n - 1);
            const uint32_t d1 = dist_bits(i1 < n ? i1 : n - 1);
            take(make_key(d0, i0), i0 < n);
            take(make_key(d1, i1), i1 < n);
        }
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
// launch wants enough waves to fill the chip (~8 per CU).
int32_t topk_run_length(int64_t total_queries) {
    int64_t r = total_queries / (256 * 8);
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
