// k_topk.hip -- K nearest materials per cell (the kiddo k-d tree query
// nearest_n::<SquaredEuclidean>(lab, min(N,100)) at src/main.rs:348-352).
//
// Brute force over all N materials, exact: each material gets the 64-bit
// key (float_bits(sqdist) << 32 | index).  Squared distances are >= +0, so
// the bit pattern orders like the value and keys are unique; the K smallest
// keys are the K nearest materials ascending with ties broken by index.
//
// One 256-thread block per query cell:
//   1. squared distances for all N (VALU; cached in LDS when N <= 16384),
//   2. radix select of the K-th smallest key over 11-bit digits (LDS
//      histograms, usually 2-3 passes; early exit once a digit bin holds
//      exactly the remaining count),
//   3. gather the K keys <= threshold, bitonic sort in LDS, write ascending.
// VALU-bound: 8 flops + 1 LDS atomic per material per pass.
// The batched kernel serves many (library, query set) segments in one
// launch: block b answers query b - starts[j] of the segment j holding b.
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

constexpr int kTopkThreads = 256;
constexpr int kDigitBits = 11;
constexpr int kBins = 1 << kDigitBits;
constexpr int kBinsPerThread = kBins / kTopkThreads;
constexpr int32_t kDistCacheMax = 16384;

__device__ __forceinline__ uint64_t make_key(uint32_t dbits, int32_t i) {
    return ((uint64_t)dbits << 32) | (uint32_t)i;
}

// Exclusive prefix sum over the block (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wid; ++w) base += wsum[w];
    __syncthreads();
    return base + inc - v;
}

struct TopkShared {
    uint32_t hist[kBins];
    uint64_t sel[128];
    uint32_t wsum[4];
    uint32_t digit, need, cnt, fill;
};

// K nearest of query q of segment s (one 256-thread block).
template <bool kCache>
__device__ __forceinline__ void topk_block(const TopkSeg& s, int32_t q, TopkShared& sh,
                                           uint32_t* dcache) {
    const int tid = threadIdx.x;
    const float* __restrict__ mL = s.mL;
    const float* __restrict__ mA = s.mA;
    const float* __restrict__ mB = s.mB;
    const int32_t n = s.n, k = s.k;
    const float ql = s.query[3 * (int64_t)q], qa = s.query[3 * (int64_t)q + 1],
                qb = s.query[3 * (int64_t)q + 2];

    auto dist_bits = [&](int32_t i) -> uint32_t {
        if (kCache) return dcache[i];
        return __float_as_uint(sq_distance(ql, qa, qb, mL[i], mA[i], mB[i]));
    };

    if (kCache) {
        for (int32_t i = tid; i < n; i += kTopkThreads)
            dcache[i] = __float_as_uint(sq_distance(ql, qa, qb, mL[i], mA[i], mB[i]));
    }
    if (tid < 128) sh.sel[tid] = ~0ull;
    __syncthreads();

    if (n <= 128) {
        if (tid < n) sh.sel[tid] = make_key(dist_bits(tid), tid);
        __syncthreads();
    } else {
        uint64_t prefix = 0;
        uint32_t need = (uint32_t)k;
        int lo = 0;
        for (int hi = 62; hi >= 0; hi -= kDigitBits) {
            const int lo_w = hi - kDigitBits + 1 > 0 ? hi - kDigitBits + 1 : 0;
            const int width = hi - lo_w + 1;
            const uint32_t mask = (1u << width) - 1u;
            for (int b = tid; b < kBins; b += kTopkThreads) sh.hist[b] = 0;
            __syncthreads();
            for (int32_t i = tid; i < n; i += kTopkThreads) {
                const uint64_t key = make_key(dist_bits(i), i);
                if ((key >> (hi + 1)) == prefix) atomicAdd(&sh.hist[(uint32_t)(key >> lo_w) & mask], 1u);
            }
            __syncthreads();
            uint32_t local[kBinsPerThread];
            uint32_t sum = 0;
#pragma unroll
            for (int b = 0; b < kBinsPerThread; ++b) {
                local[b] = sh.hist[tid * kBinsPerThread + b];
                sum += local[b];
            }
            const uint32_t before = block_exclusive_scan(sum, sh.wsum);
            if (before < need && need <= before + sum) {
                uint32_t c = before;
#pragma unroll
                for (int b = 0; b < kBinsPerThread; ++b) {
                    if (need > c && need <= c + local[b]) {
                        sh.digit = tid * kBinsPerThread + b;
                        sh.need = need - c;
                        sh.cnt = local[b];
                    }
                    c += local[b];
                }
            }
            __syncthreads();
            prefix = (prefix << width) | sh.digit;
            need = sh.need;
            lo = lo_w;
            const bool done = sh.cnt == need;
            __syncthreads();
            if (done) break;
        }
        if (tid == 0) sh.fill = 0;
        __syncthreads();
        for (int32_t i = tid; i < n; i += kTopkThreads) {
            const uint64_t key = make_key(dist_bits(i), i);
            if ((key >> lo) <= prefix) {
                const uint32_t slot = atomicAdd(&sh.fill, 1u);
                if (slot < 128) sh.sel[slot] = key;
            }
        }
        __syncthreads();
    }

    // bitonic sort of 128 keys, ascending
    for (int size = 2; size <= 128; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (tid < 64) {
                const int i = 2 * stride * (tid / stride) + (tid % stride);
                const int j = i + stride;
                const bool asc = (i & size) == 0;
                const uint64_t a = sh.sel[i], b = sh.sel[j];
                if ((a > b) == asc) {
                    sh.sel[i] = b;
                    sh.sel[j] = a;
                }
            }
            __syncthreads();
        }
    }
    const int32_t kk = k < n ? k : n;
    const int32_t out_stride = s.out_stride;
    // ranks beyond k hold the -1 sentinel in padded (chain) rows
    if (s.out_lab && tid >= kk && tid < out_stride) s.out_idx[(int64_t)q * out_stride + tid] = -1;
    if (tid < kk) {
        const uint64_t key = sh.sel[tid];
        const int32_t m = (int32_t)(uint32_t)(key & 0xffffffffu);
        const float d = __uint_as_float((uint32_t)(key >> 32));
        const int64_t o = (int64_t)q * out_stride + tid;
        s.out_idx[o] = m;
        if (s.out_d) s.out_d[o] = d;
        if (s.out_lab) s.out_lab[o] = make_float4(mL[m], mA[m], mB[m], d);
    }
}

template <bool kCache>
__global__ __launch_bounds__(kTopkThreads) void topk_kernel(TopkSeg s) {
    extern __shared__ uint32_t dcache[];
    __shared__ TopkShared sh;
    topk_block<kCache>(s, blockIdx.x, sh, dcache);
}

template <bool kCache>
__global__ __launch_bounds__(kTopkThreads) void topk_batch_kernel(const TopkSeg* __restrict__ segs,
                                                                  const int32_t* __restrict__ starts,
                                                                  int32_t n_segs) {
    extern __shared__ uint32_t dcache[];
    __shared__ TopkShared sh;
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    topk_block<kCache>(segs[j], blockIdx.x - starts[j], sh, dcache);
}

void launch_topk(const float* mL, const float* mA, const float* mB, int32_t n, const float* query,
                 int32_t nq, int32_t k, int32_t* out_idx, float* out_d, float4* out_lab,
                 int32_t out_stride, hipStream_t stream) {
    if (k > n) k = n;
    if (nq <= 0 || n <= 0 || k <= 0) return;
    const TopkSeg s{mL, mA, mB, query, out_idx, out_d, out_lab, n, nq, k, out_stride};
    if (n <= kDistCacheMax) {
        const size_t shmem = sizeof(uint32_t) * (size_t)n;
        hipLaunchKernelGGL(topk_kernel<true>, dim3(nq), dim3(kTopkThreads), shmem, stream, s);
    } else {
        hipLaunchKernelGGL(topk_kernel<false>, dim3(nq), dim3(kTopkThreads), 0, stream, s);
    }
}

// Segments with n > 0 only (a block per query); max_n decides the distance
// cache for the whole launch.
void launch_topk_batch(const TopkSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                       int32_t total_queries, int32_t max_n, hipStream_t stream) {
    if (n_segs <= 0 || total_queries <= 0) return;
    if (max_n <= kDistCacheMax) {
        const size_t shmem = sizeof(uint32_t) * (size_t)max_n;
        hipLaunchKernelGGL(topk_batch_kernel<true>, dim3(total_queries), dim3(kTopkThreads), shmem,
                           stream, d_segs, d_starts, n_segs);
    } else {
        hipLaunchKernelGGL(topk_batch_kernel<false>, dim3(total_queries), dim3(kTopkThreads), 0,
                           stream, d_segs, d_starts, n_segs);
    }
}

hipError_t raise_dynamic_lds(const void* kernel, int max_lds) {
    hipFuncAttributes attr;
    hipError_t e = hipFuncGetAttributes(&attr, kernel);
    if (e != hipSuccess) return e;
    const int dyn = max_lds - (int)attr.sharedSizeBytes;
    if (dyn <= 0) return hipErrorInvalidValue;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
}

hipError_t topk_configure(int max_lds) {
    hipError_t e = raise_dynamic_lds(reinterpret_cast<const void*>(&topk_kernel<true>), max_lds);
    if (e != hipSuccess) return e;
    return raise_dynamic_lds(reinterpret_cast<const void*>(&topk_batch_kernel<true>), max_lds);
}

hipError_t configure_kernel_limits(int max_lds) {
    hipError_t e = topk_configure(max_lds);
    if (e == hipSuccess) e = place_configure(max_lds);
    if (e == hipSuccess) e = anneal_configure(max_lds);
    if (e == hipSuccess) e = resize_configure(max_lds);
    return e;
}

}  // namespace mh
