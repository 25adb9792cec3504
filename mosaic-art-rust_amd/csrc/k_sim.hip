// k_sim.hip -- the material x material Lab distance matrix
// (SimilarityDatabase::build_similarities, src/similarity.rs:69-82, with
// calculate_lab_distance :141-146).
//
// HBM write-bound: 4 B stored per entry, 24 B of Lab read per 256 x 32
// tile.  Each 256-thread block owns a 32-row x 256-column tile; a thread
// keeps 4 column Labs in registers and streams float4 rows with
// non-temporal stores (the matrix is written once and read rarely).
// Bits: dist(i,j) == dist(j,i) and dist(i,i) == 0, so the full square equals
// the reference's packed triangle plus its mirror and zero diagonal.
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kSimRows = 32;
constexpr int kSimCols = 256;

template <bool kVec4>
__global__ __launch_bounds__(256) void sim_full_kernel(const float4* __restrict__ lab, int32_t n,
                                                       float* __restrict__ out) {
    const int tid = threadIdx.x;
    const int32_t col0 = blockIdx.x * kSimCols + (tid & 63) * 4;
    const int32_t row0 = blockIdx.y * kSimRows + (tid >> 6);
    float4 c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = lab[min(col0 + j, n - 1)];
#pragma unroll 4
    for (int r = 0; r < kSimRows; r += 4) {
        const int32_t row = row0 + r;
        if (row >= n) break;
        const float4 p = lab[row];
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = lab_distance(p, c[j]);
        float* dst = out + (int64_t)row * n + col0;
        if (kVec4 && col0 + 3 < n) {
            const v4f q = {v[0], v[1], v[2], v[3]};
            __builtin_nontemporal_store(q, reinterpret_cast<v4f*>(dst));
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (col0 + j < n) __builtin_nontemporal_store(v[j], dst + j);
        }
    }
}

// Packed upper triangle in the reference's order: row i holds j = i+1..n-1
// at i*n - i*(i+1)/2 + (j - i - 1) (src/similarity.rs:100).
__global__ __launch_bounds__(256) void sim_packed_kernel(const float4* __restrict__ lab, int32_t n,
                                                         float* __restrict__ out) {
    const int32_t i = blockIdx.x;
    const int64_t off = (int64_t)i * n - (int64_t)i * (i + 1) / 2;
    const float4 p = lab[i];
    for (int32_t j = i + 1 + threadIdx.x; j < n; j += 256) out[off + (j - i - 1)] = lab_distance(p, lab[j]);
}

// Adjacency-term matrix term[i][j] = 1/(dist(i,j) + 1): the value every
// penalty / cost / swap-delta term takes (src/adjacency.rs:82, :110, :174),
// bit-identical to computing it on the fly.  Same tiling as the distance
// matrix.
__global__ __launch_bounds__(256) void term_full_kernel(const float4* __restrict__ lab, int32_t n,
                                                        float* __restrict__ out) {
    const int tid = threadIdx.x;
    const int32_t col0 = blockIdx.x * kSimCols + (tid & 63);
    const int32_t row0 = blockIdx.y * kSimRows + (tid >> 6);
    float4 c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = lab[min(col0 + 64 * j, n - 1)];
    for (int r = 0; r < kSimRows; r += 4) {
        const int32_t row = row0 + r;
        if (row >= n) break;
        const float4 p = lab[row];
        float* dst = out + (int64_t)row * n;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (col0 + 64 * j < n) dst[col0 + 64 * j] = adj_term(p, c[j]);
    }
}

void launch_term_full(const float4* lab, int32_t n, float* out, hipStream_t stream) {
    if (n <= 0) return;
    dim3 grid((n + kSimCols - 1) / kSimCols, (n + kSimRows - 1) / kSimRows);
    hipLaunchKernelGGL(term_full_kernel, grid, dim3(256), 0, stream, lab, n, out);
}

void launch_sim_full(const float4* lab, int32_t n, float* out, hipStream_t stream) {
    if (n <= 0) return;
    dim3 grid((n + kSimCols - 1) / kSimCols, (n + kSimRows - 1) / kSimRows);
    if (n % 4 == 0)
        hipLaunchKernelGGL(sim_full_kernel<true>, grid, dim3(256), 0, stream, lab, n, out);
    else
        hipLaunchKernelGGL(sim_full_kernel<false>, grid, dim3(256), 0, stream, lab, n, out);
}

void launch_sim_packed(const float4* lab, int32_t n, float* out, hipStream_t stream) {
    if (n <= 1) return;
    hipLaunchKernelGGL(sim_packed_kernel, dim3(n - 1), dim3(256), 0, stream, lab, n, out);
}

}  // namespace mh
