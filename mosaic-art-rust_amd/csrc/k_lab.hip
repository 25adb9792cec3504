// k_lab.hip -- average Lab of image cells (calculate_average_lab,
// src/lib.rs:94-119, applied per cell at src/main.rs:536-541 and per
// material at src/main.rs:280).
//
// The reference folds per-pixel Lab values serially in row-major order, so
// the f32 sums are order-dependent.  The kernel keeps that order: a
// 256-thread workgroup owns kCellsPerBlock cells; all threads convert a
// chunk of kChunk pixels per cell to Lab (the expensive part: LUT + 3x cbrt)
// into LDS, then one lane per (cell, channel) adds the chunk in pixel order
// to a register accumulator.  Chunks repeat until the cell is done.
//
// The batched kernel runs the same blocks for many images at once: block b
// belongs to the segment j with starts[j] <= b < starts[j+1].
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

// 4 cells per block: the double-precision cube roots are dependent chains,
// so enough waves per SIMD to hide them matter more than amortising the
// block's setup (16 cells per block ran 1.6x longer on an 80x60 grid).
constexpr int kCellsPerBlock = 4;
constexpr int kChunk = 256;
constexpr int kPitch = kChunk + 1;  // odd pitch: the 48 summing lanes hit distinct banks
constexpr int kLabThreads = 256;

struct LabShared {
    float lut[256];
    float lab[3][kCellsPerBlock][kPitch];
};

// Cell c's pixel (px, py) lives at
//   rgb + (c / per_row) * block_row_stride + (c % per_row) * block_col_stride
//       + py * row_stride + px * 3.
__device__ __forceinline__ void cell_lab_block(const CellSeg& s, int32_t blk,
                                               const float* __restrict__ lut_g, LabShared& sh) {
    const int tid = threadIdx.x;
    sh.lut[tid] = lut_g[tid];

    const int32_t cell0 = blk * kCellsPerBlock;
    const int32_t cw = s.cw, ch = s.ch, n_cells = s.n_cells, per_row = s.per_row;
    const int32_t npx = cw * ch;
    // summing lane -> (cell, channel)
    const int sum_cell = tid / 3, sum_ch = tid % 3;
    const bool summer = tid < 3 * kCellsPerBlock && cell0 + sum_cell < n_cells;
    float acc = 0.0f;
    // integer RGB sums of the pixels this thread converts (calculate_average_rgb,
    // src/main.rs:687-708, when out_rgb is requested): thread t sees pixel t of
    // every chunk of each of the block's cells
    const bool want_rgb = s.out_rgb != nullptr;
    uint32_t rsum[kCellsPerBlock][3];
#pragma unroll
    for (int g = 0; g < kCellsPerBlock; ++g) rsum[g][0] = rsum[g][1] = rsum[g][2] = 0u;

    __syncthreads();
    for (int32_t base = 0; base < npx; base += kChunk) {
        const int32_t len = min(kChunk, npx - base);
        static_assert(kChunk == kLabThreads, "one pixel per thread and cell per chunk");
#pragma unroll
        for (int g = 0; g < kCellsPerBlock; ++g) {
            const int p = tid;
            const int32_t cell = cell0 + g;
            if (cell < n_cells && p < len) {
                const int32_t pix = base + p;
                const int32_t py = pix / cw, px = pix - py * cw;
                const uint8_t* src = s.rgb + (int64_t)(cell / per_row) * s.block_row_stride +
                                     (int64_t)(cell % per_row) * s.block_col_stride +
                                     (int64_t)py * s.row_stride + (int64_t)px * 3;
                const uint8_t r = src[0], gg = src[1], b = src[2];
                float L, A, B;
                rgb8_to_lab(sh.lut, r, gg, b, L, A, B);
                sh.lab[0][g][p] = L;
                sh.lab[1][g][p] = A;
                sh.lab[2][g][p] = B;
                if (want_rgb) {
                    rsum[g][0] += r;
                    rsum[g][1] += gg;
                    rsum[g][2] += b;
                }
            }
        }
        __syncthreads();
        if (summer) {
            const float* v = sh.lab[sum_ch][sum_cell];
            for (int32_t p = 0; p < len; ++p) acc = acc + v[p];
        }
        __syncthreads();
    }
    if (summer) {
        const float total = (float)((uint32_t)cw * (uint32_t)ch);
        s.out[(int64_t)(cell0 + sum_cell) * 3 + sum_ch] = acc / total;
    }
    if (want_rgb) {
        // per-thread partial sums through LDS (the Lab staging area is free
        // now), then one lane per (cell, channel) adds the 256 partials;
        // integer sums, so the order is free
        uint32_t* part = reinterpret_cast<uint32_t*>(&sh.lab[0][0][0]);
        static_assert(sizeof(sh.lab) >= 3 * kCellsPerBlock * (kLabThreads + 1) * sizeof(uint32_t),
                      "partials fit the staging area");
#pragma unroll
        for (int g = 0; g < kCellsPerBlock; ++g)
#pragma unroll
            for (int c = 0; c < 3; ++c) part[(g * 3 + c) * (kLabThreads + 1) + tid] = rsum[g][c];
        __syncthreads();
        if (summer) {
            const uint32_t* v = part + (sum_cell * 3 + sum_ch) * (kLabThreads + 1);
            uint64_t t = 0;
            for (int i = 0; i < kLabThreads; ++i) t += v[i];
            s.out_rgb[(int64_t)(cell0 + sum_cell) * 3 + sum_ch] = (uint8_t)(t / (uint64_t)npx);
        }
    }
}

__global__ __launch_bounds__(kLabThreads) void cell_lab_kernel(CellSeg s,
                                                               const float* __restrict__ lut) {
    __shared__ LabShared sh;
    cell_lab_block(s, blockIdx.x, lut, sh);
}

__global__ __launch_bounds__(kLabThreads) void cell_lab_batch_kernel(
    const CellSeg* __restrict__ segs, const int32_t* __restrict__ starts, int32_t n_segs,
    const float* __restrict__ lut) {
    __shared__ LabShared sh;
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    cell_lab_block(segs[j], blockIdx.x - starts[j], lut, sh);
}

int32_t cell_lab_blocks(int32_t n_cells) {
    return n_cells > 0 ? (n_cells + kCellsPerBlock - 1) / kCellsPerBlock : 0;
}

void launch_cell_lab(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                     int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                     int32_t n_cells, const float* lut, float* out, hipStream_t stream,
                     uint8_t* out_rgb) {
    if (n_cells <= 0) return;
    const CellSeg s{rgb, row_stride, block_row_stride, block_col_stride, per_row, cw, ch, n_cells,
                    out, out_rgb};
    hipLaunchKernelGGL(cell_lab_kernel, dim3(cell_lab_blocks(n_cells)), dim3(kLabThreads), 0,
                       stream, s, lut);
}

void launch_cell_lab_batch(const CellSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                           int32_t total_blocks, const float* lut, hipStream_t stream) {
    if (n_segs <= 0 || total_blocks <= 0) return;
    hipLaunchKernelGGL(cell_lab_batch_kernel, dim3(total_blocks), dim3(kLabThreads), 0, stream,
                       d_segs, d_starts, n_segs, lut);
}

}  // namespace mh
