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

constexpr int kCellsPerBlock = 16;
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

    __syncthreads();
    for (int32_t base = 0; base < npx; base += kChunk) {
        const int32_t len = min(kChunk, npx - base);
        for (int i = tid; i < kCellsPerBlock * kChunk; i += kLabThreads) {
            const int g = i / kChunk, p = i % kChunk;
            const int32_t cell = cell0 + g;
            if (cell < n_cells && p < len) {
                const int32_t pix = base + p;
                const int32_t py = pix / cw, px = pix - py * cw;
                const uint8_t* src = s.rgb + (int64_t)(cell / per_row) * s.block_row_stride +
                                     (int64_t)(cell % per_row) * s.block_col_stride +
                                     (int64_t)py * s.row_stride + (int64_t)px * 3;
                float L, A, B;
                rgb8_to_lab(sh.lut, src[0], src[1], src[2], L, A, B);
                sh.lab[0][g][p] = L;
                sh.lab[1][g][p] = A;
                sh.lab[2][g][p] = B;
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
                     int32_t n_cells, const float* lut, float* out, hipStream_t stream) {
    if (n_cells <= 0) return;
    const CellSeg s{rgb, row_stride, block_row_stride, block_col_stride, per_row, cw, ch, n_cells,
                    out, nullptr};
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
