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
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

constexpr int kCellsPerBlock = 16;
constexpr int kChunk = 256;
constexpr int kPitch = kChunk + 1;  // odd pitch: the 48 summing lanes hit distinct banks
constexpr int kLabThreads = 256;

// Cell c's pixel (px, py) lives at
//   rgb + (c / per_row) * block_row_stride + (c % per_row) * block_col_stride
//       + py * row_stride + px * 3.
__global__ __launch_bounds__(kLabThreads) void cell_lab_kernel(
    const uint8_t* __restrict__ rgb, int64_t row_stride, int32_t per_row, int64_t block_row_stride,
    int64_t block_col_stride, int32_t cw, int32_t ch, int32_t n_cells,
    const float* __restrict__ lut_g, float* __restrict__ out) {
    __shared__ float lut[256];
    __shared__ float lab_s[3][kCellsPerBlock][kPitch];

    const int tid = threadIdx.x;
    lut[tid] = lut_g[tid];

    const int32_t cell0 = blockIdx.x * kCellsPerBlock;
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
                const uint8_t* src = rgb + (int64_t)(cell / per_row) * block_row_stride +
                                     (int64_t)(cell % per_row) * block_col_stride +
                                     (int64_t)py * row_stride + (int64_t)px * 3;
                float L, A, B;
                rgb8_to_lab(lut, src[0], src[1], src[2], L, A, B);
                lab_s[0][g][p] = L;
                lab_s[1][g][p] = A;
                lab_s[2][g][p] = B;
            }
        }
        __syncthreads();
        if (summer) {
            const float* v = lab_s[sum_ch][sum_cell];
            for (int32_t p = 0; p < len; ++p) acc = acc + v[p];
        }
        __syncthreads();
    }
    if (summer) {
        const float total = (float)((uint32_t)cw * (uint32_t)ch);
        out[(int64_t)(cell0 + sum_cell) * 3 + sum_ch] = acc / total;
    }
}

void launch_cell_lab(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                     int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                     int32_t n_cells, const float* lut, float* out, hipStream_t stream) {
    if (n_cells <= 0) return;
    const int blocks = (n_cells + kCellsPerBlock - 1) / kCellsPerBlock;
    hipLaunchKernelGGL(cell_lab_kernel, dim3(blocks), dim3(kLabThreads), 0, stream, rgb, row_stride,
                       per_row, block_row_stride, block_col_stride, cw, ch, n_cells, lut, out);
}

}  // namespace mh
