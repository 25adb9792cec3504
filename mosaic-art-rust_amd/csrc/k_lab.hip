// k_lab.hip -- average Lab of image cells (calculate_average_lab,
// src/lib.rs:94-119, applied per cell at src/main.rs:536-541 and per
// material at src/main.rs:280).
//
// The reference folds per-pixel Lab values serially in row-major order, so
// the f32 sums are order-dependent.  The kernel keeps that order: a
// 256-thread workgroup owns a few cells; all threads convert a chunk of
// pixels per cell to Lab (the expensive part: LUT + 3x cbrt) into LDS, then
// one lane per (cell, channel) adds the chunk in pixel order to a register
// accumulator.  Chunks repeat until the cell is done.
//
// The batched kernel runs the same blocks for many images at once: block b
// belongs to the segment j with starts[j] <= b < starts[j+1].
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

// A block owns kCells cells and converts kChunk pixels of each per round
// (kCells * kChunk = 1024: four pixels per thread).  4 cells of 256-pixel
// chunks for small launches: the double-precision cube roots are dependent
// chains, so enough blocks to spread over the device matter more than the
// block's setup.  16 cells of 64-pixel chunks for large launches: the serial
// fold runs on 48 lanes of one wave instead of 12, a quarter of the summing
// instructions per cell.
constexpr int kLabThreads = 256;
constexpr int kLabPixels = 1024;  // pixels converted per block and round
// Cell g's chunk starts at g * (kChunk + 1) of a channel's array (an odd
// pitch, so the summing lanes of one channel hit distinct banks); the
// arrays are 1048 floats apart (at most 16 * 65 used), which offsets the
// channels by 24 banks.
constexpr int kLabArray = 1048;

struct LabShared {
    float lut[256];
    float lab[3][kLabArray];
};

// Cell c's pixel (px, py) lives at
//   rgb + (c / per_row) * block_row_stride + (c % per_row) * block_col_stride
//       + py * row_stride + px * 3.
// Pixel index -> (px, py) by umulhi with cw_magic = floor((2^32 - 1) / cw) + 1,
// exact while pix * cw < 2^32 (pix < cw * ch), else by division.
template <int kCells, bool kOff32>
__device__ __forceinline__ void cell_lab_block(const CellSeg& s, int32_t blk,
                                               const float* __restrict__ lut_g, LabShared& sh) {
    constexpr int kChunk = kLabPixels / kCells;   // pixels per cell and round
    constexpr int kPer = kLabPixels / kLabThreads;  // pixels per thread and round
    constexpr int kSpan = kLabThreads / kChunk;     // cells covered by one pass of the threads
    static_assert(kChunk * kCells == kLabPixels && kSpan * kChunk == kLabThreads, "layout");
    static_assert(kCells * (kChunk + 1) <= kLabArray, "chunks fit a channel's array");
    const int tid = threadIdx.x;
    sh.lut[tid] = gload(lut_g, tid);

    const int32_t cell0 = blk * kCells;
    const int32_t cw = s.cw, ch = s.ch, n_cells = s.n_cells, per_row = s.per_row;
    const int32_t npx = cw * ch;
    const uint32_t cw_magic = 0xffffffffu / (uint32_t)cw + 1u;
    const bool magic_ok = (uint64_t)(uint32_t)npx * (uint32_t)cw < (1ull << 32);
    const uint32_t rs32 = (uint32_t)s.row_stride;
    // summing lane -> (cell, channel)
    const int sum_cell = tid / 3, sum_ch = tid % 3;
    const bool summer = tid < 3 * kCells && cell0 + sum_cell < n_cells;
    float acc = 0.0f;
    // this thread's pixel p_own of cells g_j = j * kSpan + wv, wv = tid / kChunk;
    // kChunk is a multiple of 64, so wv is the same on every lane of a wave
    // and the cells' base addresses are scalar
    static_assert(kChunk % 64 == 0, "a wave's lanes share their cells");
    const int p_own = tid % kChunk;
    const int wv = __builtin_amdgcn_readfirstlane(tid / kChunk);
    // integer RGB sums of the pixels this thread converts (calculate_average_rgb,
    // src/main.rs:687-708, when out_rgb is requested)
    const bool want_rgb = s.out_rgb != nullptr;
    uint32_t rsum[kPer][3];
    const MH_GAS uint8_t* cell_base[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        rsum[j][0] = rsum[j][1] = rsum[j][2] = 0u;
        const int32_t cell = cell0 + j * kSpan + wv;
        const int32_t c = cell < n_cells ? cell : n_cells - 1;
        cell_base[j] = (const MH_GAS uint8_t*)s.rgb + (int64_t)(c / per_row) * s.block_row_stride +
                       (int64_t)(c % per_row) * s.block_col_stride;
    }

    __syncthreads();
    for (int32_t base = 0; base < npx; base += kChunk) {
        const int32_t len = min(kChunk, npx - base);
        // the pixel's offset within its cell, the same for the kPer cells
        const int32_t pix = base + p_own;
        const int32_t py = magic_ok ? (int32_t)__umulhi((uint32_t)pix, cw_magic) : pix / cw;
        const int32_t px = pix - py * cw;
        const uint32_t o32 = (uint32_t)py * rs32 + (uint32_t)px * 3u;
        const int64_t o64 = (int64_t)py * s.row_stride + px * 3;
        // every pixel's loads first (raw values kept until the conversion,
        // so they are all in flight together), then the conversions
        typedef uint16_t u16_any __attribute__((aligned(1)));
        uint32_t rg[kPer], bk[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            rg[j] = bk[j] = 0u;
            if (cell0 + j * kSpan + wv < n_cells && p_own < len) {
                const MH_GAS uint8_t* src = kOff32 ? cell_base[j] + o32 : cell_base[j] + o64;
                rg[j] = *(const MH_GAS u16_any*)src;
                bk[j] = src[2];
            }
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int g = j * kSpan + wv;
            if (cell0 + g < n_cells && p_own < len) {
                const uint32_t r = rg[j] & 0xffu, gg = rg[j] >> 8, b = bk[j];
                float L, A, B;
                rgb8_to_lab(sh.lut, r, gg, b, L, A, B);
                const int slot = g * (kChunk + 1) + p_own;  // cell g's chunk, pixel order
                sh.lab[0][slot] = L;
                sh.lab[1][slot] = A;
                sh.lab[2][slot] = B;
                if (want_rgb) {
                    rsum[j][0] += r;
                    rsum[j][1] += gg;
                    rsum[j][2] += b;
                }
            }
        }
        __syncthreads();
        if (summer) {
            const float* v = sh.lab[sum_ch] + sum_cell * (kChunk + 1);
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
        // now), then one lane per (cell, channel) adds the kChunk partials
        // of its cell; integer sums, so the order is free
        uint32_t* part = reinterpret_cast<uint32_t*>(&sh.lab[0][0]);
        static_assert(sizeof(sh.lab) >= kPer * 3 * (kLabThreads + 1) * sizeof(uint32_t),
                      "partials fit the staging area");
#pragma unroll
        for (int j = 0; j < kPer; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) part[(j * 3 + c) * (kLabThreads + 1) + tid] = rsum[j][c];
        __syncthreads();
        if (summer) {
            const int j = sum_cell / kSpan, t0 = (sum_cell % kSpan) * kChunk;
            const uint32_t* v = part + (j * 3 + sum_ch) * (kLabThreads + 1) + t0;
            uint64_t t = 0;
            for (int i = 0; i < kChunk; ++i) t += v[i];
            s.out_rgb[(int64_t)(cell0 + sum_cell) * 3 + sum_ch] = (uint8_t)(t / (uint64_t)npx);
        }
    }
}

// kOff32: a pixel's byte offset within its cell fits 31 bits (a uniform
// choice per segment), so loads take the cell's scalar base plus a 32-bit
// offset.
template <int kCells>
__device__ __forceinline__ void cell_lab_run(const CellSeg& s, int32_t blk,
                                             const float* __restrict__ lut_g, LabShared& sh) {
    if (s.row_stride >= 0 && (uint64_t)s.ch * (uint64_t)s.row_stride < (1ull << 31))
        cell_lab_block<kCells, true>(s, blk, lut_g, sh);
    else
        cell_lab_block<kCells, false>(s, blk, lut_g, sh);
}

template <int kCells>
__global__ __launch_bounds__(kLabThreads) void cell_lab_kernel(CellSeg s,
                                                               const float* __restrict__ lut) {
    __shared__ LabShared sh;
    cell_lab_run<kCells>(s, blockIdx.x, lut, sh);
}

template <int kCells>
__global__ __launch_bounds__(kLabThreads) void cell_lab_batch_kernel(
    const CellSeg* __restrict__ segs, const int32_t* __restrict__ starts, int32_t n_segs,
    const float* __restrict__ lut) {
    __shared__ LabShared sh;
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    cell_lab_run<kCells>(segs[j], blockIdx.x - starts[j], lut, sh);
}

int32_t cell_lab_cells_per_block(int64_t total_cells) { return total_cells >= 65536 ? 16 : 4; }

int32_t cell_lab_blocks(int32_t n_cells, int32_t cells_per_block) {
    return n_cells > 0 ? (n_cells + cells_per_block - 1) / cells_per_block : 0;
}

void launch_cell_lab(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                     int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                     int32_t n_cells, const float* lut, float* out, hipStream_t stream,
                     uint8_t* out_rgb) {
    if (n_cells <= 0) return;
    const CellSeg s{rgb, row_stride, block_row_stride, block_col_stride, per_row, cw, ch, n_cells,
                    out, out_rgb};
    const int32_t cpb = cell_lab_cells_per_block(n_cells);
    const dim3 grid(cell_lab_blocks(n_cells, cpb));
    if (cpb == 16)
        hipLaunchKernelGGL(cell_lab_kernel<16>, grid, dim3(kLabThreads), 0, stream, s, lut);
    else
        hipLaunchKernelGGL(cell_lab_kernel<4>, grid, dim3(kLabThreads), 0, stream, s, lut);
}

void launch_cell_lab_batch(const CellSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                           int32_t total_blocks, int32_t cells_per_block, const float* lut,
                           hipStream_t stream) {
    if (n_segs <= 0 || total_blocks <= 0) return;
    if (cells_per_block == 16)
        hipLaunchKernelGGL(cell_lab_batch_kernel<16>, dim3(total_blocks), dim3(kLabThreads), 0,
                           stream, d_segs, d_starts, n_segs, lut);
    else
        hipLaunchKernelGGL(cell_lab_batch_kernel<4>, dim3(total_blocks), dim3(kLabThreads), 0,
                           stream, d_segs, d_starts, n_segs, lut);
}

}  // namespace mh
