// k_composite.hip -- tile resize, colour adjustment and compositing
// (MosaicGenerator::resize_image src/main.rs:710-738,
// calculate_average_rgb :687-708, calculate_optimal_adjustment
// src/color_adjustment.rs:99-154, ColorAdjustment::apply_to_image :40-96,
// and the composite loops src/main.rs:548-590 / :625-664).
//
// The resize depends only on (material, cell size), so it runs once per
// material per job instead of once per cell: a horizontal Lanczos3 pass to a
// u8 intermediate, then a vertical pass, both integer fixed-point with 16-bit
// taps and a per-axis precision that the host computes.  Compositing is one coalesced per-pixel pass: output
// pixel -> cell -> material -> resized pixel -> per-cell adjustment -> u8.
// HBM-bound: ~9 B per output pixel (3 B target read for the cell average, 3 B
// tile read, 3 B written).
#include "mh_device.hpp"
#include "mh_kernels.hpp"

namespace mh {

// A pass's sum has `prec` fraction bits: keep the integer part, clamped.
__device__ __forceinline__ uint32_t clip8(int32_t acc, int32_t prec) {
    const int32_t v = acc >> prec;
    return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// Horizontal pass: src [n][sh][sw][3] -> tmp [n][sh][dw][3].
__global__ void resize_h_kernel(const uint8_t* __restrict__ src, int32_t n, int32_t sw, int32_t sh,
                                int32_t dw, const int32_t* __restrict__ lo,
                                const int32_t* __restrict__ cnt, const int32_t* __restrict__ kk,
                                int32_t ks, int32_t prec, uint8_t* __restrict__ tmp) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)n * sh * dw;
    if (idx >= total) return;
    const int32_t o = (int32_t)(idx % dw);
    const int64_t row = idx / dw;  // m*sh + y
    const uint8_t* s = src + (row * sw + lo[o]) * 3;
    const int32_t* k = kk + (int64_t)o * ks;
    int32_t a0 = 1 << (prec - 1), a1 = a0, a2 = a0;
    const int32_t c = cnt[o];
    for (int32_t t = 0; t < c; ++t) {
        const int32_t w = k[t];
        a0 += (int32_t)s[3 * t] * w;
        a1 += (int32_t)s[3 * t + 1] * w;
        a2 += (int32_t)s[3 * t + 2] * w;
    }
    uint8_t* d = tmp + idx * 3;
    d[0] = (uint8_t)clip8(a0, prec);
    d[1] = (uint8_t)clip8(a1, prec);
    d[2] = (uint8_t)clip8(a2, prec);
}

// Vertical pass: tmp [n][sh][dw][3] -> dst [slot(m)][dh][dw][3], slot(m) =
// dst_index[m] (or m without an index).
__global__ void resize_v_kernel(const uint8_t* __restrict__ tmp, int32_t n, int32_t sh, int32_t dw,
                                int32_t dh, const int32_t* __restrict__ lo,
                                const int32_t* __restrict__ cnt, const int32_t* __restrict__ kk,
                                int32_t ks, int32_t prec, const int32_t* __restrict__ dst_index,
                                uint8_t* __restrict__ dst) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)n * dh * dw;
    if (idx >= total) return;
    const int32_t x = (int32_t)(idx % dw);
    const int64_t r = idx / dw;
    const int32_t o = (int32_t)(r % dh);
    const int64_t m = r / dh;
    const uint8_t* s = tmp + ((m * sh + lo[o]) * dw + x) * 3;
    const int32_t* k = kk + (int64_t)o * ks;
    const int64_t step = (int64_t)dw * 3;
    int32_t a0 = 1 << (prec - 1), a1 = a0, a2 = a0;
    const int32_t c = cnt[o];
    for (int32_t t = 0; t < c; ++t) {
        const int32_t w = k[t];
        a0 += (int32_t)s[t * step] * w;
        a1 += (int32_t)s[t * step + 1] * w;
        a2 += (int32_t)s[t * step + 2] * w;
    }
    const int64_t slot = dst_index ? (int64_t)dst_index[m] : m;
    uint8_t* d = dst + ((slot * dh + o) * dw + x) * 3;
    d[0] = (uint8_t)clip8(a0, prec);
    d[1] = (uint8_t)clip8(a1, prec);
    d[2] = (uint8_t)clip8(a2, prec);
}

// Both passes for one material per 256-thread block, staged in LDS: the
// source image arrives with coalesced dword loads, the horizontal pass writes
// a u8 intermediate in LDS, the vertical pass writes the tile.  Same
// arithmetic as the two-kernel path.
constexpr int kFusedResizeMaxLds = 96 * 1024;

__global__ __launch_bounds__(256) void resize_fused_kernel(
    const uint8_t* __restrict__ src, int32_t sw, int32_t sh, int32_t dw, int32_t dh,
    const int32_t* __restrict__ hlo, const int32_t* __restrict__ hcnt,
    const int32_t* __restrict__ hk, int32_t hks, int32_t hprec, const int32_t* __restrict__ vlo,
    const int32_t* __restrict__ vcnt, const int32_t* __restrict__ vk, int32_t vks, int32_t vprec,
    const int32_t* __restrict__ dst_index, uint8_t* __restrict__ dst) {
    extern __shared__ __align__(16) uint8_t rsm[];
    const int32_t m = blockIdx.x;
    const int32_t src_bytes = sw * sh * 3;
    const int32_t src_pad = (src_bytes + 15) & ~15;
    uint8_t* s_src = rsm;
    uint8_t* s_tmp = rsm + src_pad;
    const uint8_t* g = src + (int64_t)m * src_bytes;
    const int tid = threadIdx.x;
    if ((src_bytes & 3) == 0) {
        const uint32_t* g4 = reinterpret_cast<const uint32_t*>(g);
        uint32_t* s4 = reinterpret_cast<uint32_t*>(s_src);
        for (int32_t i = tid; i < src_bytes / 4; i += 256) s4[i] = g4[i];
    } else {
        for (int32_t i = tid; i < src_bytes; i += 256) s_src[i] = g[i];
    }
    __syncthreads();
    for (int32_t idx = tid; idx < sh * dw; idx += 256) {
        const int32_t y = idx / dw, o = idx - y * dw;
        const uint8_t* row = s_src + (y * sw + hlo[o]) * 3;
        const int32_t* k = hk + (int64_t)o * hks;
        int32_t a0 = 1 << (hprec - 1), a1 = a0, a2 = a0;
        const int32_t c = hcnt[o];
        for (int32_t t = 0; t < c; ++t) {
            const int32_t w = k[t];
            a0 += (int32_t)row[3 * t] * w;
            a1 += (int32_t)row[3 * t + 1] * w;
            a2 += (int32_t)row[3 * t + 2] * w;
        }
        uint8_t* d = s_tmp + idx * 3;
        d[0] = (uint8_t)clip8(a0, hprec);
        d[1] = (uint8_t)clip8(a1, hprec);
        d[2] = (uint8_t)clip8(a2, hprec);
    }
    __syncthreads();
    uint8_t* out = dst + (int64_t)(dst_index ? dst_index[m] : m) * dh * dw * 3;
    for (int32_t idx = tid; idx < dh * dw; idx += 256) {
        const int32_t o = idx / dw, x = idx - o * dw;
        const uint8_t* col = s_tmp + (vlo[o] * dw + x) * 3;
        const int32_t* k = vk + (int64_t)o * vks;
        const int32_t step = dw * 3;
        int32_t a0 = 1 << (vprec - 1), a1 = a0, a2 = a0;
        const int32_t c = vcnt[o];
        for (int32_t t = 0; t < c; ++t) {
            const int32_t w = k[t];
            a0 += (int32_t)col[t * step] * w;
            a1 += (int32_t)col[t * step + 1] * w;
            a2 += (int32_t)col[t * step + 2] * w;
        }
        out[idx * 3] = (uint8_t)clip8(a0, vprec);
        out[idx * 3 + 1] = (uint8_t)clip8(a1, vprec);
        out[idx * 3 + 2] = (uint8_t)clip8(a2, vprec);
    }
}

size_t resize_fused_lds_bytes(int32_t sw, int32_t sh, int32_t dw) {
    return (size_t)(((sw * sh * 3) + 15) & ~15) + (size_t)sh * dw * 3;
}

void launch_resize(const uint8_t* src, int32_t n, int32_t sw, int32_t sh, int32_t dw, int32_t dh,
                   const ResizeTaps& h, const ResizeTaps& v, uint8_t* tmp, uint8_t* dst,
                   hipStream_t stream, const int32_t* dst_index) {
    if (n <= 0 || dw <= 0 || dh <= 0) return;
    const size_t lds = resize_fused_lds_bytes(sw, sh, dw);
    if (lds <= (size_t)kFusedResizeMaxLds) {
        hipLaunchKernelGGL(resize_fused_kernel, dim3(n), dim3(256), lds, stream, src, sw, sh, dw, dh,
                           h.lo, h.cnt, h.k, h.ksize, h.prec, v.lo, v.cnt, v.k, v.ksize, v.prec,
                           dst_index, dst);
        return;
    }
    const int64_t th = (int64_t)n * sh * dw;
    hipLaunchKernelGGL(resize_h_kernel, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, stream, src,
                       n, sw, sh, dw, h.lo, h.cnt, h.k, h.ksize, h.prec, tmp);
    const int64_t tv = (int64_t)n * dh * dw;
    hipLaunchKernelGGL(resize_v_kernel, dim3((unsigned)((tv + 255) / 256)), dim3(256), 0, stream, tmp,
                       n, sh, dw, dh, v.lo, v.cnt, v.k, v.ksize, v.prec, dst_index, dst);
}

// Integer average RGB of cells (calculate_average_rgb): one 256-thread
// block per cell; sums are exact so the reduction order is free.  Cell
// addressing as in the Lab kernel.
__device__ __forceinline__ void avg_rgb_block(const CellSeg& s, int32_t cell) {
    __shared__ uint64_t part[3][4];
    const int32_t cw = s.cw;
    const uint8_t* base = s.rgb + (int64_t)(cell / s.per_row) * s.block_row_stride +
                          (int64_t)(cell % s.per_row) * s.block_col_stride;
    const int32_t npx = cw * s.ch;
    uint64_t s0 = 0, s1 = 0, s2 = 0;
    for (int32_t p = threadIdx.x; p < npx; p += 256) {
        const int32_t py = p / cw, px = p - py * cw;
        const uint8_t* q = base + (int64_t)py * s.row_stride + (int64_t)px * 3;
        s0 += q[0];
        s1 += q[1];
        s2 += q[2];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wid] = s0;
        part[1][wid] = s1;
        part[2][wid] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint64_t t = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] +
                           part[threadIdx.x][3];
        s.out_rgb[3 * (int64_t)cell + threadIdx.x] = (uint8_t)(t / (uint64_t)npx);
    }
}

__global__ __launch_bounds__(256) void avg_rgb_kernel(CellSeg s) { avg_rgb_block(s, blockIdx.x); }

void launch_avg_rgb(const uint8_t* rgb, int64_t row_stride, int32_t per_row,
                    int64_t block_row_stride, int64_t block_col_stride, int32_t cw, int32_t ch,
                    int32_t n_cells, uint8_t* out, hipStream_t stream) {
    if (n_cells <= 0 || cw <= 0 || ch <= 0) return;
    const CellSeg s{rgb, row_stride, block_row_stride, block_col_stride, per_row, cw, ch, n_cells,
                    nullptr, out};
    hipLaunchKernelGGL(avg_rgb_kernel, dim3(n_cells), dim3(256), 0, stream, s);
}

// Per-cell adjustment: calculate_optimal_adjustment(tile_avg[grid[c]],
// target_avg[c], strength).  Empty cells get a NaN marker in .y.
__device__ __forceinline__ void cell_adjust_one(const int32_t* __restrict__ grid, int32_t c,
                                                const uint8_t* __restrict__ tile_avg,
                                                const uint8_t* __restrict__ target_avg,
                                                float strength, float4* __restrict__ out) {
    const int32_t m = grid[c];
    if (m == kEmpty) {
        out[c] = make_float4(0.0f, __int_as_float(0x7fc00000), 0.0f, 1.0f);
        return;
    }
    const uint8_t* t = tile_avg + 3 * (int64_t)m;
    const uint8_t* g = target_avg + 3 * (int64_t)c;
    out[c] = optimal_adjustment(t[0], t[1], t[2], g[0], g[1], g[2], strength);
}

__global__ void cell_adjust_kernel(const int32_t* __restrict__ grid, int32_t cells,
                                   const uint8_t* __restrict__ tile_avg,
                                   const uint8_t* __restrict__ target_avg, float strength,
                                   float4* __restrict__ out) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    cell_adjust_one(grid, c, tile_avg, target_avg, strength, out);
}

void launch_cell_adjust(const int32_t* grid, int32_t cells, const uint8_t* tile_avg,
                        const uint8_t* target_avg, float strength, float4* out,
                        hipStream_t stream) {
    if (cells <= 0) return;
    hipLaunchKernelGGL(cell_adjust_kernel, dim3((cells + 255) / 256), dim3(256), 0, stream, grid,
                       cells, tile_avg, target_avg, strength, out);
}

// The contrast stage of ColorAdjustment::adjust_pixel at contrast 1 for
// each channel value (contrast1_entry); blocks copy it to LDS.
__device__ const Contrast1Table kContrast1 = make_contrast1_table();

// Pixels per thread: a block of bs threads covers kPx * bs pixels of one
// output row, pixel X = X0 + k * bs (k < kPx), so the row's setup is shared
// by kPx pixels per thread and every k-step stores a contiguous run.  One
// job's launch (a few Mpx) takes 4, so that its blocks fill the device; the
// batched launch, with many jobs' rows, takes 8, which halves the per-thread
// share of the setup (segment search, row and column divisions, the table
// copy): the served stream 57.3k -> 59.0k jobs/s, against 16.9 us instead
// of 14.9 us for a medium job's own launch.  12 and 16 pixels per thread
// lose more to idle lanes at the ends of short rows than they save (57.0k
// and 54.5-55.0k jobs/s).
constexpr int32_t kCompPxJob = 4;
constexpr int32_t kCompPxBatch = 8;

static int32_t composite_row_blocks(int32_t ow, int32_t oh, int32_t threads, int32_t px) {
    return ow > 0 && oh > 0 ? ((ow + threads * px - 1) / (threads * px)) * oh : 0;
}

// Block b of segment s: row Y = b / per_row of the gw*tw x gh*th mosaic,
// pixels X0 + k * bs with X0 = (b % per_row) * bs * kCompPx + thread and
// per_row = ceil(ow / (bs * kCompPx)), ow = gw * tw.  Everything about the
// row is uniform: its cell row's grid entries and adjustments, the tiles'
// pixel row dy = Y % th, the output row.  A thread divides once for its
// first pixel's cell column cx and byte column dx3 = 3 (X0 - cx tw) and steps
// the others by bs = q tw + r.  kNarrow: every material index and the tile
// size th tw 3 are below 2^24 and n th tw 3 < 2^32 (composite_narrow), so a
// tile byte offset is one 24-bit multiply-add; otherwise 64-bit.
template <bool kNarrow, int32_t kCompPx>
__device__ __forceinline__ void composite_block(const CompSeg& s, int32_t b, float* ctab) {
    const int32_t tw = s.tw, th = s.th, ow = s.gw * tw;
    const int32_t bs = (int32_t)blockDim.x;
    const int32_t per_row = (ow + bs * kCompPx - 1) / (bs * kCompPx);
    const int32_t Y = b / per_row;
    const int32_t cy = Y / th, dy = Y - cy * th;
    const int64_t row_cell = (int64_t)cy * s.gw;
    const int32_t* __restrict__ grid = s.grid + row_cell;
    const float4* __restrict__ adj = s.adjust ? s.adj + row_cell : nullptr;
    const uint32_t tile_bytes = (uint32_t)th * (uint32_t)tw * 3u;
    const uint8_t* __restrict__ tiles = s.tiles + (int64_t)dy * tw * 3;
    uint8_t* __restrict__ out = s.out + (int64_t)Y * ow * 3;
    if (s.adjust) {
        // blockDim.x >= 64: one float4 of the table per thread of the first 64
        if (threadIdx.x < 64)
            reinterpret_cast<float4*>(ctab)[threadIdx.x] =
                gload4(reinterpret_cast<const float4*>(kContrast1.v), threadIdx.x);
        __syncthreads();
    }
    const int32_t X0 = (b - Y * per_row) * bs * kCompPx + (int32_t)threadIdx.x;
    // the pixels go through in stages, so that each stage's loads are in
    // flight together: cells, then tile pixels and adjustments, then the
    // arithmetic and stores
    int32_t m[kCompPx], cxk[kCompPx];
    uint32_t dxk[kCompPx];
    {
        // X0 / tw as umulhi(X0, floor((2^32 - 1) / tw) + 1), exact while
        // X0 * tw < 2^32 (then X0 m / 2^32 exceeds X0 / tw by less than 1 / tw)
        const uint32_t tw_magic = 0xffffffffu / (uint32_t)tw + 1u;
        const bool magic_ok =
            (uint64_t)((uint32_t)ow + (uint32_t)(bs * kCompPx)) * (uint32_t)tw < (1ull << 32);
        int32_t cx = magic_ok ? (int32_t)__umulhi((uint32_t)X0, tw_magic)
                              : (int32_t)((uint32_t)X0 / (uint32_t)tw);
        uint32_t dx3 = (uint32_t)(X0 - cx * tw) * 3u;
        const int32_t q = bs / tw;
        const uint32_t r3 = (uint32_t)(bs - q * tw) * 3u, tw3 = (uint32_t)tw * 3u;
#pragma unroll
        for (int32_t k = 0; k < kCompPx; ++k) {
            cxk[k] = cx;
            dxk[k] = dx3;
            m[k] = X0 + k * bs < ow ? gload_at(grid, (uint32_t)cx * 4u) : kEmpty;
            cx += q;
            dx3 += r3;
            if (dx3 >= tw3) {
                dx3 -= tw3;
                ++cx;
            }
        }
    }
    // a pixel's red and green bytes as one (unaligned) 16-bit load, blue
    // alone; the raw loaded values are kept until the last stage, so no
    // wait is needed before the next pixel's loads
    typedef uint16_t u16_any __attribute__((aligned(1)));
    uint32_t rg[kCompPx], bk[kCompPx];
    v4f_t av[kCompPx];
#pragma unroll
    for (int32_t k = 0; k < kCompPx; ++k) {
        rg[k] = 0u;
        bk[k] = 0u;
        av[k] = v4f_t{0.0f, 1.0f, 0.0f, 1.0f};
        if (m[k] != kEmpty) {
            const MH_GAS uint8_t* p =
                kNarrow ? (const MH_GAS uint8_t*)tiles + (__umul24((uint32_t)m[k], tile_bytes) + dxk[k])
                        : (const MH_GAS uint8_t*)tiles + ((int64_t)m[k] * tile_bytes + (int64_t)dxk[k]);
            rg[k] = *(const MH_GAS u16_any*)p;
            bk[k] = p[2];
            if (s.adjust) av[k] = gload_at(reinterpret_cast<const v4f_t*>(adj), (uint32_t)cxk[k] * 16u);
        }
    }
#pragma unroll
    for (int32_t k = 0; k < kCompPx; ++k) {
        const int32_t X = X0 + k * bs;
        if (X < ow) {
            uint32_t r = rg[k] & 0xffu, g = rg[k] >> 8, bl = bk[k];
            // per-cell adjustments come from calculate_optimal_adjustment and
            // ColorAdjustment::new's clamps (cell_adjust_one), whose contrast
            // is always 1
            if (m[k] != kEmpty && s.adjust) {
                const float4 a = make_float4(av[k].x, av[k].y, av[k].z, av[k].w);
                float br, bg, bb;
                if (a.y == 1.0f) {
                    br = clamp_unit(ctab[r] + a.x);
                    bg = clamp_unit(ctab[g] + a.x);
                    bb = clamp_unit(ctab[bl] + a.x);
                } else {
                    br = bright_contrast<true>(u8_unit(r), a.x, a.y);
                    bg = bright_contrast<true>(u8_unit(g), a.x, a.y);
                    bb = bright_contrast<true>(u8_unit(bl), a.x, a.y);
                }
                adjust_tail<true>(a, br, bg, bb, r, g, bl);
            }
            MH_GAS uint8_t* d = (MH_GAS uint8_t*)out + (uint32_t)X * 3u;
            d[0] = (uint8_t)r;
            d[1] = (uint8_t)g;
            d[2] = (uint8_t)bl;
        }
    }
}

template <int32_t kPx>
__device__ __forceinline__ void composite_run(const CompSeg& s, int32_t b) {
    __shared__ __align__(16) float ctab[256];
    if (s.narrow)
        composite_block<true, kPx>(s, b, ctab);
    else
        composite_block<false, kPx>(s, b, ctab);
}

// One job: blocks of 256 threads.
__global__ __launch_bounds__(256) void composite_kernel(CompSeg s) {
    composite_run<kCompPxJob>(s, (int32_t)blockIdx.x);
}

int32_t composite_narrow(int32_t n_tiles, int32_t tw, int32_t th) {
    const uint64_t tile = (uint64_t)tw * (uint64_t)th * 3u;
    return n_tiles >= 0 && (uint64_t)n_tiles < (1ull << 24) && tile < (1ull << 24) &&
                   (uint64_t)n_tiles * tile < (1ull << 32)
               ? 1
               : 0;
}

void launch_composite(const int32_t* grid, int32_t gw, int32_t gh, int32_t tw, int32_t th,
                      const uint8_t* tiles, int32_t n_tiles, const float4* adj, int32_t adjust,
                      uint8_t* out, hipStream_t stream) {
    const int32_t blocks = composite_row_blocks(gw * tw, gh * th, 256, kCompPxJob);
    if (blocks <= 0) return;
    CompSeg s{};
    s.grid = grid;
    s.gw = gw;
    s.gh = gh;
    s.tw = tw;
    s.th = th;
    s.tiles = tiles;
    s.adjust = adjust;
    s.adj = const_cast<float4*>(adj);
    s.out = out;
    s.narrow = composite_narrow(n_tiles, tw, th);
    hipLaunchKernelGGL(composite_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, s);
}

// Batched finish of many jobs: per-cell adjustments (a thread per cell), then
// compositing (a thread per output pixel).  Block b of a launch belongs to
// segment j with starts[j] <= b < starts[j+1]; the host sizes each
// segment's block count from its cells (adjust) or output pixels (composite).
__global__ __launch_bounds__(256) void cell_adjust_batch_kernel(const CompSeg* __restrict__ segs,
                                                                const int32_t* __restrict__ starts,
                                                                int32_t n_segs) {
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    const CompSeg& s = segs[j];
    const int32_t c = (blockIdx.x - starts[j]) * 256 + threadIdx.x;
    if (c >= s.gw * s.gh) return;
    cell_adjust_one(s.grid, c, s.tile_avg, s.target_avg, s.strength, s.adj);
}

// A block of blockDim.x (64, 128 or 256) threads covers kCompPxBatch pixels
// per thread of one output row (composite_block).
__global__ __launch_bounds__(256) void composite_batch_kernel(const CompSeg* __restrict__ segs,
                                                              const int32_t* __restrict__ starts,
                                                              int32_t n_segs) {
    const int32_t j = find_segment(starts, n_segs, blockIdx.x);
    composite_run<kCompPxBatch>(segs[j], (int32_t)blockIdx.x - starts[j]);
}

int32_t cell_adjust_blocks(int32_t cells) { return cells > 0 ? (cells + 255) / 256 : 0; }

int32_t composite_blocks(int32_t ow, int32_t oh, int32_t threads) {
    return composite_row_blocks(ow, oh, threads, kCompPxBatch);
}

void launch_cell_adjust_batch(const CompSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                              int32_t total_blocks, hipStream_t stream) {
    if (n_segs <= 0 || total_blocks <= 0) return;
    hipLaunchKernelGGL(cell_adjust_batch_kernel, dim3(total_blocks), dim3(256), 0, stream, d_segs,
                       d_starts, n_segs);
}

void launch_composite_batch(const CompSeg* d_segs, const int32_t* d_starts, int32_t n_segs,
                            int32_t total_blocks, int32_t threads, hipStream_t stream) {
    if (n_segs <= 0 || total_blocks <= 0) return;
    hipLaunchKernelGGL(composite_batch_kernel, dim3(total_blocks), dim3(threads), 0, stream, d_segs,
                       d_starts, n_segs);
}

// ColorAdjustment::apply_to_image over a flat pixel array.
__global__ void adjust_pixels_kernel(float4 adj, const uint8_t* __restrict__ in, int64_t npx,
                                     uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    uint32_t r, g, b;
    adjust_pixel(adj, in[3 * i], in[3 * i + 1], in[3 * i + 2], r, g, b);
    out[3 * i] = (uint8_t)r;
    out[3 * i + 1] = (uint8_t)g;
    out[3 * i + 2] = (uint8_t)b;
}

void launch_adjust_pixels(float4 adj, const uint8_t* in, int64_t npx, uint8_t* out,
                          hipStream_t stream) {
    if (npx <= 0) return;
    hipLaunchKernelGGL(adjust_pixels_kernel, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0,
                       stream, adj, in, npx, out);
}

__global__ void optimal_adjust_kernel(const uint8_t* __restrict__ tile,
                                      const uint8_t* __restrict__ target, int32_t n, float k,
                                      float4* __restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t* t = tile + 3 * (int64_t)i;
    const uint8_t* g = target + 3 * (int64_t)i;
    out[i] = optimal_adjustment(t[0], t[1], t[2], g[0], g[1], g[2], k);
}

void launch_optimal_adjust(const uint8_t* tile, const uint8_t* target, int32_t n, float k,
                           float4* out, hipStream_t stream) {
    if (n <= 0) return;
    hipLaunchKernelGGL(optimal_adjust_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, tile,
                       target, n, k, out);
}

hipError_t resize_configure(int max_lds) {
    return raise_dynamic_lds(reinterpret_cast<const void*>(&resize_fused_kernel), max_lds);
}

}  // namespace mh
