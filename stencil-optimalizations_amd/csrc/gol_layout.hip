// Conversions between the engine's row layout and the reference's other bit
// layouts (src/algorithms/_shared/bitwise/bit_modes.hpp), used at the C ABI
// boundary when gol_config.layout is not GOL_LAYOUT_ROWS:
//   columns (BitColumns, :10-32): W-bit word (x, wy) at wy * X + x holds cells
//     (x, W*wy + b) in bit b, W = 16 / 32 / 64;
//   tiles (BitTile, :37-125): XB x YB cells per word (4x4 in 16 bits, 8x4 in
//     32, 8x8 in 64), word (wx, wy) at wy * (X / XB) + wx, cell (tx, ty) in bit
//     W - 1 - (ty * XB + tx) (BitTileCommon, :37-56).
// Rows: X/32 dwords per row, bit k of dword j = cell 32j + k.
//
// Columns <-> rows is a bit-matrix transpose: a wavefront's 64 lanes hold 64
// consecutive column words (or 64 row segments), and one ballot per bit
// gathers a 64-cell row segment (or a column word). Tiles <-> rows gathers one
// XB-bit field per tile row, bit-reversed (the tile's first cell is its most
// significant bit).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gol_kernels.hpp"

namespace golhip {
namespace {

constexpr int kLayoutBlock = 256;

template <typename Word, int W>
__global__ __launch_bounds__(kLayoutBlock) void columns_to_rows(const Word* __restrict__ cols,
                                                                uint32_t* __restrict__ rows, int64_t x_cells,
                                                                int64_t word_rows) {
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kLayoutBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t groups = (x_cells + 63) / 64;
    if (wave >= groups * word_rows) return;  // whole wavefronts leave
    const int64_t g = wave % groups, wy = wave / groups;
    const int64_t x = g * 64 + lane;
    const uint64_t word = x < x_cells ? static_cast<uint64_t>(cols[wy * x_cells + x]) : 0;
    uint64_t mine = 0;
#pragma unroll
    for (int b = 0; b < W; ++b) {
        const uint64_t m = __ballot((word >> b) & 1);
        if (lane == b) mine = m;
    }
    if (lane < W) {
        uint32_t* r = rows + (wy * W + lane) * (x_cells / 32) + g * 2;
        r[0] = static_cast<uint32_t>(mine);
        if (g * 64 + 32 < x_cells) r[1] = static_cast<uint32_t>(mine >> 32);
    }
}

template <typename Word, int W>
__global__ __launch_bounds__(kLayoutBlock) void rows_to_columns(const uint32_t* __restrict__ rows,
                                                                Word* __restrict__ cols, int64_t x_cells,
                                                                int64_t word_rows) {
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kLayoutBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t groups = (x_cells + 63) / 64;
    if (wave >= groups * word_rows) return;
    const int64_t g = wave % groups, wy = wave / groups;
    uint64_t v = 0;
    if (lane < W) {
        const uint32_t* r = rows + (wy * W + lane) * (x_cells / 32) + g * 2;
        v = r[0];
        if (g * 64 + 32 < x_cells) v |= static_cast<uint64_t>(r[1]) << 32;
    }
    uint64_t mine = 0;
#pragma unroll
    for (int c = 0; c < 64; ++c) {
        const uint64_t m = __ballot((v >> c) & 1);
        if (lane == c) mine = m;
    }
    const int64_t x = g * 64 + lane;
    if (x < x_cells) cols[wy * x_cells + x] = static_cast<Word>(mine);  // lanes >= W gave 0
}

// The XB-bit field of tile row ty, cells in increasing x.
template <int XB, int YB>
__device__ __forceinline__ uint32_t tile_row(uint64_t t, int ty) {
    const uint32_t field = static_cast<uint32_t>(t >> (XB * (YB - 1 - ty))) & ((1u << XB) - 1);
    return __builtin_bitreverse32(field) >> (32 - XB);
}

template <typename Word, int XB, int YB>
__global__ __launch_bounds__(kLayoutBlock) void tiles_to_rows(const Word* __restrict__ tiles,
                                                              uint32_t* __restrict__ rows, int64_t x_cells,
                                                              int64_t y_rows) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kLayoutBlock + threadIdx.x;
    const int64_t x_dw = x_cells / 32;
    if (i >= x_dw * y_rows) return;
    const int64_t y = i / x_dw, dx = i % x_dw;
    const int64_t tiles_x = x_cells / XB;
    constexpr int kPer = 32 / XB;  // tiles across one dword
    const Word* t = tiles + (y / YB) * tiles_x + dx * kPer;
    const int ty = static_cast<int>(y % YB);
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) out |= tile_row<XB, YB>(static_cast<uint64_t>(t[k]), ty) << (XB * k);
    rows[i] = out;
}

template <typename Word, int XB, int YB>
__global__ __launch_bounds__(kLayoutBlock) void rows_to_tiles(const uint32_t* __restrict__ rows,
                                                              Word* __restrict__ tiles, int64_t x_cells,
                                                              int64_t y_rows) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kLayoutBlock + threadIdx.x;
    const int64_t tiles_x = x_cells / XB;
    if (i >= tiles_x * (y_rows / YB)) return;
    const int64_t wy = i / tiles_x, wx = i % tiles_x;
    const int64_t x_dw = x_cells / 32;
    const int64_t x0 = wx * XB;
    const uint32_t* r = rows + wy * YB * x_dw + x0 / 32;
    const int off = static_cast<int>(x0 % 32);
    uint64_t t = 0;
#pragma unroll
    for (int ty = 0; ty < YB; ++ty) {
        const uint32_t field = (r[ty * x_dw] >> off) & ((1u << XB) - 1);
        const uint32_t rev = __builtin_bitreverse32(field) >> (32 - XB);
        t |= static_cast<uint64_t>(rev) << (XB * (YB - 1 - ty));
    }
    tiles[i] = static_cast<Word>(t);
}

unsigned blocks_for(int64_t threads) { return static_cast<unsigned>((threads + kLayoutBlock - 1) / kLayoutBlock); }

}  // namespace

hipError_t layout_to_rows(int layout, int word_bits, const void* src, uint32_t* rows, int64_t x_cells,
                          int64_t y_rows, hipStream_t stream) {
    const dim3 block(kLayoutBlock);
    if (layout == kLayoutColumns) {
        const int64_t waves = (x_cells + 63) / 64 * (y_rows / word_bits);
        const dim3 grid(blocks_for(waves * 64));
        if (word_bits == 16)
            hipLaunchKernelGGL((columns_to_rows<uint16_t, 16>), grid, block, 0, stream,
                               static_cast<const uint16_t*>(src), rows, x_cells, y_rows / 16);
        else if (word_bits == 32)
            hipLaunchKernelGGL((columns_to_rows<uint32_t, 32>), grid, block, 0, stream,
                               static_cast<const uint32_t*>(src), rows, x_cells, y_rows / 32);
        else
            hipLaunchKernelGGL((columns_to_rows<uint64_t, 64>), grid, block, 0, stream,
                               static_cast<const uint64_t*>(src), rows, x_cells, y_rows / 64);
        return hipGetLastError();
    }
    if (layout == kLayoutTiles) {
        const dim3 grid(blocks_for(x_cells / 32 * y_rows));
        if (word_bits == 16)
            hipLaunchKernelGGL((tiles_to_rows<uint16_t, 4, 4>), grid, block, 0, stream,
                               static_cast<const uint16_t*>(src), rows, x_cells, y_rows);
        else if (word_bits == 32)
            hipLaunchKernelGGL((tiles_to_rows<uint32_t, 8, 4>), grid, block, 0, stream,
                               static_cast<const uint32_t*>(src), rows, x_cells, y_rows);
        else
            hipLaunchKernelGGL((tiles_to_rows<uint64_t, 8, 8>), grid, block, 0, stream,
                               static_cast<const uint64_t*>(src), rows, x_cells, y_rows);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

hipError_t rows_to_layout(int layout, int word_bits, const uint32_t* rows, void* dst, int64_t x_cells,
                          int64_t y_rows, hipStream_t stream) {
    const dim3 block(kLayoutBlock);
    if (layout == kLayoutColumns) {
        const int64_t waves = (x_cells + 63) / 64 * (y_rows / word_bits);
        const dim3 grid(blocks_for(waves * 64));
        if (word_bits == 16)
            hipLaunchKernelGGL((rows_to_columns<uint16_t, 16>), grid, block, 0, stream, rows,
                               static_cast<uint16_t*>(dst), x_cells, y_rows / 16);
        else if (word_bits == 32)
            hipLaunchKernelGGL((rows_to_columns<uint32_t, 32>), grid, block, 0, stream, rows,
                               static_cast<uint32_t*>(dst), x_cells, y_rows / 32);
        else
            hipLaunchKernelGGL((rows_to_columns<uint64_t, 64>), grid, block, 0, stream, rows,
                               static_cast<uint64_t*>(dst), x_cells, y_rows / 64);
        return hipGetLastError();
    }
    if (layout == kLayoutTiles) {
        if (word_bits == 16)
            hipLaunchKernelGGL((rows_to_tiles<uint16_t, 4, 4>), dim3(blocks_for(x_cells / 4 * (y_rows / 4))), block,
                               0, stream, rows, static_cast<uint16_t*>(dst), x_cells, y_rows);
        else if (word_bits == 32)
            hipLaunchKernelGGL((rows_to_tiles<uint32_t, 8, 4>), dim3(blocks_for(x_cells / 8 * (y_rows / 4))), block,
                               0, stream, rows, static_cast<uint32_t*>(dst), x_cells, y_rows);
        else
            hipLaunchKernelGGL((rows_to_tiles<uint64_t, 8, 8>), dim3(blocks_for(x_cells / 8 * (y_rows / 8))), block,
                               0, stream, rows, static_cast<uint64_t*>(dst), x_cells, y_rows);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

}  // namespace golhip
