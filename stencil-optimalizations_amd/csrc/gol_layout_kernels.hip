// Kernels that advance the reference's 8x8-tile and column bit layouts in
// their own words, with no conversion to rows (GOL_FLAG_NATIVE_LAYOUT).
//
// Reference counterparts:
//   tile8_kernel   the 8x8-tile full adder, _shared/bitwise/bitwise-ops/tiled-full-adder.cuh:19-172
//                  (and the BitTile policy's word ops), in the plain launch of
//                  cuda-naive-bitwise/cuda_naive_bitwise_kernel.cu:32-79 ("tiles-64") and the
//                  temporal form of cuda-temporal-simple-tiled/temporal_simple_tiled.cu:41-132
//   col_kernel     the BitColumns policy's ops (bit_modes.hpp:10-32) in the plain launch
//                  ("cols-32/64")
//
// 8x8 tiles. Cell (tx, ty) of a tile word is bit 63 - (8 ty + tx): as
// little-endian dwords, the high dword holds rows 0..3 and the low dword rows
// 4..7, one row per byte, its first cell in the byte's top bit. Read as the
// sequence hi(t), lo(t), hi(t + 1), ... a column of tiles is a column of
// 8-cell rows, four per dword, the top row in the top byte. A lane holds one
// such column (one tile column, BT tiles tall); per dword and generation:
//   west/east planes  byte-masked shifts, the edge bit of each row from the
//                     tile to the left / right (the neighbouring lane, DPP)
//   row sums          H = w + c + e and E = w + e as bit planes (4 ops)
//   rows above/below  H of the row above / below each row: a funnel shift by
//                     8 / 24 bits with the dword above / below (v_alignbit)
//   next state        the row kernel's six v_bitop3 (n | alive) == 3
// 20 VALU ops per dword (32 cells) and generation, against 12 for rows.
//
// Columns. A W-bit column word (x, wy) holds cells (x, W wy + k) in bit k, so
// a grid column is a run of dwords, cell (x, 32 v + k) in bit k of the
// column's dword v (a 64-bit word is the dword pair (2 wy, 2 wy + 1)). The
// row scheme transposed: a lane holds one grid column; cells above and below
// come from funnel shifts with the lane's neighbouring dwords, the column
// sums of the columns left and right from the neighbouring lanes (DPP). 16
// VALU ops per dword and generation.
//
// Both kernels run T generations per launch on a tile with a halo (tiles of
// 8 cells around 8x8 tiles, lanes and dwords around columns) and clear cells
// outside the grid after every generation (the dead boundary).

#include "gol_device.hpp"
#include "gol_kernels.hpp"

namespace golhip {

namespace {

// West / east neighbour planes of a tile dword (4 rows of 8 cells).
__device__ __forceinline__ uint32_t tile_west(uint32_t left, uint32_t q) {
    return ((q >> 1) & 0x7F7F7F7Fu) | ((left << 7) & 0x80808080u);
}
__device__ __forceinline__ uint32_t tile_east(uint32_t q, uint32_t right) {
    return ((q << 1) & 0xFEFEFEFEu) | ((right >> 7) & 0x01010101u);
}

// Keep-mask of dword j from a lane's in-grid bits (bit j of inside[j / 32]):
// all ones or zero, by a signed one-bit field extract.
template <int N>
__device__ __forceinline__ uint32_t keep_bits(const uint32_t (&inside)[(N + 31) / 32], int j) {
    return static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(inside[j / 32]), j % 32, 1));
}

// In-grid bits of a lane's dwords [j_lo, j_hi), none for a lane outside.
template <int N>
__device__ __forceinline__ void inside_bits(uint32_t (&inside)[(N + 31) / 32], bool lane_in, int j_lo, int j_hi) {
#pragma unroll
    for (int k = 0; k < (N + 31) / 32; ++k) {
        uint32_t b = 0;
        for (int j = 32 * k; j < 32 * k + 32 && j < N; ++j)
            if (j >= j_lo && j < j_hi) b |= 1u << (j - 32 * k);
        inside[k] = lane_in ? b : 0u;
    }
}

__device__ __forceinline__ RowSum tile_sums(uint32_t q) {
    return row_sum(tile_west(from_lower_lane(q), q), q, tile_east(q, from_upper_lane(q)));
}

// One generation of a lane's column of N tile dwords s (top to bottom), `up`
// the dword just above s[0] and `down` the one just below s[N - 1]. Lanes 0
// and 63 see dead cells beyond the wavefront (covered by the tile's halo).
// MASK: cells outside the grid are cleared, dword j where bit j of `inside`
// is 0. Each dword is overwritten only after every read of its old value, its
// neighbours' DPP reads included (they happen in the same instruction in
// every lane).
template <int N, bool MASK>
__device__ __forceinline__ void tile8_step(uint32_t (&s)[N], uint32_t up, uint32_t down,
                                           const uint32_t (&inside)[(N + 31) / 32]) {
    uint32_t p0, p1;  // H of the dword above the current one
    row_sum_h(tile_west(from_lower_lane(up), up), up, tile_east(up, from_upper_lane(up)), p0, p1);
    RowSum cur = tile_sums(s[0]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        RowSum nxt;
        if (j + 1 < N) {
            nxt = tile_sums(s[j + 1]);
        } else {
            row_sum_h(tile_west(from_lower_lane(down), down), down, tile_east(down, from_upper_lane(down)), nxt.h0,
                      nxt.h1);
            nxt.e0 = nxt.e1 = 0;
        }
        const uint32_t a0 = __builtin_amdgcn_alignbit(p0, cur.h0, 8), a1 = __builtin_amdgcn_alignbit(p1, cur.h1, 8);
        const uint32_t b0 = __builtin_amdgcn_alignbit(cur.h0, nxt.h0, 24);
        const uint32_t b1 = __builtin_amdgcn_alignbit(cur.h1, nxt.h1, 24);
        uint32_t v = next_state(a0, a1, cur.e0, cur.e1, b0, b1, s[j]);
        if (MASK) v &= keep_bits<N>(inside, j);
        s[j] = v;
        p0 = cur.h0;
        p1 = cur.h1;
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// A workgroup of NW wavefronts advances a tile of 64 tile columns x NW * BT
// tile rows by `gens` generations and stores its interior. Wavefronts are
// stacked vertically and swap their edge dwords through LDS every generation
// (double-buffered by parity: one barrier per generation); the zero slots
// above the first and below the last wavefront are the tile's edges.
template <int BT, int NW, bool MASK>
__device__ __forceinline__ void tile8_generations(uint32_t (&s)[2 * BT], uint32_t (&inside)[(2 * BT + 31) / 32],
                                                  uint32_t (&tops)[2][NW + 1][64], uint32_t (&bottoms)[2][NW + 1][64],
                                                  int gens, int lane, int w) {
    for (int g = 0; g < gens; ++g) {
        // opaque each generation, so the per-dword masks are not hoisted out
        // of the loop (they would hold a register each)
        if (MASK) asm volatile("" : "+v"(inside[0]));
        const int par = g & 1;
        tops[par][w][lane] = s[0];
        bottoms[par][w + 1][lane] = s[2 * BT - 1];
        __syncthreads();
        const uint32_t up = bottoms[par][w][lane];
        const uint32_t down = tops[par][w + 1][lane];
        tile8_step<2 * BT, MASK>(s, up, down, inside);
    }
}

template <int BT, int NW>
__global__ __launch_bounds__(64 * NW) void tile8_kernel(const uint2* __restrict__ src, uint2* __restrict__ dst,
                                                        Tile8Geo g, int gens) {
    __shared__ uint32_t tops[2][NW + 1][64];
    __shared__ uint32_t bottoms[2][NW + 1][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int gx = static_cast<int>(blockIdx.x) % g.groups_x, gy = static_cast<int>(blockIdx.x) / g.groups_x;
    const int x0 = gx * g.stride_x - g.halo_x;  // tile column of lane 0
    const int y0 = gy * g.stride_y - g.halo_y;  // tile row of wave 0's first tile
    const int tx = x0 + lane;
    const int ty0 = y0 + w * BT;
    const bool col_in = tx >= 0 && tx < g.x_tiles;
    const int ctx = tx < 0 ? 0 : (tx >= g.x_tiles ? g.x_tiles - 1 : tx);

    uint32_t s[2 * BT];
#pragma unroll
    for (int t = 0; t < BT; ++t) {
        const int ty = ty0 + t;
        const bool in = col_in && ty >= 0 && ty < g.y_tiles;
        const int cty = ty < 0 ? 0 : (ty >= g.y_tiles ? g.y_tiles - 1 : ty);
        const uint2 v = src[static_cast<int64_t>(cty) * g.x_tiles + ctx];
        s[2 * t] = in ? v.y : 0u;      // rows 0..3
        s[2 * t + 1] = in ? v.x : 0u;  // rows 4..7
    }
    // the band's dwords inside the grid: tiles [t_lo, t_hi) (wave-uniform)
    const int t_lo = ty0 < 0 ? -ty0 : 0;
    const int t_hi = g.y_tiles - ty0 < BT ? g.y_tiles - ty0 : BT;
    uint32_t inside[(2 * BT + 31) / 32];
    inside_bits<2 * BT>(inside, col_in, 2 * t_lo, 2 * t_hi);
    if (w == 0) bottoms[0][0][lane] = bottoms[1][0][lane] = 0u;
    if (w == NW - 1) tops[0][NW][lane] = tops[1][NW][lane] = 0u;

    const bool edge = x0 < 0 || x0 + 64 > g.x_tiles || y0 < 0 || y0 + NW * BT > g.y_tiles;  // workgroup-uniform
    if (edge)
        tile8_generations<BT, NW, true>(s, inside, tops, bottoms, gens, lane, w);
    else
        tile8_generations<BT, NW, false>(s, inside, tops, bottoms, gens, lane, w);

    const bool lane_inside = lane >= g.halo_x && lane < 64 - g.halo_x && col_in;
#pragma unroll
    for (int t = 0; t < BT; ++t) {
        const int tr = w * BT + t;
        const int ty = ty0 + t;
        if (lane_inside && tr >= g.halo_y && tr < NW * BT - g.halo_y && ty >= 0 && ty < g.y_tiles)
            dst[static_cast<int64_t>(ty) * g.x_tiles + tx] = make_uint2(s[2 * t + 1], s[2 * t]);
    }
}

// ---- columns ----

// One generation of a lane's column run of N dwords (top to bottom); the
// dwords beyond the run count as dead (covered by the halo dwords). MASK: as
// tile8_step.
template <int N, bool MASK>
__device__ __forceinline__ void col_step(uint32_t (&s)[N], const uint32_t (&inside)[(N + 31) / 32]) {
    uint32_t prev = 0;  // old value of the dword above the current one
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t c = s[j];
        const uint32_t up = __builtin_amdgcn_alignbit(c, prev, 31);                    // cell (x, y - 1)
        const uint32_t dn = __builtin_amdgcn_alignbit(j + 1 < N ? s[j + 1] : 0u, c, 1);  // cell (x, y + 1)
        const RowSum r = row_sum(up, c, dn);  // column sums: up + c + dn and up + dn
        const uint32_t a0 = from_lower_lane(r.h0), a1 = from_lower_lane(r.h1);  // column x - 1
        const uint32_t b0 = from_upper_lane(r.h0), b1 = from_upper_lane(r.h1);  // column x + 1
        uint32_t v = next_state(a0, a1, r.e0, r.e1, b0, b1, c);
        if (MASK) v &= keep_bits<N>(inside, j);
        prev = c;
        s[j] = v;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Dword v of column x: 64-bit words (x, wy) at wy * X + x hold dwords 2 wy and
// 2 wy + 1; 32-bit words (x, v) at v * X + x.
template <int WB>
__device__ __forceinline__ int64_t col_index(int x, int v, int x_cells) {
    if constexpr (WB == 64)
        return (static_cast<int64_t>(v >> 1) * x_cells + x) * 2 + (v & 1);
    else
        return static_cast<int64_t>(v) * x_cells + x;
}

// Each wavefront advances its own tile (no LDS, no barrier): 64 columns x N
// dwords, of which halo_x lanes on each side and halo_v dwords above and
// below are halo; kColWaves tiles per workgroup.
template <int N, bool MASK>
__device__ __forceinline__ void col_generations(uint32_t (&s)[N], uint32_t (&inside)[(N + 31) / 32], int gens) {
    for (int g = 0; g < gens; ++g) {
        if (MASK) {
#pragma unroll
            for (int k = 0; k < (N + 31) / 32; ++k) asm volatile("" : "+v"(inside[k]));
        }
        col_step<N, MASK>(s, inside);
    }
}

template <int N, int WB>
__global__ __launch_bounds__(64 * kColWaves) void col_kernel(const uint32_t* __restrict__ src,
                                                             uint32_t* __restrict__ dst, ColGeo g, int gens) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = static_cast<int>(blockIdx.x) * kColWaves + w;
    if (tile >= g.groups_x * g.groups_v) return;  // the whole wavefront leaves together
    const int gx = tile % g.groups_x, gv = tile / g.groups_x;
    const int x0 = gx * g.stride_x - g.halo_x;
    const int v0 = gv * g.stride_v - g.halo_v;
    const int x = x0 + lane;
    const bool col_in = x >= 0 && x < g.x_cells;
    const int cx = x < 0 ? 0 : (x >= g.x_cells ? g.x_cells - 1 : x);

    // the run's dwords inside the grid: [j_lo, j_hi) (wave-uniform); the
    // loads and stores test these bits (a VGPR), not a scalar condition per
    // dword, which would hold two SGPRs each for the whole kernel
    const int j_lo = v0 < 0 ? -v0 : 0;
    const int j_hi = g.v_dwords - v0 < N ? g.v_dwords - v0 : N;
    uint32_t inside[(N + 31) / 32];
    inside_bits<N>(inside, col_in, j_lo, j_hi);
    uint32_t s[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int v = v0 + j;
        const int cv = v < 0 ? 0 : (v >= g.v_dwords ? g.v_dwords - 1 : v);
        s[j] = src[col_index<WB>(cx, cv, g.x_cells)] & keep_bits<N>(inside, j);
    }
    const bool edge = x0 < 0 || x0 + 64 > g.x_cells || v0 < 0 || v0 + N > g.v_dwords;  // wave-uniform
    if (edge)
        col_generations<N, true>(s, inside, gens);
    else
        col_generations<N, false>(s, inside, gens);

    if (!(lane >= g.halo_x && lane < 64 - g.halo_x && col_in)) return;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j >= g.halo_v && j < N - g.halo_v && keep_bits<N>(inside, j))
            dst[col_index<WB>(x, v0 + j, g.x_cells)] = s[j];
    }
}

}  // namespace

// ---- host side ----

bool plan_tile8(int64_t x_tiles, int64_t y_tiles, int gens, Tile8Geo* g, const char** why) {
    if (gens < 1 || gens > kTile8MaxSteps) {
        *why = "8x8-tile launches take 1 to 16 generations";
        return false;
    }
    const int h = (gens + 7) / 8;  // tiles of halo: 8 cells each
    constexpr int tile_rows = kTile8Rows * kTile8Waves;
    g->x_tiles = static_cast<int>(x_tiles);
    g->y_tiles = static_cast<int>(y_tiles);
    g->halo_x = x_tiles <= 64 ? 0 : h;
    g->halo_y = y_tiles <= tile_rows ? 0 : h;
    g->stride_x = 64 - 2 * g->halo_x;
    g->stride_y = tile_rows - 2 * g->halo_y;
    g->groups_x = static_cast<int>((x_tiles + g->stride_x - 1) / g->stride_x);
    g->groups_y = static_cast<int>((y_tiles + g->stride_y - 1) / g->stride_y);
    return true;
}

hipError_t launch_tile8(const Tile8Geo& g, const uint32_t* src, uint32_t* dst, int gens, hipStream_t stream) {
    const unsigned groups = static_cast<unsigned>(g.groups_x) * static_cast<unsigned>(g.groups_y);
    hipLaunchKernelGGL((tile8_kernel<kTile8Rows, kTile8Waves>), dim3(groups), dim3(64 * kTile8Waves), 0, stream,
                       reinterpret_cast<const uint2*>(src), reinterpret_cast<uint2*>(dst), g, gens);
    return hipGetLastError();
}

bool plan_columns(int64_t x_cells, int64_t y_cells, int word_bits, int gens, ColGeo* g, const char** why) {
    if (word_bits != 32 && word_bits != 64) {
        *why = "the column kernel takes 32- and 64-bit column words";
        return false;
    }
    if (gens < 1 || gens > kColMaxSteps) {
        *why = "column launches take 1 to 16 generations";
        return false;
    }
    if (y_cells % word_bits != 0) {
        *why = "Y must be a multiple of the column word";
        return false;
    }
    g->x_cells = static_cast<int>(x_cells);
    g->v_dwords = static_cast<int>(y_cells / 32);
    g->word_bits = word_bits;
    g->halo_x = x_cells <= 64 ? 0 : gens;  // a column of cells per generation
    g->halo_v = g->v_dwords <= kColDwords ? 0 : 1;  // 32 rows >= generations per launch
    g->stride_x = 64 - 2 * g->halo_x;
    g->stride_v = kColDwords - 2 * g->halo_v;
    g->groups_x = static_cast<int>((x_cells + g->stride_x - 1) / g->stride_x);
    g->groups_v = (g->v_dwords + g->stride_v - 1) / g->stride_v;
    return true;
}

hipError_t launch_columns(const ColGeo& g, const uint32_t* src, uint32_t* dst, int gens, hipStream_t stream) {
    const unsigned tiles = static_cast<unsigned>(g.groups_x) * static_cast<unsigned>(g.groups_v);
    const dim3 grid((tiles + kColWaves - 1) / kColWaves);
    if (g.word_bits == 64)
        hipLaunchKernelGGL((col_kernel<kColDwords, 64>), grid, dim3(64 * kColWaves), 0, stream, src, dst, g, gens);
    else
        hipLaunchKernelGGL((col_kernel<kColDwords, 32>), grid, dim3(64 * kColWaves), 0, stream, src, dst, g, gens);
    return hipGetLastError();
}

}  // namespace golhip
