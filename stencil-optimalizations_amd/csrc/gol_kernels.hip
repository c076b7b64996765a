// HIP kernels for gfx950: the plain one-generation stencil and the
// temporal-tile stencil (T generations per launch).
//
// Reference counterparts:
//   plain kernel     algorithms/cuda-naive-bitwise/cuda_naive_bitwise_kernel.cu:32-79 and
//                    algorithms/cuda-simple-warp-exchange/simple_warp_exchange.cu:33-85
//                    (one launch per generation, input/output ping-pong)
//   tile kernel      algorithms/cuda-temporal-rowed/temporal_rowed.cu:33-229
//                    (T generations per launch on a tile with a T-row halo)
//
// MI355X mapping. A wavefront holds a band of BR rows x 64*DW dwords in VGPRs
// (lane l owns dwords [l*DW, (l+1)*DW) of each row). Word edges between lanes
// travel by DPP wave shifts; the rows just above and below a band come from the
// neighbouring wavefronts through a double-buffered LDS slot, so a generation
// costs one workgroup barrier. A tile is NW bands stacked vertically. Tiles
// overlap by their halo: halo_rows rows (>= generations per launch) above and
// below, and one lane (32*DW cells >= generations per launch) left and right,
// and neither halo is needed along an axis where one tile spans the grid.
// Cells outside the grid are cleared after every generation, which is the
// reference's dead boundary (algorithms/cpu-naive/gol_cpu_naive.hpp:92-117).

#include "gol_device.hpp"
#include "gol_kernels.hpp"

namespace golhip {

namespace {

// Loads are branch-free: the address is clamped into the grid and the value
// selected afterwards, so the register band never flows through a branch.
// With vec_ok (x_dw and gx0 multiples of DW) a lane's segment lies wholly
// inside or wholly outside the row.
// COH: device-coherent accesses (agent-scope relaxed atomics, sc1), for cells
// that other workgroups, possibly on other XCDs, write during the same launch:
// such a store is acknowledged once every CU can see it and such a load reads
// at that level, so no cache-wide write-back or invalidate is needed.
template <int DW, bool COH = false>
__device__ __forceinline__ void load_seg(const uint32_t* __restrict__ src, int x_dw, int y_rows, int gy, int gx0,
                                         bool vec_ok, uint32_t (&out)[DW]) {
    const bool row_in = gy >= 0 && gy < y_rows;
    const int cy = gy < 0 ? 0 : (gy >= y_rows ? y_rows - 1 : gy);
    const uint32_t* row = src + static_cast<int64_t>(cy) * x_dw;
    if constexpr (COH) {
#pragma unroll
        for (int d = 0; d < DW; ++d) {
            const int gx = gx0 + d;
            const bool in = row_in && gx >= 0 && gx < x_dw;
            const uint32_t v = __hip_atomic_load(const_cast<uint32_t*>(row) + (gx < 0 ? 0 : (gx >= x_dw ? x_dw - 1 : gx)),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[d] = in ? v : 0u;
        }
    } else if (vec_ok) {
        const bool in = row_in && gx0 >= 0 && gx0 + DW <= x_dw;
        const int cx = gx0 < 0 ? 0 : (gx0 + DW > x_dw ? x_dw - DW : gx0);
        if constexpr (DW == 1) {
            const uint32_t v = row[cx];
            out[0] = in ? v : 0u;
        } else if constexpr (DW == 2) {
            const uint2 v = *reinterpret_cast<const uint2*>(row + cx);
            out[0] = in ? v.x : 0u;
            out[1] = in ? v.y : 0u;
        } else {
#pragma unroll
            for (int k = 0; k < DW / 4; ++k) {
                const uint4 v = reinterpret_cast<const uint4*>(row + cx)[k];
                out[4 * k + 0] = in ? v.x : 0u;
                out[4 * k + 1] = in ? v.y : 0u;
                out[4 * k + 2] = in ? v.z : 0u;
                out[4 * k + 3] = in ? v.w : 0u;
            }
        }
    } else {
#pragma unroll
        for (int d = 0; d < DW; ++d) {
            const int gx = gx0 + d;
            const bool in = row_in && gx >= 0 && gx < x_dw;
            const uint32_t v = row[gx < 0 ? 0 : (gx >= x_dw ? x_dw - 1 : gx)];
            out[d] = in ? v : 0u;
        }
    }
}

template <int DW, bool COH = false>
__device__ __forceinline__ void store_seg(uint32_t* __restrict__ dst, int x_dw, int gy, int gx0, bool vec_ok,
                                          const uint32_t (&v)[DW]) {
    uint32_t* row = dst + static_cast<int64_t>(gy) * x_dw;
    if constexpr (COH) {
#pragma unroll
        for (int d = 0; d < DW; ++d) {
            const int gx = gx0 + d;
            if (gx >= 0 && gx < x_dw) __hip_atomic_store(row + gx, v[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (vec_ok && gx0 >= 0 && gx0 + DW <= x_dw) {
        if constexpr (DW == 1) {
            row[gx0] = v[0];
        } else if constexpr (DW == 2) {
            *reinterpret_cast<uint2*>(row + gx0) = make_uint2(v[0], v[1]);
        } else {
#pragma unroll
            for (int k = 0; k < DW / 4; ++k) {
                reinterpret_cast<uint4*>(row + gx0)[k] =
                    make_uint4(v[4 * k + 0], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
            }
        }
    } else {
#pragma unroll
        for (int d = 0; d < DW; ++d) {
            const int gx = gx0 + d;
            if (gx >= 0 && gx < x_dw) row[gx] = v[d];
        }
    }
}

// ---------------------------------------------------------------------------
// Plain kernel: each lane owns PLAIN_RB rows x DW dwords and reads the rows
// just above and below its strip from global memory; lanes 0 and 63 also load
// the dword beyond the wavefront's edge (the DPP shift has no source there).
// ---------------------------------------------------------------------------
constexpr int PLAIN_RB = 8;
constexpr int PLAIN_WAVES = 4;

template <int DW>
__global__ __launch_bounds__(64 * PLAIN_WAVES) void plain_kernel(const uint32_t* __restrict__ src,
                                                                 uint32_t* __restrict__ dst, int x_dw, int y_rows) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int wave_x0 = blockIdx.x * (64 * DW);
    const int gx0 = wave_x0 + lane * DW;
    const int gy0 = (blockIdx.y * PLAIN_WAVES + w) * PLAIN_RB;
    if (gy0 >= y_rows) return;  // whole wavefront leaves together
    const bool vec_ok = (x_dw % DW) == 0;

    uint32_t s[PLAIN_RB][DW], above[DW], below[DW];
    load_seg<DW>(src, x_dw, y_rows, gy0 - 1, gx0, vec_ok, above);
#pragma unroll
    for (int r = 0; r < PLAIN_RB; ++r) load_seg<DW>(src, x_dw, y_rows, gy0 + r, gx0, vec_ok, s[r]);
    load_seg<DW>(src, x_dw, y_rows, gy0 + PLAIN_RB, gx0, vec_ok, below);

    // One load per row fetches both wave-edge dwords: lane 0 the dword left of
    // the wavefront, lane 63 the dword right of it (other lanes re-read their own).
    uint32_t el[PLAIN_RB + 2], er[PLAIN_RB + 2];
    const int xe = lane == 0 ? wave_x0 - 1 : (lane == 63 ? wave_x0 + 64 * DW : gx0);
    const bool xe_in = xe >= 0 && xe < x_dw;
    const int cxe = xe < 0 ? 0 : (xe >= x_dw ? x_dw - 1 : xe);
#pragma unroll
    for (int i = 0; i < PLAIN_RB + 2; ++i) {
        const int gy = gy0 - 1 + i;
        const bool in = xe_in && gy >= 0 && gy < y_rows;
        const int cy = gy < 0 ? 0 : (gy >= y_rows ? y_rows - 1 : gy);
        const uint32_t v = src[static_cast<int64_t>(cy) * x_dw + cxe];
        el[i] = (in && lane == 0) ? v : 0u;
        er[i] = (in && lane == 63) ? v : 0u;
    }
    uint32_t colmask[DW];
#pragma unroll
    for (int d = 0; d < DW; ++d) colmask[d] = ~0u;
    advance_band<DW, PLAIN_RB, false, false, true>(s, above, below, 0, PLAIN_RB, colmask, el, er);

#pragma unroll
    for (int r = 0; r < PLAIN_RB; ++r) {
        if (gy0 + r < y_rows) store_seg<DW>(dst, x_dw, gy0 + r, gx0, vec_ok, s[r]);
    }
}

// ---------------------------------------------------------------------------
// Temporal tile: one workgroup advances one tile of one grid by `gens`
// generations and writes the tile's interior. Shared by the single-grid
// launch, the active-tile launch and the batched work-item launch.
// ---------------------------------------------------------------------------

// LDS exchange of band edge rows, double-buffered by generation parity:
//   tops[p][w]        top row of wave w      (tops[p][NW] stays zero)
//   bottoms[p][w + 1] bottom row of wave w   (bottoms[p][0] stays zero)
// so wave w reads bottoms[p][w] (row above) and tops[p][w + 1] (row below)
// without a branch; the zero slots are the tile's top and bottom edges.
template <int DW, int NW>
struct EdgeRows {
    uint32_t tops[2][NW + 1][64 * DW];
    uint32_t bottoms[2][NW + 1][64 * DW];
};

template <int DW>
__device__ __forceinline__ void lds_store_seg(uint32_t* row, int lane, const uint32_t (&v)[DW]) {
    if constexpr (DW >= 4) {
#pragma unroll
        for (int k = 0; k < DW / 4; ++k)
            reinterpret_cast<uint4*>(row + lane * DW)[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int d = 0; d < DW; ++d) row[lane * DW + d] = v[d];
    }
}

template <int DW>
__device__ __forceinline__ void lds_load_seg(const uint32_t* row, int lane, uint32_t (&v)[DW]) {
    if constexpr (DW >= 4) {
#pragma unroll
        for (int k = 0; k < DW / 4; ++k) {
            const uint4 q = reinterpret_cast<const uint4*>(row + lane * DW)[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int d = 0; d < DW; ++d) v[d] = row[lane * DW + d];
    }
}

template <int DW, int NW>
__device__ __forceinline__ void clear_edge_slots(EdgeRows<DW, NW>& lds, int lane, int w) {
    uint32_t z[DW];
#pragma unroll
    for (int d = 0; d < DW; ++d) z[d] = 0;
    if (w == 0) {
        lds_store_seg<DW>(lds.bottoms[0][0], lane, z);
        lds_store_seg<DW>(lds.bottoms[1][0], lane, z);
    }
    if (w == NW - 1) {
        lds_store_seg<DW>(lds.tops[0][NW], lane, z);
        lds_store_seg<DW>(lds.tops[1][NW], lane, z);
    }
}

template <int DW, int BR, int NW, bool MASK_ROWS, bool MASK_COLS>
__device__ __forceinline__ void run_generations(uint32_t (&s)[BR][DW], EdgeRows<DW, NW>& lds, int gens, int lane,
                                                int w, int row_lo, int row_hi, const uint32_t (&colmask)[DW]) {
    for (int g = 0; g < gens; ++g) {
        const int par = g & 1;
        lds_store_seg<DW>(lds.tops[par][w], lane, s[0]);
        lds_store_seg<DW>(lds.bottoms[par][w + 1], lane, s[BR - 1]);
        __syncthreads();
        uint32_t above[DW], below[DW];
        lds_load_seg<DW>(lds.bottoms[par][w], lane, above);
        lds_load_seg<DW>(lds.tops[par][w + 1], lane, below);
        advance_band<DW, BR, MASK_ROWS, MASK_COLS>(s, above, below, row_lo, row_hi, colmask);
    }
}

// Advances one tile by gens generations and stores its interior.
template <int DW, int BR, int NW, bool COH = false>
__device__ __forceinline__ void process_tile(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                             const TileGeo& a, int tile, int gens, EdgeRows<DW, NW>& lds) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int tx = tile % a.tiles_x;
    const int ty = tile / a.tiles_x;
    const int x0 = tx * a.stride_dwords - a.halo_lanes * DW;
    const int y0 = ty * a.stride_rows - a.halo_rows;
    const int gx0 = x0 + lane * DW;
    const int gy0 = y0 + w * BR;
    const bool vec_ok = (a.x_dw % DW) == 0;

    uint32_t s[BR][DW];
#pragma unroll
    for (int r = 0; r < BR; ++r) load_seg<DW, COH>(src, a.x_dw, a.y_rows, gy0 + r, gx0, vec_ok, s[r]);
    clear_edge_slots<DW, NW>(lds, lane, w);

    uint32_t colmask[DW];
#pragma unroll
    for (int d = 0; d < DW; ++d) colmask[d] = (gx0 + d >= 0 && gx0 + d < a.x_dw) ? ~0u : 0u;
    // band rows [row_lo, row_hi) lie inside the grid (wave-uniform)
    const int row_lo = gy0 < 0 ? -gy0 : 0;
    const int row_hi = a.y_rows - gy0 < BR ? a.y_rows - gy0 : BR;

    const bool edge_y = y0 < 0 || y0 + NW * BR > a.y_rows;  // workgroup-uniform
    const bool edge_x = x0 < 0 || x0 + 64 * DW > a.x_dw;
    if (edge_x) {
        run_generations<DW, BR, NW, true, true>(s, lds, gens, lane, w, row_lo, row_hi, colmask);
    } else if (edge_y) {
        run_generations<DW, BR, NW, true, false>(s, lds, gens, lane, w, row_lo, row_hi, colmask);
    } else {
        run_generations<DW, BR, NW, false, false>(s, lds, gens, lane, w, row_lo, row_hi, colmask);
    }

    const int tile_rows = NW * BR;
    const bool lane_inside = lane >= a.halo_lanes && lane < 64 - a.halo_lanes;
#pragma unroll
    for (int r = 0; r < BR; ++r) {
        const int tr = w * BR + r;
        const int gy = gy0 + r;
        if (tr >= a.halo_rows && tr < tile_rows - a.halo_rows && gy >= 0 && gy < a.y_rows && lane_inside)
            store_seg<DW, COH>(dst, a.x_dw, gy, gx0, vec_ok, s[r]);
    }
}

template <int DW, int BR, int NW>
__global__ __launch_bounds__(64 * NW) void tile_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                       TileGeo a, int gens) {
    __shared__ EdgeRows<DW, NW> lds;
    process_tile<DW, BR, NW>(src, dst, a, blockIdx.x, gens, lds);
}

// ---------------------------------------------------------------------------
// Change tracking over small tiles. A wavefront advances one tile of DW
// dwords (32*DW cells) x tile_rows rows by up to T generations: lane l holds
// grid row y0 - T + l, so the wavefront spans the tile's rows and T halo rows
// above and below (tile_rows = 64 - 2T), as the tile's DW dwords plus halo
// dwords. For T <= 16 one halo dword serves both sides (the ring form, DW + 1
// dwords): its bits 16..31 hold cells 16..31 of the dword west of the tile and
// its bits 0..15 cells 0..15 of the dword east of it, each at its own bit, so
// the lane's dwords form a ring (the halo dword is the west neighbour of the
// tile's first dword and the east neighbour of its last) and every dword
// takes the same two funnel shifts. The seam inside the halo dword (bits 15
// and 16 see each other as neighbours) spoils one cell per generation from
// 16 cells out, so the tile's cells stay exact for 16 generations. For
// deeper blocks a whole halo dword lies on each side (DW + 2 dwords, exact
// for 32 generations). Rows above and below
// a lane's row come from the neighbouring lanes (DPP), a dword's west and east
// neighbours from the lane's own dwords, so a wavefront needs no LDS and no
// barrier, and the wavefronts of a workgroup take independent tiles. A halo
// cell is wrong after g generations only if it lies within g cells of the
// wavefront's edge; after T generations the tile's own cells are exact.
//
// The kernel reads and writes the grid tile-major (fine_word_index): the
// tiles in row-major order, each tile_rows x DW dwords, row-major inside. A
// wavefront's own rows are then one contiguous run of 4 * DW * tile_rows
// bytes, its halo rows the end of the tile above and the start of the tile
// below, and its west and east halo dwords one dword of every row of the
// neighbouring tiles' columns: 12 lines of 128 bytes for a 64 x 32 tile and
// its halo, where the row-major grid puts each of the 64 rows in a line of
// its own. The step converts the grid to this layout before its first block
// and back after its last (fine_pack_kernel, fine_unpack_kernel).
// ---------------------------------------------------------------------------
template <int N, bool MASK, bool RING>
__device__ __forceinline__ void fine_generation(uint32_t (&s)[N], uint32_t rowmask, const uint32_t (&colmask)[N]) {
    uint32_t h0[N], h1[N], e0[N], e1[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t w = west_of(j > 0 ? s[j - 1] : (RING ? s[N - 1] : 0u), s[j]);
        const uint32_t e = east_of(s[j], j + 1 < N ? s[j + 1] : (RING ? s[0] : 0u));
        const RowSum r = row_sum(w, s[j], e);
        h0[j] = r.h0;
        h1[j] = r.h1;
        e0[j] = r.e0;
        e1[j] = r.e1;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t a0 = from_lower_lane(h0[j]), a1 = from_lower_lane(h1[j]);  // row above
        const uint32_t b0 = from_upper_lane(h0[j]), b1 = from_upper_lane(h1[j]);  // row below
        uint32_t v = next_state(a0, a1, e0[j], e1[j], b0, b1, s[j]);
        if (MASK) v &= rowmask & colmask[j];
        s[j] = v;
    }
}

template <int N, bool MASK, bool RING>
__device__ __forceinline__ void fine_generations(uint32_t (&s)[N], int gens, uint32_t rowmask,
                                                 const uint32_t (&colmask)[N]) {
    for (int g = 0; g < gens; ++g) fine_generation<N, MASK, RING>(s, rowmask, colmask);
}

template <int DW>
struct FineRow {
    uint32_t v[DW + 2];
};

// The words a lane reads for one tile: its row's DW tile dwords, then the
// dwords west and east of the tile. Lane l holds row l - T of the tile's rows:
// lanes below T read the last T rows of the tile above, lanes from T +
// tile_rows the first T rows of the tile below. Where such a row or a halo
// dword lies outside the grid the address is clamped into the tile itself
// (the masks applied after the load clear what it reads). No branch around
// the loads, so all of them are in flight together, and nothing waits for
// them here. The tile's own dwords are one vector load (two for 8 dwords).
template <int DW>
__device__ __forceinline__ void fine_load(const uint32_t* __restrict__ src, const FineGeo& g, int tile,
                                          FineRow<DW>& r) {
    const int lane = threadIdx.x & 63;
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int k = lane < g.halo_rows ? -1 : (lane < g.halo_rows + g.tile_rows ? 0 : 1);  // tile row ty + k
    const int row = lane - g.halo_rows - k * g.tile_rows;                                // row in that tile
    const int tyy = ty + k >= 0 && ty + k < g.tiles_y ? ty + k : ty;
    const size_t tile_dw = static_cast<size_t>(g.tile_rows) * DW;
    const uint32_t* line = src + static_cast<size_t>(tyy) * g.tiles_x * tile_dw + static_cast<size_t>(row) * DW;
    const uint32_t* own = line + tx * tile_dw;
    if constexpr (DW == 2) {
        const uint2 w = *reinterpret_cast<const uint2*>(own);
        r.v[0] = w.x;
        r.v[1] = w.y;
    } else {
#pragma unroll
        for (int q = 0; q < DW / 4; ++q) {
            const uint4 w = reinterpret_cast<const uint4*>(own)[q];
            r.v[4 * q] = w.x;
            r.v[4 * q + 1] = w.y;
            r.v[4 * q + 2] = w.z;
            r.v[4 * q + 3] = w.w;
        }
    }
    r.v[DW] = line[(tx > 0 ? tx - 1 : tx) * tile_dw + DW - 1];
    r.v[DW + 1] = line[(tx + 1 < g.tiles_x ? tx + 1 : tx) * tile_dw];
}

// Advances one tile, whose words fine_load read, by gens generations. With
// tracking (tr.list_next set), a tile whose cells changed appends itself and
// each of its eight neighbours that a changed cell lies within T cells of
// (Chebyshev distance, T = halo_rows, the generations of a tracked block: the
// neighbour's next block sees exactly the cells within T of it) to the next
// block's list: a tile joins once, when its mark goes from 0 to 1, to
// sub-list fine_sub_list(nx, ny, DW) of the new tile, whose counter reserves
// its entry. Sub-lists spread the reservations,
// device-scope atomics that one counter would take one after another, over
// kFineLists counters in lines of their own. The wavefront clears the mark
// of a listed tile it advances (its list is done with it).
template <int DW, bool RING>
__device__ __forceinline__ void fine_tile(const FineRow<DW>& r, uint32_t* __restrict__ dst, const FineGeo& g,
                                          int gens, int tile, bool listed, const FineTrack& tr) {
    constexpr int N = RING ? DW + 1 : DW + 2;  // s[0]: (west) halo, s[1..DW]: the tile, s[DW + 1]: east halo
    const int lane = threadIdx.x & 63;
    const int tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const int x0 = tx * DW - 1;                      // the dword west of the tile
    const int y0 = ty * g.tile_rows - g.halo_rows;   // row of lane 0
    const int y = y0 + lane;
    const bool row_in = y >= 0 && y < g.y_rows;
    const bool west_in = x0 >= 0, east_in = x0 + DW + 1 < g.x_dw;  // the halo dwords lie in the grid
    const uint32_t rowmask = row_in ? ~0u : 0u;

    uint32_t s[N], colmask[N];
#pragma unroll
    for (int j = 1; j <= DW; ++j) colmask[j] = x0 + j < g.x_dw ? ~0u : 0u;
    if (RING) {
        colmask[0] = (west_in ? 0xffff0000u : 0u) | (east_in ? 0x0000ffffu : 0u);
        s[0] = ((r.v[DW] & 0xffff0000u) | (r.v[DW + 1] & 0x0000ffffu)) & colmask[0] & rowmask;
    } else {
        colmask[0] = west_in ? ~0u : 0u;
        colmask[N - 1] = east_in ? ~0u : 0u;
        s[0] = r.v[DW] & colmask[0] & rowmask;
        s[N - 1] = r.v[DW + 1] & colmask[N - 1] & rowmask;
    }
#pragma unroll
    for (int d = 0; d < DW; ++d) s[1 + d] = r.v[d] & colmask[1 + d] & rowmask;

    uint32_t before[DW];
#pragma unroll
    for (int d = 0; d < DW; ++d) before[d] = s[1 + d];

    const bool edge = y0 < 0 || y0 + 64 > g.y_rows || !west_in || !east_in;  // wave-uniform
    if (edge) fine_generations<N, true, RING>(s, gens, rowmask, colmask);
    else fine_generations<N, false, RING>(s, gens, rowmask, colmask);

    const bool own = lane >= g.halo_rows && lane < g.halo_rows + g.tile_rows && row_in;
    bool diff = false;
#pragma unroll
    for (int d = 0; d < DW; ++d) diff |= before[d] != s[1 + d];
    // changes within reach (halo_rows cells, the generations of a tracked
    // block) of the west and east neighbours: cells 0 .. reach - 1 of the
    // first dword, the last reach cells of the last one
    const int reach = g.halo_rows;
    const bool west_diff = ((before[0] ^ s[1]) & ((1u << reach) - 1u)) != 0u;
    const bool east_diff = ((before[DW - 1] ^ s[DW]) >> (32 - reach)) != 0u;
    if (own) {
        // the tile's row, padding dwords past the grid's width included (the
        // masks keep them 0)
        uint32_t* out = dst + (static_cast<size_t>(tile) * g.tile_rows + (lane - g.halo_rows)) * DW;
        if constexpr (DW == 2) {
            *reinterpret_cast<uint2*>(out) = make_uint2(s[1], s[2]);
        } else {
#pragma unroll
            for (int q = 0; q < DW / 4; ++q)
                reinterpret_cast<uint4*>(out)[q] = make_uint4(s[1 + 4 * q], s[2 + 4 * q], s[3 + 4 * q], s[4 + 4 * q]);
        }
    }
    if (!tr.list_next) return;
    if (lane == 0) {
        if (listed) tr.mark_cur[tile] = 0u;
        atomicAdd(&tr.wg_active[blockIdx.x], 1ull);
    }
    if (__ballot(own && diff) == 0) return;
    // A neighbour joins the next block's list only if a cell within its
    // reach changed (the cells it sees over the next block); the tile itself
    // joins if any of its cells changed. Bit 3 (dy + 1) + dx + 1 of `near`:
    // the neighbour at (tx + dx, ty + dy).
    const int row = lane - g.halo_rows;
    const bool high = row < reach, low = row >= g.tile_rows - reach;  // within reach of the tile above / below
    const bool west = own && west_diff, east = own && east_diff;
    const unsigned near = (__ballot(west && high) != 0 ? 1u : 0u) | (__ballot(own && diff && high) != 0 ? 2u : 0u) |
                          (__ballot(east && high) != 0 ? 4u : 0u) | (__ballot(west) != 0 ? 8u : 0u) | 16u |
                          (__ballot(east) != 0 ? 32u : 0u) | (__ballot(west && low) != 0 ? 64u : 0u) |
                          (__ballot(own && diff && low) != 0 ? 128u : 0u) | (__ballot(east && low) != 0 ? 256u : 0u);
    // lanes 0..8: the neighbours within reach join the next block's list
    if (lane < 9 && ((near >> lane) & 1u)) {
        const int nx = tx + lane % 3 - 1, ny = ty + lane / 3 - 1;
        if (nx >= 0 && nx < g.tiles_x && ny >= 0 && ny < g.tiles_y) {
            const int n = ny * g.tiles_x + nx;
            if (atomicOr(&tr.mark_next[n], 1u) == 0u) {
                const int j = fine_sub_list(nx, ny, DW);
                const int at = atomicAdd(tr.count_next + j * kFineCountStride, 1);
                tr.list_next[static_cast<size_t>(j) * tr.list_cap + at] = n;
            }
        }
    }
}

// A tracked block advances the tiles of kFineLists sub-lists: workgroup b (of
// at least kFineLists) takes sub-list b % kFineLists, the one its XCD owns,
// and its wavefront v entries r, r + step, ... with r = kFineWaves *
// (b / kFineLists) + v and step the number of wavefronts on that sub-list. An
// untracked block advances tiles w, w + W, ... < dense_tiles, wavefront
// numbers w going to the workgroups in runs of 128 per XCD (neighbouring
// tiles, which share lines, on one L2), W = kFineWaves * gridDim.x. Either
// way a wavefront reads one tile
// ahead: the loads of the next tile are in flight while the current tile
// advances (a block lists more tiles than the launch has wavefronts when much
// of the grid is active). The entry after it is read before the tile
// advances, a scalar load the wavefront waits for. The launch resets the
// counters of the list after next.
template <int DW, bool RING>
__global__ __launch_bounds__(64 * kFineWaves) void fine_tile_kernel(const uint32_t* __restrict__ src,
                                                                    uint32_t* __restrict__ dst, FineGeo g, int gens,
                                                                    const int* __restrict__ list,
                                                                    const int* __restrict__ count, int listed,
                                                                    int dense_tiles, FineTrack tr) {
    if (tr.count_clear && blockIdx.x == 0 && threadIdx.x < kFineLists)
        tr.count_clear[threadIdx.x * kFineCountStride] = 0;
    const int nwg = static_cast<int>(gridDim.x);
    const int b = static_cast<int>(blockIdx.x);
    const int v = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    // Untracked: workgroup b runs on XCD b % 8; runs of 128 consecutive
    // wavefront numbers go to 32 workgroups of one XCD, the runs dealt to the
    // XCDs in turn. A bijection onto the workgroups when their number is a
    // multiple of 256; else workgroup b takes numbers 4b .. 4b + 3.
    const int group = (nwg & 255) == 0 ? ((b >> 8) * 8 + (b & 7)) * 32 + ((b >> 3) & 31) : b;
    // sub-list, first entry and step of a tracked block; for an untracked one
    // the "sub-list" is the whole range of tiles
    const int sub = b % kFineLists;
    const int cap = listed ? tr.list_cap : dense_tiles;
    const int step = listed ? kFineWaves * ((nwg - sub + kFineLists - 1) / kFineLists) : nwg * kFineWaves;
    int at = listed ? (b / kFineLists) * kFineWaves + v : group * kFineWaves + v;
    const int* entries = list + static_cast<size_t>(sub) * tr.list_cap;
    auto tile_at = [&](int i) { return listed ? __builtin_amdgcn_readfirstlane(entries[i]) : i; };
    // The first entry is read together with the sub-list's length, both loads
    // unconditional (list and count are valid also for a block that advances
    // every tile), the entry's index clamped into the sub-list: the wavefront
    // waits once, for both (the empty asm takes both values, so nothing that
    // uses one is placed between the loads).
    int first = __builtin_amdgcn_readfirstlane(entries[at < tr.list_cap ? at : tr.list_cap - 1]);
    int length = __builtin_amdgcn_readfirstlane(count[sub * kFineCountStride]);
    asm volatile("" : "+s"(first), "+s"(length));
    // The first workgroup of each sub-list hands its length to the host, with
    // the block's tag in the same 64-bit store (host memory the device writes
    // through: no copy, no event and no wait in the stream).
    if (tr.sample && listed && b < kFineLists && threadIdx.x == 0)
        __hip_atomic_store(tr.sample + sub,
                           (static_cast<unsigned long long>(static_cast<unsigned>(tr.sample_tag)) << 32) |
                               static_cast<unsigned>(length),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const int n = listed ? length : cap;
    if (at >= n) return;
    int tile = listed ? first : at;
    int next_tile = at + step < n ? tile_at(at + step) : -1;
    FineRow<DW> cur;
    fine_load<DW>(src, g, tile, cur);
    for (;;) {
        // the words arrive here, not earlier: nothing the compiler would hoist
        // to the loads that fetch them makes the wavefront wait for them there
#pragma unroll
        for (int i = 0; i < DW + 2; ++i) asm volatile("" : "+v"(cur.v[i]));
        FineRow<DW> nxt;
        int after = -1;
        if (next_tile >= 0) {
            fine_load<DW>(src, g, next_tile, nxt);
            if (at + 2 * step < n) after = tile_at(at + 2 * step);
        }
        // keeps the loads here, issued before the tile advances (the compiler
        // would otherwise sink them to their use); it reads no loaded word, so
        // nothing waits for them
        asm volatile("" ::: "memory");
        fine_tile<DW, RING>(cur, dst, g, gens, tile, listed != 0, tr);
        if (next_tile < 0) return;
        at += step;
        tile = next_tile;
        next_tile = after;
        cur = nxt;
    }
}

// Batched grids: workgroup i advances work item i (a tile of one grid).
template <int DW, int BR, int NW>
__global__ __launch_bounds__(64 * NW) void item_kernel(const BatchGrid* __restrict__ grids,
                                                       const WorkItem* __restrict__ items) {
    __shared__ EdgeRows<DW, NW> lds;
    const WorkItem it = items[blockIdx.x];
    const BatchGrid& g = grids[it.grid];
    const uint32_t* src = g.buf[it.parity];
    uint32_t* dst = g.buf[it.parity ^ 1];
    process_tile<DW, BR, NW>(src, dst, g.geo, it.tile, it.gens, lds);
}

// ---------------------------------------------------------------------------
// Narrow grids (batches and servers): lane l of wavefront w holds grid row
// 64*w + l as DW dwords, the whole grid in registers for all the item's
// generations, no halo. A row's west and east neighbours come from the lane's
// own dwords, the rows above and below from the neighbouring lanes (DPP); the
// rows across a wavefront edge come from the neighbouring wavefront's edge
// lanes, which publish their 3-cell row sums in double-buffered LDS slots (one
// barrier per generation). Rows and dwords outside the grid are cleared after
// every generation (dead boundary). Narrow grids of one row width share a
// launch sized for the tallest: wavefronts wholly below a grid only take part
// in the barriers.
// ---------------------------------------------------------------------------
template <int DW>
struct NarrowEdges {
    uint32_t v[2][kNarrowMaxWaves][4][DW];  // [parity][wave][top h0, top h1, bottom h0, bottom h1]
};

// Advances one narrow item; every wavefront of the workgroup takes part in
// it.gens barriers.
template <int DW, bool COH = false>
__device__ __forceinline__ void narrow_item(const BatchGrid& bg, const WorkItem& it, NarrowEdges<DW>& lds) {
    auto& edge = lds.v;
    const uint32_t* src = bg.buf[it.parity];
    uint32_t* dst = bg.buf[it.parity ^ 1];
    const int x_dw = bg.geo.x_dw, y_rows = bg.geo.y_rows;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));  // wave-uniform
    const int waves_used = (y_rows + 63) >> 6;
    if (w >= waves_used) {  // below the grid: keep the barrier count
        for (int g = 0; g < it.gens; ++g) __syncthreads();
        return;
    }
    const int y = w * 64 + lane;
    const bool row_in = y < y_rows;
    const uint32_t rowmask = row_in ? ~0u : 0u;
    const uint32_t* row = src + static_cast<size_t>(row_in ? y : 0) * x_dw;
    uint32_t s[DW];
#pragma unroll
    for (int j = 0; j < DW; ++j) {
        const uint32_t* p = row + (j < x_dw ? j : 0);
        const uint32_t v = COH ? __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : *p;
        s[j] = (j < x_dw ? v : 0u) & rowmask;
    }
    const bool mask_cols = x_dw < DW;  // wave-uniform
    for (int g = 0; g < it.gens; ++g) {
        uint32_t h0[DW], h1[DW], e0[DW], e1[DW];
#pragma unroll
        for (int j = 0; j < DW; ++j) {
            const uint32_t wp = west_of(j > 0 ? s[j - 1] : 0u, s[j]);
            const uint32_t ep = east_of(s[j], j + 1 < DW ? s[j + 1] : 0u);
            const RowSum r = row_sum(wp, s[j], ep);
            h0[j] = r.h0;
            h1[j] = r.h1;
            e0[j] = r.e0;
            e1[j] = r.e1;
        }
        const int par = g & 1;
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < DW; ++j) {
                edge[par][w][0][j] = h0[j];
                edge[par][w][1][j] = h1[j];
            }
        } else if (lane == 63) {
#pragma unroll
            for (int j = 0; j < DW; ++j) {
                edge[par][w][2][j] = h0[j];
                edge[par][w][3][j] = h1[j];
            }
        }
        __syncthreads();
        const int up = w > 0 ? w - 1 : 0, down = w + 1 < waves_used ? w + 1 : w;
        const uint32_t up_ok = w > 0 ? ~0u : 0u, down_ok = w + 1 < waves_used ? ~0u : 0u;
#pragma unroll
        for (int j = 0; j < DW; ++j) {
            // lane 0 takes the row above from the wavefront above, lane 63 the
            // row below from the wavefront below (dead beyond the grid)
            const uint32_t a0 = from_lower_lane_or(h0[j], edge[par][up][2][j] & up_ok);
            const uint32_t a1 = from_lower_lane_or(h1[j], edge[par][up][3][j] & up_ok);
            const uint32_t b0 = from_upper_lane_or(h0[j], edge[par][down][0][j] & down_ok);
            const uint32_t b1 = from_upper_lane_or(h1[j], edge[par][down][1][j] & down_ok);
            s[j] = next_state(a0, a1, e0[j], e1[j], b0, b1, s[j]) & rowmask;
        }
        if (mask_cols) {
#pragma unroll
            for (int j = 0; j < DW; ++j) s[j] = j < x_dw ? s[j] : 0u;
        }
    }
    if (row_in) {
#pragma unroll
        for (int j = 0; j < DW; ++j)
            if (j < x_dw) {
                uint32_t* p = dst + static_cast<size_t>(y) * x_dw + j;
                if (COH)
                    __hip_atomic_store(p, s[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    *p = s[j];
            }
    }
}

template <int DW>
__global__ __launch_bounds__(64 * kNarrowMaxWaves) void narrow_item_kernel(const BatchGrid* __restrict__ grids,
                                                                           const WorkItem* __restrict__ items) {
    __shared__ NarrowEdges<DW> lds;
    const WorkItem it = items[blockIdx.x];
    narrow_item<DW>(grids[it.grid], it, lds);
}

// ---------------------------------------------------------------------------
// Persistent batch launch: the workgroups of one launch drain a device work
// queue of items of one kernel shape. Thread 0 takes the next ticket with an
// atomic add; an item that continues a grid (after > 0) first waits until that
// many of the grid's items have completed. Items load and store their cells
// with device-coherent accesses (COH above), so the predecessors' cells are
// visible to an item on any XCD once their stores are acknowledged: a
// completed item counts itself in its grid's counter only after every
// wavefront has waited for its stores (s_waitcnt 0) and the workgroup has met
// at a barrier, and a waiting item loads its cells only after it has read the
// count. This replaces an agent-scope release/acquire pair per item: each of
// those writes back or invalidates a whole L2, which made the 4096-grid batch
// 11 % slower (209 ms against 188 ms). Every item is queued after the items it waits for, so the
// unfinished item with the lowest ticket can always run and the queue drains.
// Waits are bounded by device time: one that runs out raises *error, and
// every workgroup leaves at its next step. A wait can legitimately last as
// long as two whole-grid items of up to 2^20 generations (its predecessor may
// be the next ticket of a workgroup still running another item), so the bound
// is elapsed time, not a poll count.
// ---------------------------------------------------------------------------

// A failed wait raises *error for the host and sets the top bit of the ticket
// counter, so every later ticket is past the queue's end and each workgroup
// stops at its next step without reading the error word.
constexpr unsigned int kTicketStop = 1u << 31;

__device__ bool wait_grid_done(const QueueArgs& q, unsigned int* done, unsigned int target) {
    const unsigned long long start = wall_clock64();
    for (int i = 0;; ++i) {
        if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        if (__hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return false;
        if (q.spin_limit > 0 ? i >= q.spin_limit : wall_clock64() - start > q.wait_ticks) break;
        __builtin_amdgcn_s_sleep(2);
    }
    __hip_atomic_fetch_or(q.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(q.next, kTicketStop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}

constexpr unsigned int kNoTicket = ~0u;

// Thread 0's part of a queue step: count the workgroup's previous item
// (prev_grid >= 0) in its grid's counter, then take `ticket` (taken while the
// previous item ran; kNoTicket: take one now) and wait until its item's
// predecessors have completed. Returns the item's slot, or -1 once the queue
// is drained or a wait failed. Completion and the wait form one region on
// purpose: two thread-0 branches with only the loop's back edge between them
// get threaded into one by the compiler, which leaves the other lanes of
// wavefront 0 looping over the barrier without thread 0.
__device__ __forceinline__ int queue_step(const QueueArgs& q, int prev_grid, unsigned int ticket) {
    if (prev_grid >= 0)
        __hip_atomic_fetch_add(q.done + prev_grid, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int t = ticket != kNoTicket
                               ? ticket
                               : __hip_atomic_fetch_add(q.next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t >= static_cast<unsigned int>(q.n_items)) return -1;
    const WorkItem& it = q.items[t];
    if (it.after > 0 && !wait_grid_done(q, q.done + it.grid, static_cast<unsigned int>(it.after) + q.wait_ahead))
        return -1;
    return static_cast<int>(t);
}

// The drain loop of a persistent launch: `process` advances one item with
// every thread of the workgroup. Thread 0 takes the next ticket as an item
// starts, so the atomic's round trip overlaps the item. Before an item is
// counted, each wavefront waits for its stores to be acknowledged and the
// workgroup meets at a barrier, so the count follows every wavefront's
// stores. A ticket held while an item runs is always higher than that item's,
// so the unfinished item with the lowest ticket is never behind another item.
template <typename Process>
__device__ __forceinline__ void drain_queue(const QueueArgs& q, int* slot_lds, Process&& process) {
    int prev_grid = -1;
    unsigned int ticket = kNoTicket;
    for (;;) {
        if (threadIdx.x == 0) *slot_lds = queue_step(q, prev_grid, ticket);
        __syncthreads();
        const int slot = __builtin_amdgcn_readfirstlane(*slot_lds);
        if (slot < 0) return;
        if (threadIdx.x == 0) ticket = __hip_atomic_fetch_add(q.next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const WorkItem it = q.items[slot];
        process(it);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        prev_grid = it.grid;
    }
}

template <int DW, int BR, int NW>
__global__ __launch_bounds__(64 * NW) void queue_item_kernel(QueueArgs q) {
    __shared__ EdgeRows<DW, NW> lds;
    __shared__ int slot_lds;
    drain_queue(q, &slot_lds, [&](const WorkItem& it) {
        const BatchGrid& g = q.grids[it.grid];
        process_tile<DW, BR, NW, true>(g.buf[it.parity], g.buf[it.parity ^ 1], g.geo, it.tile, it.gens, lds);
    });
}

template <int DW>
__global__ __launch_bounds__(64 * kNarrowMaxWaves) void queue_narrow_kernel(QueueArgs q) {
    __shared__ NarrowEdges<DW> lds;
    __shared__ int slot_lds;
    drain_queue(q, &slot_lds, [&](const WorkItem& it) { narrow_item<DW, true>(q.grids[it.grid], it, lds); });
}

__global__ void zero_kernel(unsigned int* words, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0u;
}

// ---------------------------------------------------------------------------
// Persistent slab kernel: one cooperative launch advances a grid whose rows a
// single tile spans (tiles_x == 1) by many blocks of T generations. Workgroup
// ty keeps its tile (interior rows plus T halo rows above and below) in VGPRs
// for the whole launch. After each block it exports its first and last T
// interior rows to an exchange buffer, publishes the block count, waits for
// the two neighbouring slabs to publish theirs and imports their rows as its
// new halos. The tile load and store happen once per launch instead of once
// per block. Two exchange buffers alternate by block parity: a slab rewrites
// its rows of block b only at block b + 2, after both neighbours published
// block b + 1, which they do only after importing its block b rows. The
// counters carry an epoch across launches and are never reset. Waits are
// bounded: a slab that waits too long raises *error and every slab stops, so
// the launch always ends.
// ---------------------------------------------------------------------------
constexpr int kSlabSpinLimit = 1 << 20;  // polls before a wait gives up (SlabArgs::spin_limit 0)

// Relaxed polling of a neighbour's block counter.
__device__ bool wait_counter(const unsigned long long* p, unsigned long long target, unsigned int* error,
                             int spin_limit) {
    for (int i = 0; i < spin_limit; ++i) {
        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}

// Halo rows cross XCDs, each with its own L2, through device-scope (agent)
// stores and loads (sc1): such a store is acknowledged once every CU of the
// device can see it, and such a load reads at that level, so the exchange
// needs no cache-wide write-back or invalidate. Such stores are written
// through one by one, so the exchange rows are laid out lane-interleaved: a
// row is 32*DW dword pairs, and pair k of lane i (dwords 2k, 2k+1 of its
// segment) sits at slot k*64 + i, so each store instruction of a wavefront
// writes 512 contiguous bytes. Slots past the grid's width hold zeros.
template <int DW>
__device__ __forceinline__ void publish_seg(unsigned long long* row, int lane, const uint32_t (&v)[DW]) {
    static_assert(DW % 2 == 0, "slab rows are published as dword pairs");
#pragma unroll
    for (int k = 0; k < DW / 2; ++k) {
        const unsigned long long pair =
            static_cast<unsigned long long>(v[2 * k]) | (static_cast<unsigned long long>(v[2 * k + 1]) << 32);
        __hip_atomic_store(row + k * 64 + lane, pair, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ unsigned long long fetch_pair(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The block loop of one slab; the row and column masks are fixed for the
// slab, so each instantiation holds one copy of the band update.
//
// Between blocks the exchange runs beside the next block's first generation:
//   1. after a block, each wavefront holding exported rows stores them
//      (device scope) and goes on;
//   2. each exporting wavefront waits for its stores' acknowledgement
//      (s_waitcnt 0) and then adds one to the slab's counter (device-scope
//      atomic);
//   3. each wavefront holding halo rows polls its neighbour's counter (lane 0)
//      until all the neighbour's exporting wavefronts have counted the block,
//      then loads the rows into its registers.
// Where every band boundary lies among the own rows (band_rows > T, the
// 80-row slab at T <= 8), no band reads a halo row across a boundary, so the
// next block's first generation starts with its LDS edge exchange and barrier
// at once: after that barrier the exporting wavefronts count and the
// importing ones import, and then each advances its band, while the others
// advance theirs at once. Otherwise every exporting wavefront counts, and
// every importing one imports, before that barrier. No wait depends on a
// barrier the neighbour has not passed (a slab counts before it waits, and
// the barrier before the count needs no import), so the chain of waits
// between slabs always ends. Waits are bounded; a
// wavefront whose wait runs out clears keep_going, and every wavefront leaves
// after the next barrier.
template <int DW, int BR, int NW, bool MASK_ROWS, bool MASK_COLS>
__device__ __forceinline__ bool slab_blocks(const SlabArgs& a, uint32_t (&s)[BR][DW], EdgeRows<DW, NW>& lds,
                                            int* keep_going, int lane, int w, int row_lo, int row_hi,
                                            const uint32_t (&colmask)[DW]) {
    constexpr int RP = 32 * DW;  // dword pairs per exchanged row (one wavefront's 64 segments)
    const int ty = blockIdx.x;
    const int T = a.steps;
    const int tile_rows = NW * BR;
    const size_t plane = static_cast<size_t>(T) * RP;  // pairs of one side: T rows
    const int total_blocks = a.blocks + (a.last_gens > 0 ? 1 : 0);
    // wavefronts holding halo rows import them (wave-uniform)
    const bool import_top = ty > 0 && w * BR < T;
    const bool import_bottom = ty + 1 < a.tiles_y && (w + 1) * BR > tile_rows - T;
    const bool imports = import_top || import_bottom;
    const bool exports = (w * BR < 2 * T && (w + 1) * BR > T) ||
                         (w * BR < tile_rows - T && (w + 1) * BR > tile_rows - 2 * T);
    const unsigned long long per_block = static_cast<unsigned long long>(slab_publishers(BR, NW, T));
    const int spins = a.spin_limit > 0 ? a.spin_limit : kSlabSpinLimit;
    auto count_publication = [&]() {
        __builtin_amdgcn_s_waitcnt(0);  // this wavefront's exported rows are acknowledged
        if (lane == 0) __hip_atomic_fetch_add(&a.flags[ty], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // Subscribe and import: lane 0 polls the neighbour the rows come from
    // until its exporting wavefronts have all counted block b - 1; the
    // wavefront then loads the rows straight into its registers (device-scope
    // loads issued after the counter load returned, so they see the rows the
    // neighbour's acknowledged stores wrote). Lane i's dwords 2k, 2k+1 are
    // pair slot k*64 + i of a row.
    auto import_rows = [&](int par_b, int b) {
        const unsigned long long wanted = a.epoch + static_cast<unsigned long long>(b) * per_block + a.wait_ahead;
        int ok = 1;
        if (lane == 0) {
            if (import_top) ok = wait_counter(&a.flags[ty - 1], wanted, a.error, spins) ? 1 : 0;
            if (ok && import_bottom) ok = wait_counter(&a.flags[ty + 1], wanted, a.error, spins) ? 1 : 0;
        }
        ok = __builtin_amdgcn_readfirstlane(ok);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (ok) {
            const unsigned long long* up =
                a.xbuf + (static_cast<size_t>(par_b) * a.tiles_y + (ty > 0 ? ty - 1 : 0)) * 2 * plane + plane;
            const unsigned long long* dn =
                a.xbuf + (static_cast<size_t>(par_b) * a.tiles_y + (ty + 1 < a.tiles_y ? ty + 1 : ty)) * 2 * plane;
#pragma unroll
            for (int r = 0; r < BR; ++r) {
                const int tr = w * BR + r;
                const bool top = import_top && tr < T;
                const bool bottom = import_bottom && tr >= tile_rows - T;
                if (top || bottom) {
                    const unsigned long long* row =
                        top ? up + static_cast<size_t>(tr) * RP : dn + static_cast<size_t>(tr - (tile_rows - T)) * RP;
                    // the neighbour's rows are already clear outside the grid
                    // (its band update masks them), so each pair loads
                    // straight into the band's registers
#pragma unroll
                    for (int k = 0; k < DW / 2; ++k) {
                        const unsigned long long v = fetch_pair(row + k * 64 + lane);
                        s[r][2 * k] = static_cast<uint32_t>(v);
                        s[r][2 * k + 1] = static_cast<uint32_t>(v >> 32);
                    }
                }
            }
        } else if (lane == 0) {
            *keep_going = 0;
        }
    };
    // Whether the import may follow the first generation's LDS edge exchange:
    // only if no band edge row that another band reads is a halo row, i.e.
    // every band boundary lies among the own rows (band_rows > T). Otherwise
    // every exporting wavefront counts and every importing one imports before
    // that exchange.
    const bool late = BR > T;
    // The wavefronts on the exchange's path (exporting or importing) issue
    // ahead of the interior ones sharing their SIMD, so they finish each
    // generation first: the exported rows are stored, and acknowledged,
    // earlier, and the imported rows are advanced as soon as they arrive.
    if (exports || imports) __builtin_amdgcn_s_setprio(1);
    for (int b = 0; b < total_blocks; ++b) {
        const int gens = b < a.blocks ? T : a.last_gens;
        const int par_b = (b - 1) & 1;  // exchange parity of the previous block's rows
        // The block's first generation, with the import (its own copy of the
        // band update: a branch around the import inside the generation loop
        // would make every generation reshuffle the band's registers).
        {
            if (b > 0 && !late) {
                if (exports) count_publication();
                if (imports) import_rows(par_b, b);
            }
            lds_store_seg<DW>(lds.tops[0][w], lane, s[0]);
            lds_store_seg<DW>(lds.bottoms[0][w + 1], lane, s[BR - 1]);
            __syncthreads();
            if (!*keep_going) return false;  // a wait ran out: the whole workgroup leaves; the error is recorded
            uint32_t above[DW], below[DW];
            lds_load_seg<DW>(lds.bottoms[0][w], lane, above);
            lds_load_seg<DW>(lds.tops[0][w + 1], lane, below);
            if (b > 0 && late) {
                if (exports) count_publication();
                if (imports) import_rows(par_b, b);
            }
            advance_band<DW, BR, MASK_ROWS, MASK_COLS>(s, above, below, row_lo, row_hi, colmask);
        }
        // the block's other generations
        for (int g = 1; g < gens; ++g) {
            const int par = g & 1;
            lds_store_seg<DW>(lds.tops[par][w], lane, s[0]);
            lds_store_seg<DW>(lds.bottoms[par][w + 1], lane, s[BR - 1]);
            __syncthreads();
            if (!*keep_going) return false;
            uint32_t above[DW], below[DW];
            lds_load_seg<DW>(lds.bottoms[par][w], lane, above);
            lds_load_seg<DW>(lds.tops[par][w + 1], lane, below);
            advance_band<DW, BR, MASK_ROWS, MASK_COLS>(s, above, below, row_lo, row_hi, colmask);
        }
        if (b + 1 == total_blocks) break;

        // publish this block's exported rows (device-scope stores; counted in
        // the next block's first generation)
        const int par = b & 1;
        unsigned long long* mine = a.xbuf + (static_cast<size_t>(par) * a.tiles_y + ty) * 2 * plane;
#pragma unroll
        for (int r = 0; r < BR; ++r) {
            const int tr = w * BR + r;
            if (tr >= T && tr < 2 * T) publish_seg<DW>(mine + static_cast<size_t>(tr - T) * RP, lane, s[r]);
            if (tr >= tile_rows - 2 * T && tr < tile_rows - T)
                publish_seg<DW>(mine + plane + static_cast<size_t>(tr - (tile_rows - 2 * T)) * RP, lane, s[r]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();
    return *keep_going != 0;
}

template <int DW, int BR, int NW, bool MASK_COLS>
__global__ __launch_bounds__(64 * NW) void slab_kernel(SlabArgs a) {
    __shared__ EdgeRows<DW, NW> lds;
    __shared__ int keep_going;
    const TileGeo& g = a.geo;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int ty = blockIdx.x;
    const int tile_rows = NW * BR;
    const int y0 = ty * g.stride_rows - g.halo_rows;
    const int gx0 = lane * DW;
    const int gy0 = y0 + w * BR;
    const bool vec_ok = (g.x_dw % DW) == 0;

    uint32_t s[BR][DW];
#pragma unroll
    for (int r = 0; r < BR; ++r) load_seg<DW>(a.src, g.x_dw, g.y_rows, gy0 + r, gx0, vec_ok, s[r]);
    clear_edge_slots<DW, NW>(lds, lane, w);
    if (threadIdx.x == 0) keep_going = 1;

    uint32_t colmask[DW];
#pragma unroll
    for (int d = 0; d < DW; ++d) colmask[d] = (gx0 + d < g.x_dw) ? ~0u : 0u;
    const int row_lo = gy0 < 0 ? -gy0 : 0;
    const int row_hi = g.y_rows - gy0 < BR ? g.y_rows - gy0 : BR;
    const bool edge_y = y0 < 0 || y0 + tile_rows > g.y_rows;  // workgroup-uniform
    const bool done = edge_y ? slab_blocks<DW, BR, NW, true, MASK_COLS>(a, s, lds, &keep_going, lane, w, row_lo,
                                                                         row_hi, colmask)
                             : slab_blocks<DW, BR, NW, false, MASK_COLS>(a, s, lds, &keep_going, lane, w, row_lo,
                                                                          row_hi, colmask);
    if (!done) return;
#pragma unroll
    for (int r = 0; r < BR; ++r) {
        const int tr = w * BR + r;
        const int gy = gy0 + r;
        if (tr >= g.halo_rows && tr < tile_rows - g.halo_rows && gy >= 0 && gy < g.y_rows)
            store_seg<DW>(a.dst, g.x_dw, gy, gx0, vec_ok, s[r]);
    }
}

}  // namespace

const TileShape kTileShapes[] = {
    {8, 8, 8}, {8, 10, 8}, {8, 9, 8}, {8, 5, 16}, {8, 4, 16}, {4, 8, 8}, {4, 8, 16}, {2, 16, 8}, {1, 8, 8}, {1, 16, 8},
    {1, 24, 8}, {1, 32, 8}, {1, 32, 16},
};

TileGeo to_device_geo(const TileGeometry& g, int64_t x_dw, int64_t y_rows) {
    TileGeo a;
    a.x_dw = static_cast<int>(x_dw);
    a.y_rows = static_cast<int>(y_rows);
    a.halo_rows = g.halo_rows;
    a.halo_lanes = g.halo_lanes;
    a.stride_rows = g.stride_rows;
    a.stride_dwords = g.stride_dwords;
    a.tiles_x = g.tiles_x;
    return a;
}
const int kNumTileShapes = static_cast<int>(sizeof(kTileShapes) / sizeof(kTileShapes[0]));

bool plan_tiles(const TileShape& shape, int64_t x_dw, int64_t y_rows, int max_gens, TileGeometry* g,
                const char** why) {
    const int DW = shape.lane_dwords, BR = shape.band_rows, NW = shape.waves;
    g->lane_dwords = DW;
    g->band_rows = BR;
    g->waves = NW;
    g->tile_rows = NW * BR;
    g->tile_dwords = 64 * DW;
    if (x_dw <= g->tile_dwords) {
        g->halo_lanes = 0;
        g->stride_dwords = g->tile_dwords;
        g->tiles_x = 1;
    } else {
        if (max_gens > 32 * DW) {
            *why = "generations per launch exceed the one-lane x halo (32 * lane_dwords cells)";
            return false;
        }
        g->halo_lanes = 1;
        g->stride_dwords = (64 - 2) * DW;
        g->tiles_x = static_cast<int>((x_dw + g->stride_dwords - 1) / g->stride_dwords);
    }
    if (y_rows <= g->tile_rows) {
        g->halo_rows = 0;
        g->stride_rows = g->tile_rows;
        g->tiles_y = 1;
    } else {
        if (g->tile_rows - 2 * max_gens < 1) {
            *why = "generations per launch leave no interior rows in the tile";
            return false;
        }
        g->halo_rows = max_gens;
        g->stride_rows = g->tile_rows - 2 * max_gens;
        g->tiles_y = static_cast<int>((y_rows + g->stride_rows - 1) / g->stride_rows);
    }
    return true;
}

#define GOLHIP_FOR_EACH_SHAPE(X) \
    X(8, 8, 8)                     \
    X(8, 10, 8)                    \
    X(8, 9, 8)                     \
    X(8, 5, 16)                    \
    X(8, 4, 16)                    \
    X(4, 8, 8)                     \
    X(4, 8, 16)                    \
    X(2, 16, 8)                    \
    X(1, 8, 8)                     \
    X(1, 16, 8)                    \
    X(1, 24, 8)                    \
    X(1, 32, 8)                    \
    X(1, 32, 16)

hipError_t launch_tiles(const TileGeometry& g, const uint32_t* src, uint32_t* dst, int64_t x_dw, int64_t y_rows,
                        int gens, hipStream_t stream) {
    const unsigned blocks = static_cast<unsigned>(g.tiles_x) * static_cast<unsigned>(g.tiles_y);
#define GOLHIP_LAUNCH(DW, BR, NW)                                                                        \
    if (g.lane_dwords == DW && g.band_rows == BR && g.waves == NW) {                                     \
        const TileGeo a = to_device_geo(g, x_dw, y_rows);                                                \
        hipLaunchKernelGGL((tile_kernel<DW, BR, NW>), dim3(blocks), dim3(64 * NW), 0, stream, src, dst, a, gens); \
        return hipGetLastError();                                                                        \
    }
    GOLHIP_FOR_EACH_SHAPE(GOLHIP_LAUNCH)
#undef GOLHIP_LAUNCH
    return hipErrorInvalidValue;
}

static const void* tile_function(const TileShape& shape) {
#define GOLHIP_TILE_FN(DW, BR, NW)                                               \
    if (shape.lane_dwords == DW && shape.band_rows == BR && shape.waves == NW)   \
        return reinterpret_cast<const void*>(&tile_kernel<DW, BR, NW>);
    GOLHIP_FOR_EACH_SHAPE(GOLHIP_TILE_FN)
#undef GOLHIP_TILE_FN
    return nullptr;
}

int tile_occupancy(const TileShape& shape) {
    const void* f = tile_function(shape);
    int per_cu = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * shape.waves, 0) != hipSuccess)
        return 0;
    return per_cu;
}

// The change-tracking kernel for tiles of dw dwords, the ring form or not.
template <int DW>
static auto fine_kernel_for(bool ring) {
    return ring ? &fine_tile_kernel<DW, true> : &fine_tile_kernel<DW, false>;
}

static bool fine_ring(int halo_rows) { return halo_rows <= kFineRingMaxSteps; }

hipError_t launch_fine_tiles(const FineGeo& g, const uint32_t* src, uint32_t* dst, int gens, const int* list,
                             const int* count, bool listed, const FineTrack& track, hipStream_t stream) {
    if ((g.tile_dwords != 2 && g.tile_dwords != 4 && g.tile_dwords != 8) || gens > g.halo_rows ||
        g.halo_rows > kFineMaxSteps || !list || !count)
        return hipErrorInvalidValue;
    const int n = g.tiles_x * g.tiles_y;
    if (fine_workgroups(g) < kFineMinWorkgroups) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(fine_workgroups(g))), block(64 * kFineWaves);
    FineTrack t = track;
    t.list_cap = fine_list_cap(g);
    const bool ring = fine_ring(g.halo_rows);
    if (g.tile_dwords == 2)
        hipLaunchKernelGGL(fine_kernel_for<2>(ring), grid, block, 0, stream, src, dst, g, gens, list, count,
                           listed ? 1 : 0, n, t);
    else if (g.tile_dwords == 4)
        hipLaunchKernelGGL(fine_kernel_for<4>(ring), grid, block, 0, stream, src, dst, g, gens, list, count,
                           listed ? 1 : 0, n, t);
    else
        hipLaunchKernelGGL(fine_kernel_for<8>(ring), grid, block, 0, stream, src, dst, g, gens, list, count,
                           listed ? 1 : 0, n, t);
    return hipGetLastError();
}

int fine_resident_workgroups(int tile_dwords, int halo_rows) {
    const bool ring = fine_ring(halo_rows);
    const void* f = tile_dwords == 2   ? reinterpret_cast<const void*>(fine_kernel_for<2>(ring))
                    : tile_dwords == 4 ? reinterpret_cast<const void*>(fine_kernel_for<4>(ring))
                                       : reinterpret_cast<const void*>(fine_kernel_for<8>(ring));
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * kFineWaves, 0) != hipSuccess)
        return 0;
    return cus * per_cu;
}

// Row-major grid <-> the change-tracking kernel's tile-major words. Packing
// walks the padded extent (whole tiles) in row order, so its reads are
// coalesced, and writes 0 where the tiles reach past the grid; unpacking
// walks the grid in row order, so its writes are.
__global__ __launch_bounds__(256) void fine_pack_kernel(const uint32_t* __restrict__ rows,
                                                        uint32_t* __restrict__ tiles, FineGeo g) {
    const int64_t width = static_cast<int64_t>(g.tiles_x) * g.tile_dwords;
    const int64_t total = width * g.tiles_y * g.tile_rows;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t y = i / width;
        const int x = static_cast<int>(i - y * width);
        tiles[fine_word_index(g, x, y)] = y < g.y_rows && x < g.x_dw ? rows[y * g.x_dw + x] : 0u;
    }
}

__global__ __launch_bounds__(256) void fine_unpack_kernel(const uint32_t* __restrict__ tiles,
                                                          uint32_t* __restrict__ rows, FineGeo g) {
    const int64_t total = static_cast<int64_t>(g.x_dw) * g.y_rows;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t y = i / g.x_dw;
        const int x = static_cast<int>(i - y * g.x_dw);
        rows[i] = tiles[fine_word_index(g, x, y)];
    }
}

hipError_t launch_fine_layout(const FineGeo& g, const uint32_t* from, uint32_t* to, bool to_tiles,
                              hipStream_t stream) {
    const dim3 grid(2048), block(256);
    if (to_tiles)
        hipLaunchKernelGGL(fine_pack_kernel, grid, block, 0, stream, from, to, g);
    else
        hipLaunchKernelGGL(fine_unpack_kernel, grid, block, 0, stream, from, to, g);
    return hipGetLastError();
}

bool narrow_shape_for(int64_t x_dw, int64_t y_rows, TileShape* s) {
    if (x_dw > kNarrowMaxDwords || y_rows > 64 * kNarrowMaxWaves) return false;
    int dw = 2;
    while (dw < x_dw) dw *= 2;
    *s = TileShape{dw, 0, static_cast<int>((y_rows + 63) / 64)};
    return true;
}

TileGeometry narrow_geometry(const TileShape& s, int64_t x_dw, int64_t y_rows) {
    TileGeometry g{};
    g.lane_dwords = s.lane_dwords;
    g.band_rows = 0;
    g.waves = s.waves;
    g.tile_rows = 64 * s.waves;
    g.tile_dwords = s.lane_dwords;
    g.halo_rows = 0;
    g.halo_lanes = 0;
    g.stride_rows = static_cast<int>(y_rows);
    g.stride_dwords = static_cast<int>(x_dw);
    g.tiles_x = 1;
    g.tiles_y = 1;
    return g;
}

hipError_t launch_items(const TileShape& shape, const BatchGrid* grids, const WorkItem* items, int n_items,
                        hipStream_t stream) {
    if (n_items <= 0) return hipSuccess;
    if (is_narrow(shape)) {
        if (shape.waves < 1 || shape.waves > kNarrowMaxWaves) return hipErrorInvalidValue;
        const dim3 grid(static_cast<unsigned>(n_items)), block(64 * shape.waves);
        switch (shape.lane_dwords) {
            case 2: hipLaunchKernelGGL(narrow_item_kernel<2>, grid, block, 0, stream, grids, items); break;
            case 4: hipLaunchKernelGGL(narrow_item_kernel<4>, grid, block, 0, stream, grids, items); break;
            case 8: hipLaunchKernelGGL(narrow_item_kernel<8>, grid, block, 0, stream, grids, items); break;
            case 16: hipLaunchKernelGGL(narrow_item_kernel<16>, grid, block, 0, stream, grids, items); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
#define GOLHIP_LAUNCH(DW, BR, NW)                                                                        \
    if (shape.lane_dwords == DW && shape.band_rows == BR && shape.waves == NW) {                         \
        hipLaunchKernelGGL((item_kernel<DW, BR, NW>), dim3(n_items), dim3(64 * NW), 0, stream, grids, items); \
        return hipGetLastError();                                                                        \
    }
    GOLHIP_FOR_EACH_SHAPE(GOLHIP_LAUNCH)
#undef GOLHIP_LAUNCH
    return hipErrorInvalidValue;
}

static const void* queue_function(const TileShape& shape) {
    if (is_narrow(shape)) {
        switch (shape.lane_dwords) {
            case 2: return reinterpret_cast<const void*>(&queue_narrow_kernel<2>);
            case 4: return reinterpret_cast<const void*>(&queue_narrow_kernel<4>);
            case 8: return reinterpret_cast<const void*>(&queue_narrow_kernel<8>);
            case 16: return reinterpret_cast<const void*>(&queue_narrow_kernel<16>);
            default: return nullptr;
        }
    }
#define GOLHIP_QUEUE_FN(DW, BR, NW)                                                  \
    if (shape.lane_dwords == DW && shape.band_rows == BR && shape.waves == NW)       \
        return reinterpret_cast<const void*>(&queue_item_kernel<DW, BR, NW>);
    GOLHIP_FOR_EACH_SHAPE(GOLHIP_QUEUE_FN)
#undef GOLHIP_QUEUE_FN
    return nullptr;
}

int queue_capacity(const TileShape& shape) {
    const void* f = queue_function(shape);
    int dev = 0, cus = 0, per_cu = 0;
    if (!f || shape.waves < 1 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * shape.waves, 0) != hipSuccess)
        return 0;
    return cus * per_cu;
}

hipError_t launch_queue(const TileShape& shape, const QueueArgs& q, int workgroups, hipStream_t stream) {
    if (workgroups <= 0 || q.n_items <= 0) return hipSuccess;
    if (q.n_items > (1 << 30)) return hipErrorInvalidValue;  // tickets stay below the stop bit
    if (is_narrow(shape) && (shape.waves < 1 || shape.waves > kNarrowMaxWaves)) return hipErrorInvalidValue;
    const void* f = queue_function(shape);
    if (!f) return hipErrorInvalidValue;
    QueueArgs a = q;
    void* args[] = {&a};
    return hipLaunchKernel(f, dim3(static_cast<unsigned>(workgroups)), dim3(64 * shape.waves), args, 0, stream);
}

hipError_t launch_zero(unsigned int* words, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int blocks = n / 256 + 1 < 1024 ? n / 256 + 1 : 1024;
    hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, words, n);
    return hipGetLastError();
}

// The slab kernels compiled: full-width lanes of 8 dwords (rows of up to 16384 cells).
// 72-row slabs (9-row bands) at 4 generations per block keep 64 own rows each,
// as 80-row slabs at 8 do, with half the halo rows to recompute.
#define GOLHIP_FOR_EACH_SLAB_SHAPE(X) \
    X(8, 10, 8)                        \
    X(8, 9, 8)                         \
    X(8, 5, 16)                        \
    X(8, 4, 16)                        \
    X(8, 8, 8)

static const void* slab_function(const TileShape& shape, bool mask_cols) {
    const void* f = nullptr;
#define GOLHIP_SLAB_FN(DW, BR, NW)                                                                   \
    if (shape.lane_dwords == DW && shape.band_rows == BR && shape.waves == NW)                       \
        f = mask_cols ? reinterpret_cast<const void*>(&slab_kernel<DW, BR, NW, true>)                \
                      : reinterpret_cast<const void*>(&slab_kernel<DW, BR, NW, false>);
    GOLHIP_FOR_EACH_SLAB_SHAPE(GOLHIP_SLAB_FN)
#undef GOLHIP_SLAB_FN
    return f;
}

bool slab_capacity(const TileShape& shape, bool mask_cols, int* workgroups) {
    const void* f = slab_function(shape, mask_cols);
    if (!f) return false;
    int dev = 0, cus = 0, coop = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return false;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * shape.waves, 0) != hipSuccess) return false;
    *workgroups = per_cu * cus;
    return true;
}

hipError_t launch_slab(const TileShape& shape, const SlabArgs& args, hipStream_t stream) {
    const void* f = slab_function(shape, 64 * shape.lane_dwords > args.geo.x_dw);
    if (!f) return hipErrorInvalidValue;
    SlabArgs copy = args;
    void* kargs[] = {&copy};
    return hipLaunchCooperativeKernel(f, dim3(static_cast<unsigned>(args.tiles_y)), dim3(64 * shape.waves), kargs, 0,
                                      stream);
}

hipError_t launch_plain(int lane_dwords, const uint32_t* src, uint32_t* dst, int64_t x_dw, int64_t y_rows,
                        hipStream_t stream) {
    const int xd = static_cast<int>(x_dw), yr = static_cast<int>(y_rows);
    const unsigned gy = static_cast<unsigned>((y_rows + PLAIN_WAVES * PLAIN_RB - 1) / (PLAIN_WAVES * PLAIN_RB));
    auto gx = [&](int dw) { return static_cast<unsigned>((x_dw + 64 * dw - 1) / (64 * dw)); };
    switch (lane_dwords) {
        case 1:
            hipLaunchKernelGGL(plain_kernel<1>, dim3(gx(1), gy), dim3(64 * PLAIN_WAVES), 0, stream, src, dst, xd, yr);
            break;
        case 2:
            hipLaunchKernelGGL(plain_kernel<2>, dim3(gx(2), gy), dim3(64 * PLAIN_WAVES), 0, stream, src, dst, xd, yr);
            break;
        case 4:
            hipLaunchKernelGGL(plain_kernel<4>, dim3(gx(4), gy), dim3(64 * PLAIN_WAVES), 0, stream, src, dst, xd, yr);
            break;
        case 8:
            hipLaunchKernelGGL(plain_kernel<8>, dim3(gx(8), gy), dim3(64 * PLAIN_WAVES), 0, stream, src, dst, xd, yr);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace golhip
