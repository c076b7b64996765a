// Device-side bit logic of the B3/S23 stencil on row-packed cells.
//
// Layout (shared by rows-64 and rows-32, see below): a grid row of X cells is
// X/32 little-endian dwords; bit k of dword j is cell 32*j + k. A rows-64 word
// (wx, y) with bit k = cell 64*wx + k (algorithms/_shared/bitwise/bit_modes.hpp:127-163,
// general_bit_grid.hpp:217-246) is, in memory, exactly the dword pair
// (2*wx, 2*wx+1) of this layout, so one kernel serves both word widths.
//
// Counting scheme (written for gfx950's 3-input VALU ops: v_xor3_b32,
// v_bitop3_b32 and the funnel shift v_alignbit_b32):
//   per row r:  w = west neighbour plane, e = east neighbour plane
//               H(r) = w + c + e  as two bit planes (h0, h1)     (3-cell sum)
//               E(r) = w + e      as two bit planes (e0, e1)     (2-cell sum)
//   per cell:   n = H(r-1) + E(r) + H(r+1)   (eight neighbours, 0..8)
//               next = (n | alive) == 3
// Every row's H and E are built once and used by three output rows; the
// final combine costs six 3-input ops per dword. The reference computes the
// same rule per word with an explicit 9-word neighbourhood
// (_shared/bitwise/bitwise-ops/adder.cuh:17-51) or with column sums and lane
// shuffles (warp-exchange-full-adder-rows.cuh:19-88); this formulation is the
// row-sum dual of the latter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace golhip {

// v_bitop3_b32: any 3-input bitwise function, given by its truth table with
// the operands' weights a = 0xF0, b = 0xCC, c = 0xAA (e.g. a ^ b ^ c = 0x96).
template <uint32_t TABLE>
__device__ __forceinline__ uint32_t bitop3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:%4" : "=v"(d) : "v"(a), "v"(b), "v"(c), "n"(TABLE));
    return d;
}

constexpr uint32_t kXor3 = 0x96;      // a ^ b ^ c
constexpr uint32_t kMaj3 = 0xE8;      // majority(a, b, c): carry of a + b + c
constexpr uint32_t kXorAndNot = 0x14; // (a ^ b) & ~c
constexpr uint32_t kOrAnd = 0xA8;     // (a | b) & c

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return bitop3<kXor3>(a, b, c); }
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) { return bitop3<kMaj3>(a, b, c); }

// West neighbour plane of dword c: bit k holds cell k-1; bit 0 takes bit 31 of
// the dword to the left.
__device__ __forceinline__ uint32_t west_of(uint32_t left_dword, uint32_t c) {
    return __builtin_amdgcn_alignbit(c, left_dword, 31);
}

// East neighbour plane of dword c: bit k holds cell k+1; bit 31 takes bit 0 of
// the dword to the right.
__device__ __forceinline__ uint32_t east_of(uint32_t c, uint32_t right_dword) {
    return __builtin_amdgcn_alignbit(right_dword, c, 1);
}

struct RowSum {
    uint32_t h0, h1;  // west + centre + east, bits 0 and 1
    uint32_t e0, e1;  // west + east, bits 0 and 1
};

// Both sums of one dword from its west and east planes: 4 ops.
__device__ __forceinline__ RowSum row_sum(uint32_t w, uint32_t c, uint32_t e) {
    RowSum s;
    s.e0 = w ^ e;
    s.e1 = w & e;
    s.h0 = s.e0 ^ c;
    s.h1 = maj3(w, c, e);
    return s;
}

// Only the 3-cell sum (rows above and below a band need no E): 2 ops.
__device__ __forceinline__ void row_sum_h(uint32_t w, uint32_t c, uint32_t e, uint32_t& h0, uint32_t& h1) {
    h0 = xor3(w, c, e);
    h1 = maj3(w, c, e);
}

// n = A + E + B with A, B in 0..3 and E in 0..2:
//   n = l0 + 2*(lc + p) + 4*q,  l0/lc = sum/carry of the low bits, p/q of the high bits.
// (n | alive) == 3  <=>  (l0 | alive) & (lc + p == 1) & (q == 0).   6 ops.
// The six ops are one asm block: the compiler treats each inline-asm result
// conservatively and would separate dependent blocks with wait states.
__device__ __forceinline__ uint32_t next_state(uint32_t a0, uint32_t a1, uint32_t e0, uint32_t e1,
                                               uint32_t b0, uint32_t b1, uint32_t alive) {
    uint32_t out, l0, lc, p, q;
    asm("v_bitop3_b32 %1, %5, %6, %7 bitop3:0x96\n\t"   // l0 = a0 ^ e0 ^ b0
        "v_bitop3_b32 %2, %5, %6, %7 bitop3:0xe8\n\t"   // lc = maj(a0, e0, b0)
        "v_bitop3_b32 %3, %8, %9, %10 bitop3:0x96\n\t"  // p  = a1 ^ e1 ^ b1
        "v_bitop3_b32 %4, %8, %9, %10 bitop3:0xe8\n\t"  // q  = maj(a1, e1, b1)
        "v_bitop3_b32 %4, %2, %3, %4 bitop3:0x14\n\t"   // t  = (lc ^ p) & ~q
        "v_bitop3_b32 %0, %1, %11, %4 bitop3:0xa8"        // (l0 | alive) & t
        : "=v"(out), "=&v"(l0), "=&v"(lc), "=&v"(p), "=&v"(q)
        : "v"(a0), "v"(e0), "v"(b0), "v"(a1), "v"(e1), "v"(b1), "v"(alive));
    return out;
}

// Cross-lane word edges over the whole wave64 (DPP wave_shr:1 / wave_shl:1).
// Lane i receives v from lane i-1; lane 0 receives 0.
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x138, 0xf, 0xf, true));
}
// Lane i receives v from lane i+1; lane 63 receives 0.
__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x130, 0xf, 0xf, true));
}
// Same, but the lane at the wave edge receives `edge` instead of 0.
__device__ __forceinline__ uint32_t from_lower_lane_or(uint32_t v, uint32_t edge) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ uint32_t from_upper_lane_or(uint32_t v, uint32_t edge) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(edge), static_cast<int>(v), 0x130, 0xf, 0xf, false));
}

// Advance a band of BR rows x DW dwords held by one lane (the wave holds BR
// rows x 64*DW dwords) by one generation, in place.
//   above/below: the rows just outside the band (same dword columns).
//   MASK_ROWS: band rows outside [row_lo, row_hi) (wave-uniform) lie outside
//   the grid. Only the row next to the grid (row_lo - 1, row_hi) is cleared
//   after the update: rows further out reach the grid only through it, so
//   the grid sees a dead boundary while they hold whatever they evolve to.
//   The clearing is a wave-uniform branch per row, no VALU work per dword.
//   MASK_COLS: dwords with colmask 0 lie outside the grid and are cleared.
//   EDGE_IN: wave_edge_l/r[i] hold the dword just left of lane 0's first dword
//   / right of lane 63's last dword for rows above (i=0), band (1..BR) and
//   below (BR+1); only lanes 0 / 63 read them. Without EDGE_IN the wave edges
//   see dead cells (a tile edge, covered by the tile's halo).
//
// Columns are processed left to right, each from the top row down, keeping
// the running row sums of one column. Before a dword is overwritten, the west
// plane of the dword to its right is formed (wnext), so the new value can
// take the old one's register. Lane edges need the neighbouring lane's
// pre-update values: the west edge (left lane's last column) is read while the
// first column is processed, before that lane writes its last column; the east
// edge (right lane's first column) is written during the first column pass, so
// bit 0 of each of its rows is shifted into one dword before any write.
template <int DW, int BR, bool MASK_ROWS, bool MASK_COLS, bool EDGE_IN = false>
__device__ __forceinline__ void advance_band(uint32_t (&s)[BR][DW], const uint32_t (&above)[DW],
                                             const uint32_t (&below)[DW], int row_lo, int row_hi,
                                             const uint32_t (&colmask)[DW],
                                             const uint32_t* wave_edge_l = nullptr,
                                             const uint32_t* wave_edge_r = nullptr) {
    static_assert(DW == 1 || BR + 2 <= 32, "east-edge bits of a band are packed into one dword");
    auto edge_from_left = [&](uint32_t v, int i) -> uint32_t {
        return EDGE_IN ? from_lower_lane_or(v, wave_edge_l[i]) : from_lower_lane(v);
    };
    auto edge_from_right = [&](uint32_t v, int i) -> uint32_t {
        return EDGE_IN ? from_upper_lane_or(v, wave_edge_r[i]) : from_upper_lane(v);
    };

    // East edges for DW > 1: row i (0 = above, 1..BR = band, BR+1 = below) of
    // the right lane's first column contributes its bit 0, which ends in bit
    // (30 - BR + i) of east_bits.
    uint32_t east_bits = 0;
    if (DW > 1) {
        uint32_t pack = __builtin_amdgcn_alignbit(above[0], 0u, 1);
#pragma unroll
        for (int r = 0; r < BR; ++r) pack = __builtin_amdgcn_alignbit(s[r][0], pack, 1);
        pack = __builtin_amdgcn_alignbit(below[0], pack, 1);
        if (EDGE_IN) {
            uint32_t epack = __builtin_amdgcn_alignbit(wave_edge_r[0], 0u, 1);
#pragma unroll
            for (int i = 1; i < BR + 2; ++i) epack = __builtin_amdgcn_alignbit(wave_edge_r[i], epack, 1);
            east_bits = from_upper_lane_or(pack, epack);
        } else {
            east_bits = from_upper_lane(pack);
        }
    }
    auto east_edge = [&](int i) -> uint32_t { return east_bits >> (30 - BR + i); };

    uint32_t wnext[BR];  // west plane of column d of each band row (formed in pass d-1)
#pragma unroll
    for (int d = 0; d < DW; ++d) {
        uint32_t pa0, pa1;  // H of the row above the current one
        {
            const uint32_t c = above[d];
            const uint32_t w = west_of(d == 0 ? edge_from_left(above[DW - 1], 0) : above[d - 1], c);
            const uint32_t e = east_of(c, d < DW - 1 ? above[d + 1] : (DW == 1 ? edge_from_right(c, 0) : east_edge(0)));
            row_sum_h(w, c, e, pa0, pa1);
        }
        auto band_row_sum = [&](int r) -> RowSum {
            const uint32_t c = s[r][d];
            const uint32_t w = d == 0 ? west_of(edge_from_left(s[r][DW - 1], r + 1), c) : wnext[r];
            const uint32_t e =
                east_of(c, d < DW - 1 ? s[r][d + 1] : (DW == 1 ? edge_from_right(c, r + 1) : east_edge(r + 1)));
            return row_sum(w, c, e);
        };
        RowSum cur = band_row_sum(0);
#pragma unroll
        for (int r = 0; r < BR; ++r) {
            RowSum nxt;
            if (r + 1 < BR) {
                nxt = band_row_sum(r + 1);
            } else {
                const uint32_t c = below[d];
                const uint32_t w = west_of(d == 0 ? edge_from_left(below[DW - 1], BR + 1) : below[d - 1], c);
                const uint32_t e =
                    east_of(c, d < DW - 1 ? below[d + 1] : (DW == 1 ? edge_from_right(c, BR + 1) : east_edge(BR + 1)));
                row_sum_h(w, c, e, nxt.h0, nxt.h1);
                nxt.e0 = nxt.e1 = 0;
            }
            if (d < DW - 1) wnext[r] = west_of(s[r][d], s[r][d + 1]);
            uint32_t v = next_state(pa0, pa1, cur.e0, cur.e1, nxt.h0, nxt.h1, s[r][d]);
            if (MASK_COLS) v &= colmask[d];
            s[r][d] = v;
            pa0 = cur.h0;
            pa1 = cur.h1;
            cur = nxt;
            // keep the scheduler from hoisting later rows' cross-lane reads
            // (it would hold every edge of the band live at once)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (MASK_ROWS) {
        // within a generation only old values are read, so clearing the
        // boundary row now equals clearing it as it is computed
#pragma unroll
        for (int r = 0; r < BR; ++r) {
            if (r == row_lo - 1 || r == row_hi) {
#pragma unroll
                for (int d = 0; d < DW; ++d) s[r][d] = 0u;
            }
        }
    }
}

}  // namespace golhip
