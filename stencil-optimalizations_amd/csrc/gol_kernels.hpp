// Launch-side view of the stencil kernels (host code only sees this).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace golhip {

// Geometry of one temporal-tile launch over one grid.
struct TileGeometry {
    int lane_dwords;  // DW: dwords per lane (32*DW cells per lane, 64 lanes per wave)
    int band_rows;    // BR: rows per wavefront
    int waves;        // NW: wavefronts per workgroup, stacked vertically
    int tile_rows;    // NW*BR
    int tile_dwords;  // 64*DW
    int halo_rows;    // rows of halo above and below (0 when one tile spans the grid height)
    int halo_lanes;   // lanes of halo left and right (0 when one tile spans the grid width)
    int stride_rows;  // interior rows per tile
    int stride_dwords;// interior dwords per tile
    int tiles_x, tiles_y;
};

// Device-side geometry of one grid under one tile shape.
struct TileGeo {
    int x_dw, y_rows;
    int halo_rows, halo_lanes;
    int stride_rows, stride_dwords, tiles_x;
};

// One grid of a batch: its ping-pong buffers and tile geometry.
struct BatchGrid {
    uint32_t* buf[2];
    TileGeo geo;
};

// One unit of batched work: advance tile `tile` of grid `grid` by `gens`
// generations from buf[parity] into buf[parity ^ 1]. `after` is used by the
// persistent queue launch only: the item starts once that many items of its
// grid have completed in the run (the earlier rounds it reads or overwrites).
struct WorkItem {
    int grid;
    int tile;
    int gens;
    int parity;
    int after;
};

// Tile kernel shapes compiled into the library: (DW, BR, NW).
struct TileShape {
    int lane_dwords, band_rows, waves;
};
extern const TileShape kTileShapes[];
extern const int kNumTileShapes;

// Narrow grids in batches and servers (at most kNarrowMaxDwords dwords wide and
// 64 * kNarrowMaxWaves rows tall): lane l of wavefront w holds grid row
// 64 * w + l whole, the grid stays in registers for the item's generations and
// no halo is needed. A TileShape with band_rows == 0 names this kernel:
// {dwords per row (2, 4, 8 or 16), 0, wavefronts (1..8)}.
constexpr int kNarrowMaxDwords = 16;
constexpr int kNarrowMaxWaves = 8;
inline bool is_narrow(const TileShape& s) { return s.band_rows == 0; }
// The launch a grid's items go to: one per tile shape, and one per row width
// for narrow grids (launched with the tallest grid's wavefront count).
struct LaunchKey {
    int lane_dwords, band_rows, waves;
    bool operator<(const LaunchKey& o) const {
        if (lane_dwords != o.lane_dwords) return lane_dwords < o.lane_dwords;
        if (band_rows != o.band_rows) return band_rows < o.band_rows;
        return waves < o.waves;
    }
};
inline LaunchKey launch_key(const TileShape& s) {
    return is_narrow(s) ? LaunchKey{s.lane_dwords, 0, 0} : LaunchKey{s.lane_dwords, s.band_rows, s.waves};
}
// The narrow shape for a grid, if the grid is narrow enough.
bool narrow_shape_for(int64_t x_dw, int64_t y_rows, TileShape* s);
// Geometry of a narrow grid: one tile, no halo.
TileGeometry narrow_geometry(const TileShape& s, int64_t x_dw, int64_t y_rows);

// Compute the geometry for a grid of x_dw dwords x y_rows rows advanced by up
// to max_gens generations per launch. Returns false (with a reason) when the
// shape cannot cover max_gens generations.
bool plan_tiles(const TileShape& shape, int64_t x_dw, int64_t y_rows, int max_gens, TileGeometry* g,
                const char** why);

// One launch: advance src by gens (<= the planned max_gens) generations into dst.
hipError_t launch_tiles(const TileGeometry& g, const uint32_t* src, uint32_t* dst, int64_t x_dw, int64_t y_rows,
                        int gens, hipStream_t stream);

TileGeo to_device_geo(const TileGeometry& g, int64_t x_dw, int64_t y_rows);

// Workgroups of the tile kernel for `shape` one CU holds at once on the current device (0 if unknown).
int tile_occupancy(const TileShape& shape);

// Change tracking over small tiles: one wavefront per tile of tile_dwords
// (2, 4 or 8: 64, 128 or 256 cells) x tile_rows rows, lane l holding row y0 + l of
// the tile's rows and halo_rows (= T) halo rows above and below:
// tile_rows = 64 - 2T.
constexpr int kFineWaves = 4;      // wavefronts (tiles) per workgroup
constexpr int kFineDefaultDW = 2;  // tile width in dwords when gol_config.lane_dwords is 0
constexpr int kFineWideDW = 4;     // the width busy stretches switch to when lane_dwords is 0
// Width switching (lane_dwords 0): the host reads a block's list length
// every kFineChunk blocks, kFineLag - 1 chunks behind the GPU, and moves to
// 128-cell tiles when it exceeds kFineWidePercent of the 64-cell tiles, back
// when it falls below kFineNarrowPercent of the 128-cell tiles.
constexpr int kFineChunk = 32;
constexpr int kFineLag = 3;
constexpr int kFineWidePercent = 7;
constexpr int kFineNarrowPercent = 5;
constexpr int kFineMaxSteps = 21;  // T <= tile_rows: a tile is at least as tall as T generations reach
// Blocks of T <= 16 generations need at most 16 halo cells on each side of a
// tile: one dword holds both (the ring form, DW + 1 dwords per lane); deeper
// blocks keep a whole halo dword on each side (DW + 2).
constexpr int kFineRingMaxSteps = 16;

struct FineGeo {
    int x_dw, y_rows;
    int tile_dwords;  // 2, 4 or 8
    int tiles_x, tiles_y;
    int tile_rows;    // 64 - 2 * halo_rows
    int halo_rows;    // T: generations per block
    int workgroups;   // launch size: min(tiles / kFineWaves, resident workgroups)
};

// Workgroups of a change-tracking launch (kFineWaves wavefronts each).
inline int fine_workgroups(const FineGeo& g) { return g.workgroups; }

// The change-tracking kernel keeps the grid tile-major: tile (tx, ty) is
// tile_rows x tile_dwords dwords, row-major, at dword (ty * tiles_x + tx) *
// tile_rows * tile_dwords; the tiles past the grid's right and bottom edges
// are padded with zero dwords. Index of dword x of row y.
__host__ __device__ inline size_t fine_word_index(const FineGeo& g, int x, int64_t y) {
    const int tx = x / g.tile_dwords, ty = static_cast<int>(y / g.tile_rows);
    const int r = static_cast<int>(y - static_cast<int64_t>(ty) * g.tile_rows);
    return ((static_cast<size_t>(ty) * g.tiles_x + tx) * g.tile_rows + r) * g.tile_dwords + (x - tx * g.tile_dwords);
}
// Dwords of the tile-major grid.
inline size_t fine_words(const FineGeo& g) {
    return static_cast<size_t>(g.tiles_x) * g.tiles_y * g.tile_rows * g.tile_dwords;
}

// Workgroups the change-tracking kernel for tiles of tile_dwords and blocks of
// halo_rows generations can hold resident at once on the current device (0 if
// unknown).
int fine_resident_workgroups(int tile_dwords, int halo_rows);

// Converts a row-major grid (g.y_rows rows of g.x_dw dwords) to the
// tile-major words (to_tiles) or back.
hipError_t launch_fine_layout(const FineGeo& g, const uint32_t* from, uint32_t* to, bool to_tiles,
                              hipStream_t stream);

// A block's list of tiles is kFineLists sub-lists, one per XCD: sub-list j
// is taken by the workgroups b with b % kFineLists == j, which run on XCD j.
// Tile (tx, ty) goes to sub-list (c + 5 ty) mod kFineLists, c = the tile's
// 128-byte line column (tx * tile_dwords / 32): the tiles that share the
// lines of a row stay on one XCD, which fetches those lines once, and a strip
// of activity along a row, a column or a diagonal spreads over all the
// sub-lists. In each row of tiles a sub-list takes every kFineLists-th line
// column, so it holds at most list_cap = tiles_y * (32 / tile_dwords) *
// ceil(line columns / kFineLists) entries. The length of sub-list j is
// counts[j * kFineCountStride] (each counter in a 128-byte line of its own).
constexpr int kFineLists = 8;
constexpr int kFineCountStride = 32;
constexpr int kFineCounts = kFineLists * kFineCountStride;  // ints of one list's counters
__host__ __device__ inline int fine_sub_list(int tx, int ty, int tile_dwords) {
    return (((tx * tile_dwords) >> 5) + 5 * ty) & (kFineLists - 1);
}
inline int fine_list_cap(const FineGeo& g) {
    const int per_line = 32 / g.tile_dwords;
    const int columns = (g.tiles_x + per_line - 1) / per_line;
    return g.tiles_y * per_line * ((columns + kFineLists - 1) / kFineLists);
}
// Workgroups a tracked launch needs at least: one per sub-list.
constexpr int kFineMinWorkgroups = kFineLists;

// Change-tracking state of one block (all null: no tracking).
struct FineTrack {
    uint32_t* mark_cur;             // marks of the list being advanced: cleared as its tiles are taken
    uint32_t* mark_next;            // marks of the next block's list (all 0 when the block starts)
    int* list_next;                 // next block's list (kFineLists sub-lists), built by appending
    int* count_next;                // its sub-lists' lengths (0 when the block starts)
    int* count_clear;               // lengths of the list after next: reset to 0
    unsigned long long* wg_active;  // tiles advanced, per workgroup
    unsigned long long* sample;     // or null: host-visible memory that receives, per sub-list j,
                                    // sample_tag << 32 | the length of the sub-list this block advances
    int sample_tag;                 // > 0
    int list_cap;                   // entries per sub-list: set by launch_fine_tiles (fine_list_cap)
};

// One change-tracking block: advance the tiles of list's sub-lists (lengths
// in count) when listed, else every tile, by gens (<= halo_rows)
// generations; with track set, the tiles whose cells changed append
// themselves and the neighbours a changed cell lies within halo_rows cells
// of to the next list. list (kFineLists * track.list_cap
// entries) and count (kFineLists counters) are valid device pointers either
// way: the kernel reads both before it knows which it uses.
hipError_t launch_fine_tiles(const FineGeo& g, const uint32_t* src, uint32_t* dst, int gens, const int* list,
                             const int* count, bool listed, const FineTrack& track, hipStream_t stream);

// Batched grids: one workgroup per work item, all items of one tile shape.
hipError_t launch_items(const TileShape& shape, const BatchGrid* grids, const WorkItem* items, int n_items,
                        hipStream_t stream);

// Persistent batch launch: `workgroups` workgroups drain a device work queue
// of items of one kernel shape (a narrow shape launches 64 * waves threads).
// Every item must follow, in the queue, the items of its grid it waits for
// (WorkItem::after). next and done start at 0; a wait that runs out of polls
// raises *error, and every workgroup then leaves.
struct QueueArgs {
    const BatchGrid* grids;
    const WorkItem* items;
    int n_items;
    int spin_limit;           // > 0 (testing): polls before a wait gives up, instead of wait_ticks
    unsigned long long wait_ticks;  // device wall-clock ticks (wall_clock64) before a wait gives up
    unsigned int* next;       // ticket counter
    unsigned int* done;       // per grid: items completed in this run
    unsigned int* error;
    unsigned int wait_ahead;  // 0; testing the failure path: waits for counts no item reaches
};
hipError_t launch_queue(const TileShape& shape, const QueueArgs& q, int workgroups, hipStream_t stream);
// Workgroups of the queue kernel for `shape` resident at once on the current device (0 if unknown).
int queue_capacity(const TileShape& shape);
// Sets n words to 0 (a kernel, so that it may be recorded into a graph).
hipError_t launch_zero(unsigned int* words, int n, hipStream_t stream);

// Persistent slab launch over a grid that one tile spans in x (tiles_x == 1).
constexpr int kSlabMaxSteps = 8;  // deepest halo (generations per block) a slab launch takes
struct SlabArgs {
    const uint32_t* src;
    uint32_t* dst;
    TileGeo geo;                // halo_rows == steps
    int steps;                  // T: generations per block (the halo depth), at most kSlabMaxSteps (8)
    int blocks;                 // full blocks of T generations
    int last_gens;              // generations of a final shorter block (0: none)
    int tiles_y;                // workgroups
    unsigned long long* xbuf;   // [2][tiles_y][2][T][32 * DW] exchanged halo rows, lane-interleaved pairs
    unsigned long long* flags;  // [tiles_y] publications counted so far (slab_publishers per block), offset by epoch
    unsigned long long epoch;   // counter value every slab holds before this launch
    unsigned int* error;        // raised when a neighbour wait runs out
    int spin_limit;             // polls before a wait gives up; 0: the built-in limit
    unsigned long long wait_ahead;  // 0; testing the failure path: waits for counts no slab publishes
};

// Wavefronts of a slab holding exported rows ([T, 2T) or [R - 2T, R - T) of
// its R = band_rows * waves rows): each counts its own publication, so a
// slab's block counter grows by this many per exchanged block. The same for
// every slab (every slab publishes both sides).
__host__ __device__ inline int slab_publishers(int band_rows, int waves, int steps) {
    const int R = band_rows * waves, T = steps;
    int n = 0;
    for (int w = 0; w < waves; ++w) {
        const int b0 = w * band_rows, b1 = b0 + band_rows;
        if ((b0 < 2 * T && b1 > T) || (b0 < R - T && b1 > R - 2 * T)) ++n;
    }
    return n;
}

// Workgroups of the slab kernel that can be resident at once (false if the
// device cannot launch cooperatively); mask_cols selects the variant for rows
// narrower than a wavefront's 64 lanes.
bool slab_capacity(const TileShape& shape, bool mask_cols, int* workgroups);
hipError_t launch_slab(const TileShape& shape, const SlabArgs& args, hipStream_t stream);

// Bit layouts of the C ABI besides rows (gol_config.layout; the values of
// enum gol_layout): conversion from a layout's words to the row layout and
// back, on the stream. word_bits 16, 32 or 64; X a multiple of 32; Y a
// multiple of the word's rows (W for columns, 4 or 8 for tiles).
constexpr int kLayoutColumns = 1;
constexpr int kLayoutTiles = 2;
hipError_t layout_to_rows(int layout, int word_bits, const void* src, uint32_t* rows, int64_t x_cells,
                          int64_t y_rows, hipStream_t stream);
hipError_t rows_to_layout(int layout, int word_bits, const uint32_t* rows, void* dst, int64_t x_cells,
                          int64_t y_rows, hipStream_t stream);

// Native kernels of the reference's 8x8-tile and column layouts
// (GOL_FLAG_NATIVE_LAYOUT): the grid stays in the layout's words.
//   8x8 tiles, 64-bit words: a workgroup of kTile8Waves wavefronts, each lane
//   one tile column kTile8Rows tiles tall, so 64 tile columns x
//   kTile8Rows * kTile8Waves tile rows per workgroup, with a halo of
//   ceil(T / 8) tiles around it; X a multiple of 32 cells (4 tiles), Y of 8.
constexpr int kTile8Rows = 8;
constexpr int kTile8Waves = 8;
constexpr int kTile8MaxSteps = 16;
struct Tile8Geo {
    int x_tiles, y_tiles;    // grid size in tiles
    int halo_x, halo_y;      // halo tiles left and right / above and below (0 where one tile spans the grid)
    int stride_x, stride_y;  // interior tiles per workgroup
    int groups_x, groups_y;  // workgroups across and down
};
bool plan_tile8(int64_t x_tiles, int64_t y_tiles, int gens, Tile8Geo* g, const char** why);
// Advance gens (<= the planned) generations from src into dst (tile words).
hipError_t launch_tile8(const Tile8Geo& g, const uint32_t* src, uint32_t* dst, int gens, hipStream_t stream);

//   columns, 32- or 64-bit words: each wavefront its own tile of 64 grid
//   columns (one per lane) x kColDwords dwords of 32 rows, with T lanes of
//   halo left and right and one dword (32 rows) above and below; Y a multiple
//   of the word.
constexpr int kColDwords = 34;
constexpr int kColWaves = 4;
constexpr int kColMaxSteps = 16;
struct ColGeo {
    int x_cells, v_dwords;   // grid columns; dwords of 32 rows per column (Y / 32)
    int word_bits;           // 32 or 64
    int halo_x, halo_v;      // halo lanes left and right / dwords above and below
    int stride_x, stride_v;  // interior columns / dwords per tile
    int groups_x, groups_v;  // tiles across and down
};
bool plan_columns(int64_t x_cells, int64_t y_cells, int word_bits, int gens, ColGeo* g, const char** why);
hipError_t launch_columns(const ColGeo& g, const uint32_t* src, uint32_t* dst, int gens, hipStream_t stream);

// Plain kernel: one generation per launch, global memory to global memory.
// lane_dwords in {1,2,4,8}.
hipError_t launch_plain(int lane_dwords, const uint32_t* src, uint32_t* dst, int64_t x_dw, int64_t y_rows,
                        hipStream_t stream);

}  // namespace golhip
