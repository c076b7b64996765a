// C ABI over the HIP kernels: one handle = one grid's device state, its
// ping-pong buffers and its stream. Declarations and the reference
// interfaces each call replaces are in include/golhip.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/golhip.h"
#include "gol_capi_internal.hpp"
#include "gol_kernels.hpp"

namespace golhip {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int fail_hip(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? GOL_ERR_OUT_OF_MEMORY : GOL_ERR_HIP, "%s: %s (%d)", what,
                hipGetErrorString(e), static_cast<int>(e));
}

void clear_error() { g_last_error.clear(); }

const char* last_error() { return g_last_error.c_str(); }

// Default tile shape by row width: the narrowest lane width whose wavefront
// spans the row, so one tile covers the full width when it can (no x halo).
TileShape default_shape(int64_t x_dw, int64_t y_rows) {
    if (x_dw <= 64) return (y_rows <= 512 && y_rows > 128) ? TileShape{1, 32, 16} : TileShape{1, 16, 8};
    if (x_dw <= 128) return TileShape{2, 16, 8};
    if (x_dw <= 256) return TileShape{4, 8, 8};
    return TileShape{8, 10, 8};
}

// Time model of temporal launches on MI355X, used to choose a tile shape,
// generations per launch and the launch form for a single grid. On 16384^2
// random grids it stays within about 10 % of the measured launch times of every
// shape and depth tried:
//   compute: each CU advances its tiles one after another; a dword-generation
//            costs about 12.7 + 8 / band_rows VALU ops (13.5 at 10 rows: the
//            rows just above and below a band are summed once per band) at
//            64 lanes per clock, 2.4 GHz;
//   memory:  every tile loads its rows with halo and stores its interior, at
//            about 5 TB/s of HBM bandwidth. With one tile per CU at a time
//            (occupancy 1) this does not overlap compute; with several
//            resident per CU and more tiles than CUs, other tiles' compute hides
//            it except for about one tile's share;
//   launch:  about 4 us between dependent launches;
//   persistent slabs: one tile per CU for the whole run, no per-block load,
//            store or launch; the halo exchange runs beside the next block's
//            first generation and leaves about 0.6 + 0.2 * T us per block
//            exposed (with the compute term, the measured 14.5 us per block of
//            4 generations on 72-row slabs and 7.5 us per block of 2 on
//            16384^2; 80-row slabs at 8 generations measure 31.9 us, 0.9 us
//            above the model).
double model_us_per_generation(const TileShape& s, const TileGeometry& g, int steps, int occupancy,
                               bool persistent) {
    const double cus = 256.0;
    const double tiles = static_cast<double>(g.tiles_x) * g.tiles_y;
    const double waves_of_tiles = std::ceil(tiles / cus);
    const double dwords_per_tile = static_cast<double>(s.band_rows) * s.waves * 64.0 * s.lane_dwords;
    const double ops = 12.7 + 8.0 / s.band_rows;
    const double compute_us = waves_of_tiles * dwords_per_tile * steps * ops / 64.0 / 2400.0;
    if (persistent) return (compute_us + 0.6 + 0.2 * steps) / steps;
    const double bytes = tiles * (static_cast<double>(g.tile_rows) + g.stride_rows) * 64.0 * s.lane_dwords * 4.0;
    const double memory_us = bytes / 5.0e6;
    const double busy = occupancy >= 2 && waves_of_tiles >= 2
                            ? std::max(compute_us, memory_us) + memory_us / waves_of_tiles
                            : compute_us + memory_us;
    return (busy + 4.0) / steps;
}

// Whether the persistent slab launch takes this shape and depth for the grid:
// one tile spans the width, every slab is resident at once, the interiors are
// at least as tall as the exchanged halos and blocks are at most kSlabMaxSteps
// generations.
bool slab_fits(const TileShape& s, const TileGeometry& g, int64_t x_dw, int steps) {
    if (g.tiles_x != 1 || g.tiles_y < 2 || g.stride_rows < steps || steps > kSlabMaxSteps || s.lane_dwords % 2 != 0 ||
        x_dw > 64 * s.lane_dwords)
        return false;
    int capacity = 0;
    return slab_capacity(s, 64 * s.lane_dwords > x_dw, &capacity) && capacity >= g.tiles_y;
}

// Shape and generations per launch with the lowest modelled time per
// generation; allow_persistent prices shapes the slab launch takes as slabs.
bool choose_temporal(int64_t x_dw, int64_t y_rows, int fixed_dw, int fixed_br, int fixed_nw, int fixed_steps,
                     bool allow_persistent, TileShape* best_shape, int* best_steps) {
    static const int kSteps[] = {2, 4, 6, 8, 10, 12, 16, 24, 32};
    double best = 1e300;
    bool found = false;
    for (int i = 0; i < kNumTileShapes; ++i) {
        const TileShape& s = kTileShapes[i];
        if ((fixed_dw && s.lane_dwords != fixed_dw) || (fixed_br && s.band_rows != fixed_br) ||
            (fixed_nw && s.waves != fixed_nw))
            continue;
        const bool one_tile = x_dw <= 64 * s.lane_dwords && y_rows <= s.band_rows * s.waves;
        const int occupancy = tile_occupancy(s);
        for (int steps : kSteps) {
            if (fixed_steps && steps != fixed_steps) continue;
            TileGeometry g;
            const char* why = nullptr;
            const int plan_steps = one_tile ? 1 : steps;
            if (!plan_tiles(s, x_dw, y_rows, plan_steps, &g, &why)) continue;
            const bool slab = allow_persistent && !one_tile && slab_fits(s, g, x_dw, steps);
            // narrow rows on wide lanes waste lanes: charge the full tile width
            const double t = model_us_per_generation(s, g, one_tile ? 64 : steps, occupancy, slab);
            if (t < best) {
                best = t;
                *best_shape = s;
                *best_steps = one_tile ? 0 : steps;
                found = true;
            }
        }
    }
    return found;
}

bool shape_compiled(const TileShape& s) {
    for (int i = 0; i < kNumTileShapes; ++i) {
        if (kTileShapes[i].lane_dwords == s.lane_dwords && kTileShapes[i].band_rows == s.band_rows &&
            kTileShapes[i].waves == s.waves)
            return true;
    }
    return false;
}

}  // namespace golhip

using namespace golhip;

struct gol_handle {
    gol_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    uint32_t* buf[2] = {nullptr, nullptr};
    int cur = 0;
    int64_t x_dw = 0, y_rows = 0;
    size_t n_dwords = 0;
    void* staging = nullptr;  // words of a non-row layout on the device (n_dwords * 4 bytes)
    uint32_t kernel = GOL_KERNEL_TEMPORAL;
    uint32_t plain_dw = 1;
    TileShape shape{1, 16, 8};
    int steps = 8;        // generations per temporal launch
    bool one_tile = false;  // one tile spans the grid: no halo, any number of generations per launch
    TileGeometry geom{};
    uint64_t launches = 0;
    uint64_t generations = 0;
    // change tracking (GOL_KERNEL_ACTIVE): small tiles, one wavefront each
    FineGeo fine{};               // the tile geometry in use (fine_geo[fine_sel])
    FineGeo fine_geo[2]{};        // 64-cell tiles, and 128-cell tiles when the width adapts
    int fine_sel = 0;
    bool fine_adapt = false;      // lane_dwords left to the library: the width follows the activity
    unsigned long long* fine_sample = nullptr;  // [kFineLag][kFineLists] pinned, coherent: tag << 32 | length
    unsigned long long* fine_sample_dev = nullptr;  // the same memory as the device addresses it
    int fine_sample_tag[kFineLag] = {};  // the tag each chunk's sample carries (0: no sample)
    int fine_sample_sel[kFineLag] = {};  // the geometry each sample was taken in
    int fine_tags = 0;                   // tags handed out so far
    uint64_t fine_switches = 0;   // tile-width changes so far
    uint32_t* fbuf[2] = {nullptr, nullptr};  // the grid tile-major (fine_word_index), fine_words dwords each
    int fcur = 0;                 // fbuf[fcur] holds the state while a step runs
    uint32_t* marks = nullptr;    // [3][n_tiles] tile is on list k (lists rotate by block)
    int* lists = nullptr;         // [3][kFineLists][fine_list_cap] active-tile lists
    int* counts = nullptr;        // [3][kFineLists][kFineCountStride] sub-list lengths
    unsigned long long* wg_active = nullptr;  // [fine_workgroups] tiles advanced, per workgroup
    int n_tiles = 0;
    uint64_t blocks = 0;          // change-tracking blocks run
    bool flags_valid = false;     // false: the next block treats every tile as active
    // persistent slab launch (GOL_KERNEL_TEMPORAL, one tile spans the width)
    bool persistent = false;
    unsigned long long* xbuf = nullptr;   // [2][tiles_y][2][steps][32 * lane_dwords] dword pairs
    unsigned long long* slab_flags = nullptr;  // [tiles_y]
    unsigned int* slab_error = nullptr;
    unsigned long long slab_epoch = 0;
    bool slab_check = false;              // a slab launch ran since the error flag was last read
    // native layout kernels (GOL_FLAG_NATIVE_LAYOUT): the buffers hold the layout's words
    uint32_t native = 0;  // 0, or the enum gol_layout advanced in its own words
    Tile8Geo t8{};
    ColGeo cols{};
};

namespace {

int validate_config(const gol_config* c) {
    if (!c) return fail(GOL_ERR_INVALID_ARGUMENT, "config is null");
    if (c->layout > GOL_LAYOUT_TILES) return fail(GOL_ERR_INVALID_ARGUMENT, "unknown layout %u", c->layout);
    if (c->layout == GOL_LAYOUT_ROWS && c->word_bits != 32 && c->word_bits != 64)
        return fail(GOL_ERR_INVALID_ARGUMENT, "word_bits must be 32 or 64, got %u", c->word_bits);
    if (c->layout != GOL_LAYOUT_ROWS && c->word_bits != 16 && c->word_bits != 32 && c->word_bits != 64)
        return fail(GOL_ERR_INVALID_ARGUMENT, "word_bits must be 16, 32 or 64, got %u", c->word_bits);
    if (c->x_cells == 0 || c->y_cells == 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "grid dimensions must be positive");
    const uint32_t x_multiple = c->layout == GOL_LAYOUT_ROWS ? c->word_bits : 32;
    if (c->x_cells % x_multiple != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions X must be a multiple of %u", x_multiple);
    // words of the columns and tiles layouts span several rows
    const uint32_t y_multiple =
        c->layout == GOL_LAYOUT_COLUMNS ? c->word_bits : (c->layout == GOL_LAYOUT_TILES ? (c->word_bits == 64 ? 8 : 4) : 1);
    if (c->y_cells % y_multiple != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions Y must be a multiple of %u", y_multiple);
    if (c->x_cells / 32 > (1u << 26) || c->y_cells > (1u << 30) || (c->x_cells / 32) * c->y_cells > (1ull << 34))
        return fail(GOL_ERR_INVALID_ARGUMENT, "grid too large");
    if (c->kernel > GOL_KERNEL_ACTIVE) return fail(GOL_ERR_INVALID_ARGUMENT, "unknown kernel %u", c->kernel);
    if (c->lane_dwords != 0 && c->lane_dwords != 1 && c->lane_dwords != 2 && c->lane_dwords != 4 &&
        c->lane_dwords != 8)
        return fail(GOL_ERR_INVALID_ARGUMENT, "lane_dwords must be 0, 1, 2, 4 or 8");
    if (c->flags & GOL_FLAG_NATIVE_LAYOUT) {
        const bool tiles = c->layout == GOL_LAYOUT_TILES && c->word_bits == 64;
        const bool cols = c->layout == GOL_LAYOUT_COLUMNS && (c->word_bits == 32 || c->word_bits == 64);
        if (!tiles && !cols)
            return fail(GOL_ERR_INVALID_ARGUMENT, "GOL_FLAG_NATIVE_LAYOUT takes 64-bit 8x8 tiles or 32- / 64-bit "
                                                  "columns");
        if (c->kernel == GOL_KERNEL_ACTIVE)
            return fail(GOL_ERR_INVALID_ARGUMENT, "GOL_FLAG_NATIVE_LAYOUT runs the plain or temporal kernel, not "
                                                  "change tracking");
    }
    return GOL_OK;
}

// Resolve kernel, tile shape and generations per launch for a handle.
// Change-tracking tiles of dw dwords for blocks of `steps` generations; the
// launch holds no more workgroups than the device holds at once (wavefronts
// loop over list slots).
FineGeo fine_geometry(int64_t x_dw, int64_t y_rows, int dw, int steps) {
    FineGeo g{};
    g.x_dw = static_cast<int>(x_dw);
    g.y_rows = static_cast<int>(y_rows);
    g.tile_dwords = dw;
    g.halo_rows = steps;
    g.tile_rows = 64 - 2 * steps;
    g.tiles_x = static_cast<int>((x_dw + dw - 1) / dw);
    g.tiles_y = static_cast<int>((y_rows + g.tile_rows - 1) / g.tile_rows);
    const int all = (g.tiles_x * g.tiles_y + kFineWaves - 1) / kFineWaves;
    const int resident = fine_resident_workgroups(dw, steps);
    g.workgroups = std::max(resident > 0 ? std::min(all, resident) : all, kFineMinWorkgroups);
    return g;
}

int plan(gol_handle* h) {
    const gol_config& c = h->cfg;
    h->kernel = c.kernel == GOL_KERNEL_AUTO ? GOL_KERNEL_TEMPORAL : c.kernel;
    if (c.flags & GOL_FLAG_NATIVE_LAYOUT) {
        // the layout's own words: one generation per launch (plain) or up to
        // 16; by default 16 for tiles (a halo of 2 tiles) and 8 for columns
        // (8 halo lanes a side), the fastest of 4..16 on 16384^2
        const int fallback = c.layout == GOL_LAYOUT_TILES ? 16 : 8;
        const int steps =
            h->kernel == GOL_KERNEL_PLAIN ? 1 : (c.temporal_steps ? static_cast<int>(c.temporal_steps) : fallback);
        const char* why = nullptr;
        h->native = c.layout;
        h->steps = steps;
        const bool ok = c.layout == GOL_LAYOUT_TILES
                            ? plan_tile8(static_cast<int64_t>(c.x_cells / 8), static_cast<int64_t>(c.y_cells / 8),
                                         steps, &h->t8, &why)
                            : plan_columns(static_cast<int64_t>(c.x_cells), static_cast<int64_t>(c.y_cells),
                                           static_cast<int>(c.word_bits), steps, &h->cols, &why);
        if (!ok) return fail(GOL_ERR_INVALID_ARGUMENT, "temporal_steps %d: %s", steps, why);
        h->n_tiles = c.layout == GOL_LAYOUT_TILES ? h->t8.groups_x * h->t8.groups_y : h->cols.groups_x * h->cols.groups_v;
        return GOL_OK;
    }
    if (h->kernel == GOL_KERNEL_PLAIN) {
        uint32_t dw = c.lane_dwords;
        if (dw == 0) dw = h->x_dw <= 64 ? 1 : (h->x_dw <= 128 ? 2 : 4);
        h->plain_dw = dw;
        return GOL_OK;
    }
    if (h->kernel == GOL_KERNEL_ACTIVE) {
        // Change tracking skips quiet regions tile by tile, so its tiles are
        // small: 32 * lane_dwords cells (lane_dwords 2, 4 or 8) x (64 - 2T)
        // rows, one wavefront each. kFineRingMaxSteps (16) generations per
        // block by default, the deepest block one shared halo dword covers:
        // oscillators of period 1, 2, 4, 8 and 16 return to their state at
        // block boundaries and count as unchanged; those of period 3 (pulsars)
        // keep their neighbourhood listed, which 12 would avoid at the cost of
        // taller halos and more blocks. band_rows and waves do not apply.
        const int steps = c.temporal_steps ? static_cast<int>(c.temporal_steps) : kFineRingMaxSteps;
        if (steps < 1 || steps > kFineMaxSteps)
            return fail(GOL_ERR_INVALID_ARGUMENT, "temporal_steps %d: change tracking takes 1 to %d generations per block",
                        steps, kFineMaxSteps);
        const int dw = c.lane_dwords ? static_cast<int>(c.lane_dwords) : kFineDefaultDW;
        if (dw != 2 && dw != 4 && dw != 8)
            return fail(GOL_ERR_INVALID_ARGUMENT, "lane_dwords %d: change-tracking tiles are 2, 4 or 8 dwords wide",
                        dw);
        h->steps = steps;
        h->fine_geo[0] = fine_geometry(h->x_dw, h->y_rows, dw, steps);
        // With lane_dwords left to the library the tile width follows the
        // grid's activity: 64-cell tiles track sparse activity best, 128-cell
        // tiles (a fifth less VALU work per cell) win once a block lists many
        // tiles of large busy regions (enqueue_active).
        h->fine_adapt = c.lane_dwords == 0 && h->x_dw >= 2 * kFineWideDW;
        h->fine_geo[1] = fine_geometry(h->x_dw, h->y_rows, h->fine_adapt ? kFineWideDW : dw, steps);
        h->fine_sel = 0;
        h->fine = h->fine_geo[0];
        h->n_tiles = h->fine.tiles_x * h->fine.tiles_y;
        h->shape = TileShape{dw, 1, kFineWaves};
        return GOL_OK;
    }
    TileShape s = default_shape(h->x_dw, h->y_rows);
    int modelled_steps = -1;
    if (h->kernel == GOL_KERNEL_TEMPORAL) {
        TileShape m;
        int ms = 0;
        if (choose_temporal(h->x_dw, h->y_rows, static_cast<int>(c.lane_dwords), static_cast<int>(c.band_rows),
                            static_cast<int>(c.waves), static_cast<int>(c.temporal_steps),
                            !(c.flags & GOL_FLAG_NO_PERSISTENT), &m, &ms)) {
            s = m;
            modelled_steps = ms;
        }
    }
    if (c.lane_dwords) s.lane_dwords = static_cast<int>(c.lane_dwords);
    if (c.band_rows) s.band_rows = static_cast<int>(c.band_rows);
    if (c.waves) s.waves = static_cast<int>(c.waves);
    if (!shape_compiled(s))
        return fail(GOL_ERR_RUNTIME, "tile shape (lane_dwords %d, band_rows %d, waves %d) is not compiled in",
                    s.lane_dwords, s.band_rows, s.waves);
    h->shape = s;
    h->one_tile = h->x_dw <= 64 * s.lane_dwords && h->y_rows <= s.band_rows * s.waves;
    int steps = static_cast<int>(c.temporal_steps);
    if (steps == 0 && modelled_steps > 0 && !h->one_tile) steps = modelled_steps;
    if (steps == 0) steps = h->one_tile ? 1024 : 8;
    if (steps < 1 || steps > 1 << 20) return fail(GOL_ERR_INVALID_ARGUMENT, "temporal_steps out of range");
    const char* why = nullptr;
    if (!plan_tiles(s, h->x_dw, h->y_rows, h->one_tile ? 1 : steps, &h->geom, &why))
        return fail(GOL_ERR_RUNTIME, "temporal_steps %d does not fit the tile: %s", steps, why);
    h->steps = steps;
    h->n_tiles = h->geom.tiles_x * h->geom.tiles_y;
    // persistent slabs: every tile resident at once, one tile across the width,
    // interiors at least as tall as the exchanged halos
    if (h->kernel == GOL_KERNEL_TEMPORAL && !(c.flags & GOL_FLAG_NO_PERSISTENT) && !h->one_tile &&
        slab_fits(s, h->geom, h->x_dw, steps))
        h->persistent = true;
    return GOL_OK;
}

// Checks the slab error flag (a neighbour wait that ran out); waits for the stream.
int check_slab_error(gol_handle* h) {
    if (!h->slab_check) return GOL_OK;
    unsigned int err = 0;
    hipError_t e = hipMemcpyAsync(&err, h->slab_error, sizeof(err), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "slab error flag");
    h->slab_check = false;
    if (err) {
        // the block counters are left at whatever the stopped launch reached:
        // start them over, so the next launch synchronises from zero
        (void)hipMemsetAsync(h->slab_error, 0, sizeof(unsigned int), h->stream);
        (void)hipMemsetAsync(h->slab_flags, 0, h->geom.tiles_y * sizeof(unsigned long long), h->stream);
        (void)hipStreamSynchronize(h->stream);
        h->slab_epoch = 0;
        return fail(GOL_ERR_RUNTIME, "persistent launch stopped: a slab waited too long for its neighbours; "
                    "the state is undefined until the next upload (set GOL_FLAG_NO_PERSISTENT to avoid the "
                    "persistent launch)");
    }
    return GOL_OK;
}

// One cooperative launch per at most 65536 blocks: every slab stays resident
// and exchanges halos with its neighbours between blocks.
int enqueue_persistent(gol_handle* h, uint64_t generations) {
    const uint64_t T = static_cast<uint64_t>(h->steps);
    uint64_t left = generations;
    while (left > 0) {
        const uint64_t blocks = std::min<uint64_t>(left / T, 65536);
        const uint64_t last = blocks < 65536 ? left - blocks * T : 0;
        SlabArgs a;
        a.src = h->buf[h->cur];
        a.dst = h->buf[h->cur ^ 1];
        a.geo = to_device_geo(h->geom, h->x_dw, h->y_rows);
        a.steps = h->steps;
        a.blocks = static_cast<int>(blocks);
        a.last_gens = static_cast<int>(last);
        a.tiles_y = h->geom.tiles_y;
        a.xbuf = h->xbuf;
        a.flags = h->slab_flags;
        a.epoch = h->slab_epoch;
        a.error = h->slab_error;
        a.spin_limit = 0;
        a.wait_ahead = 0;
        if (h->cfg.flags & GOL_FLAG_SLAB_WAIT_PROBE) {
            // every slab waits, briefly, for a block count no slab publishes
            a.wait_ahead = 1ull << 40;
            a.spin_limit = 256;
        }
        hipError_t e = launch_slab(h->shape, a, h->stream);
        if (e != hipSuccess) return fail_hip(e, "slab_kernel launch");
        const uint64_t total = blocks + (last ? 1 : 0);
        // exchanges happen between blocks, each counted by every exporting wavefront
        h->slab_epoch += (total > 0 ? total - 1 : 0) *
                         static_cast<uint64_t>(slab_publishers(h->shape.band_rows, h->shape.waves, h->steps));
        h->cur ^= 1;
        h->launches += 1;
        const uint64_t done = blocks * T + last;
        h->generations += done;
        left -= done;
        h->slab_check = true;
    }
    return GOL_OK;
}

// One change-tracking block of h->steps generations, one launch: the tiles
// on this block's list (every tile when the lists are not valid) advance, and
// each tile that changed puts itself and the neighbours within `steps` cells
// of a changed cell on the next block's list. A skipped tile keeps the state
// it had two blocks ago in the output buffer, which equals its current state:
// no cell within reach of it changed over the last block, so the block
// starting now reproduces the same state (the tiles are at least `steps`
// cells wide and tall, the reach of `steps` generations, so that reach lies
// in the tile and its eight neighbours). Lists, marks and counters rotate
// over three slots:
// block b reads slot b % 3, appends to slot (b + 1) % 3 and resets the
// counter of slot (b + 2) % 3, which block b - 1 read.
int enqueue_active_block(gol_handle* h, unsigned long long* sample = nullptr, int sample_tag = 0) {
    const size_t n = static_cast<size_t>(h->n_tiles);
    hipError_t e;
    if (!h->flags_valid) {
        // a new state: every tile advances, and no list or mark is pending
        if ((e = hipMemsetAsync(h->marks, 0, 3 * n * sizeof(uint32_t), h->stream)) != hipSuccess ||
            (e = hipMemsetAsync(h->counts, 0, 3 * kFineCounts * sizeof(int), h->stream)) != hipSuccess)
            return fail_hip(e, "hipMemsetAsync");
    }
    const size_t k = static_cast<size_t>(h->blocks % 3), k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    const size_t list_len = kFineLists * static_cast<size_t>(fine_list_cap(h->fine));
    FineTrack tr;
    tr.mark_cur = h->marks + k * n;
    tr.mark_next = h->marks + k1 * n;
    tr.list_next = h->lists + k1 * list_len;
    tr.count_next = h->counts + k1 * kFineCounts;
    tr.count_clear = h->counts + k2 * kFineCounts;
    tr.wg_active = h->wg_active;
    tr.sample = sample;
    tr.sample_tag = sample_tag;
    e = launch_fine_tiles(h->fine, h->fbuf[h->fcur], h->fbuf[h->fcur ^ 1], h->steps, h->lists + k * list_len,
                          h->counts + k * kFineCounts, h->flags_valid, tr, h->stream);
    if (e != hipSuccess) return fail_hip(e, "fine_tile_kernel launch");
    h->flags_valid = true;
    h->fcur ^= 1;
    h->launches += 1;
    h->blocks += 1;
    h->generations += static_cast<uint64_t>(h->steps);
    return GOL_OK;
}

// Change tracking: the grid goes tile-major into fbuf[fcur] (the other
// buffer keeps the tiles the last block did not write, which equal their
// state then), blocks of `steps` generations run on the tile-major buffers,
// and the result comes back row-major into buf[cur]. The buffers never hold
// another state between steps: uploads reset the lists, so the first block
// after one writes every tile.
// The length of the list chunk slot buf's sample block advanced, once every
// sub-list's word carries the block's tag; -1 when no sample is pending. The
// block is queued chunks ago, so this is a short wait on the GPU's progress;
// should the words not arrive while the stream has gone idle, or after a long
// wait, the stream is checked for an error and the sample is skipped (the
// width stays).
int fine_read_sample(gol_handle* h, int buf, int64_t* listed) {
    *listed = -1;
    const unsigned long long* words = h->fine_sample + buf * kFineLists;
    const unsigned tag = static_cast<unsigned>(h->fine_sample_tag[buf]);
    h->fine_sample_tag[buf] = 0;
    if (tag == 0) return GOL_OK;
    auto read = [&]() {
        int64_t sum = 0;
        for (int j = 0; j < kFineLists; ++j) {
            const unsigned long long w = __atomic_load_n(words + j, __ATOMIC_ACQUIRE);
            if (static_cast<unsigned>(w >> 32) != tag) return static_cast<int64_t>(-1);
            sum += static_cast<unsigned>(w);
        }
        return sum;
    };
    const auto start = std::chrono::steady_clock::now();
    for (uint64_t i = 0;; ++i) {
        if ((*listed = read()) >= 0) return GOL_OK;
        if ((i & 1023) == 1023) {
            const hipError_t q = hipStreamQuery(h->stream);
            if (q != hipSuccess && q != hipErrorNotReady) return fail_hip(q, "change-tracking list sample");
            if (q == hipSuccess) {  // idle: the words are final
                *listed = read();
                return GOL_OK;
            }
            if (std::chrono::steady_clock::now() - start > std::chrono::seconds(10)) return GOL_OK;
        }
        std::this_thread::yield();
    }
}

// Moves the state to the other tile geometry: back to rows, then tile-major
// in the other width; the next block advances every tile.
int fine_switch_width(gol_handle* h) {
    hipError_t e = launch_fine_layout(h->fine, h->fbuf[h->fcur], h->buf[h->cur], false, h->stream);
    if (e != hipSuccess) return fail_hip(e, "change-tracking layout launch");
    h->fine_sel ^= 1;
    h->fine = h->fine_geo[h->fine_sel];
    h->n_tiles = h->fine.tiles_x * h->fine.tiles_y;
    if ((e = launch_fine_layout(h->fine, h->buf[h->cur], h->fbuf[h->fcur], true, h->stream)) != hipSuccess)
        return fail_hip(e, "change-tracking layout launch");
    h->launches += 2;
    h->flags_valid = false;
    h->fine_switches += 1;
    return GOL_OK;
}

int enqueue_active(gol_handle* h, uint64_t generations) {
    if (generations == 0) return GOL_OK;
    hipError_t e = launch_fine_layout(h->fine, h->buf[h->cur], h->fbuf[h->fcur], true, h->stream);
    if (e != hipSuccess) return fail_hip(e, "change-tracking layout launch");
    h->launches += 1;
    uint64_t left = generations;
    // Blocks go out in chunks of kFineChunk. With an adapting width, the
    // last block of each chunk writes the length of the list it advances to
    // host memory, and before chunk c the host reads chunk c - kFineLag's
    // (waiting for it leaves the chunks between queued on the GPU), and
    // switches the width when the list is long for 64-cell tiles or short for
    // 128-cell ones. Every result is exact either way; only the time depends
    // on it.
    for (uint64_t chunk = 0; left >= static_cast<uint64_t>(h->steps); ++chunk) {
        const int buf = static_cast<int>(chunk % kFineLag);
        if (h->fine_adapt && chunk >= static_cast<uint64_t>(kFineLag)) {
            int64_t listed;
            int rc = fine_read_sample(h, buf, &listed);
            if (rc) return rc;
            const int64_t tiles = static_cast<int64_t>(h->n_tiles);
            const bool sampled_here = h->fine_sample_sel[buf] == h->fine_sel;
            if (listed >= 0 && sampled_here && h->flags_valid &&
                ((h->fine_sel == 0 && listed * 100 > tiles * kFineWidePercent) ||
                 (h->fine_sel == 1 && listed * 100 < tiles * kFineNarrowPercent))) {
                int rc = fine_switch_width(h);
                if (rc) return rc;
            }
        }
        bool sampled = false;
        for (int b = 0; b < kFineChunk && left >= static_cast<uint64_t>(h->steps); ++b) {
            // the chunk's last block writes its list's length to host memory
            const bool last = b == kFineChunk - 1 || left < 2 * static_cast<uint64_t>(h->steps);
            const bool sample = h->fine_adapt && last && h->flags_valid;
            int tag = 0;
            if (sample) {
                // tags do not repeat within 10^9 samples: a word left from an
                // earlier sample never passes for this one
                h->fine_tags = h->fine_tags % 1000000000 + 1;
                tag = h->fine_tags;
            }
            int rc = enqueue_active_block(h, sample ? h->fine_sample_dev + buf * kFineLists : nullptr, tag);
            if (rc) return rc;
            if (sample) {
                sampled = true;
                h->fine_sample_tag[buf] = tag;
                h->fine_sample_sel[buf] = h->fine_sel;
            }
            left -= static_cast<uint64_t>(h->steps);
        }
        if (!sampled) h->fine_sample_tag[buf] = 0;
    }
    if (left > 0) {
        // a shorter block runs every tile: an oscillator that returns to its
        // state after `steps` generations need not after fewer
        h->flags_valid = false;
        e = launch_fine_tiles(h->fine, h->fbuf[h->fcur], h->fbuf[h->fcur ^ 1], static_cast<int>(left), h->lists,
                              h->counts, false, FineTrack{}, h->stream);
        if (e != hipSuccess) return fail_hip(e, "fine_tile_kernel launch");
        h->fcur ^= 1;
        h->launches += 1;
        h->generations += left;
    }
    e = launch_fine_layout(h->fine, h->fbuf[h->fcur], h->buf[h->cur], false, h->stream);
    if (e != hipSuccess) return fail_hip(e, "change-tracking layout launch");
    h->launches += 1;
    return GOL_OK;
}

int enqueue(gol_handle* h, uint64_t generations) {
    if (h->persistent) return enqueue_persistent(h, generations);
    uint64_t left = generations;
    if (h->native) {
        while (left > 0) {
            const int gens = static_cast<int>(std::min<uint64_t>(left, static_cast<uint64_t>(h->steps)));
            const hipError_t e = h->native == GOL_LAYOUT_TILES
                                     ? launch_tile8(h->t8, h->buf[h->cur], h->buf[h->cur ^ 1], gens, h->stream)
                                     : launch_columns(h->cols, h->buf[h->cur], h->buf[h->cur ^ 1], gens, h->stream);
            if (e != hipSuccess) return fail_hip(e, "native layout kernel launch");
            h->cur ^= 1;
            h->launches += 1;
            h->generations += static_cast<uint64_t>(gens);
            left -= static_cast<uint64_t>(gens);
        }
        return GOL_OK;
    }
    if (h->kernel == GOL_KERNEL_ACTIVE) return enqueue_active(h, generations);
    while (left > 0) {
        uint32_t* src = h->buf[h->cur];
        uint32_t* dst = h->buf[h->cur ^ 1];
        hipError_t e;
        uint64_t done;
        if (h->kernel == GOL_KERNEL_PLAIN) {
            e = launch_plain(static_cast<int>(h->plain_dw), src, dst, h->x_dw, h->y_rows, h->stream);
            done = 1;
        } else {
            done = std::min<uint64_t>(left, static_cast<uint64_t>(h->steps));
            e = launch_tiles(h->geom, src, dst, h->x_dw, h->y_rows, static_cast<int>(done), h->stream);
        }
        if (e != hipSuccess) return fail_hip(e, "kernel launch");
        h->cur ^= 1;
        h->launches += 1;
        h->generations += done;
        left -= done;
    }
    return GOL_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int now = -1;
        if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

int gol_create(const gol_config* config, gol_handle** out) {
    clear_error();
    if (!out) return fail(GOL_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int rc = validate_config(config);
    if (rc) return rc;
    int dev = config->device;
    if (dev < 0) {
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return fail_hip(e, "hipGetDeviceCount");
    if (dev >= count) return fail(GOL_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", dev, count);

    gol_handle* h = new gol_handle();
    h->cfg = *config;
    h->device = dev;
    h->x_dw = static_cast<int64_t>(config->x_cells / 32);
    h->y_rows = static_cast<int64_t>(config->y_cells);
    h->n_dwords = static_cast<size_t>(h->x_dw) * static_cast<size_t>(h->y_rows);
    DeviceGuard g(dev);
    rc = plan(h);
    if (rc) {
        delete h;
        return rc;
    }
    const size_t bytes = h->n_dwords * sizeof(uint32_t);
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&h->ev_start)) != hipSuccess || (e = hipEventCreate(&h->ev_stop)) != hipSuccess ||
        (e = hipMalloc(&h->buf[0], bytes)) != hipSuccess || (e = hipMalloc(&h->buf[1], bytes)) != hipSuccess ||
        (e = hipMemsetAsync(h->buf[0], 0, bytes, h->stream)) != hipSuccess ||
        (h->cfg.layout != GOL_LAYOUT_ROWS && !h->native && (e = hipMalloc(&h->staging, bytes)) != hipSuccess)) {
        rc = fail_hip(e, "gol_create");
        gol_destroy(h);
        return rc;
    }
    if (h->kernel == GOL_KERNEL_ACTIVE) {
        // sized for the larger of the two tile geometries
        const FineGeo& g0 = h->fine_geo[0];
        const FineGeo& g1 = h->fine_geo[1];
        const size_t nt = std::max<size_t>(static_cast<size_t>(g0.tiles_x) * g0.tiles_y,
                                           static_cast<size_t>(g1.tiles_x) * g1.tiles_y);
        const size_t nwg = static_cast<size_t>(std::max(fine_workgroups(g0), fine_workgroups(g1)));
        const size_t list_len = kFineLists * static_cast<size_t>(std::max(fine_list_cap(g0), fine_list_cap(g1)));
        const size_t fwords = std::max(fine_words(g0), fine_words(g1));
        if ((e = hipMalloc(&h->marks, 3 * nt * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMalloc(&h->lists, 3 * list_len * sizeof(int))) != hipSuccess ||
            (e = hipMalloc(&h->counts, 3 * kFineCounts * sizeof(int))) != hipSuccess ||
            (e = hipMalloc(&h->wg_active, nwg * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMemsetAsync(h->wg_active, 0, nwg * sizeof(unsigned long long), h->stream)) != hipSuccess ||
            (e = hipMalloc(&h->fbuf[0], fwords * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMalloc(&h->fbuf[1], fwords * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMemsetAsync(h->fbuf[0], 0, fwords * sizeof(uint32_t), h->stream)) != hipSuccess ||
            (e = hipMemsetAsync(h->fbuf[1], 0, fwords * sizeof(uint32_t), h->stream)) != hipSuccess ||
            (h->fine_adapt &&
             ((e = hipHostMalloc(reinterpret_cast<void**>(&h->fine_sample),
                                 kFineLag * kFineLists * sizeof(unsigned long long),
                                 hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess ||
              (e = hipHostGetDevicePointer(reinterpret_cast<void**>(&h->fine_sample_dev), h->fine_sample, 0)) !=
                  hipSuccess))) {
            rc = fail_hip(e, "gol_create (change tracking state)");
            gol_destroy(h);
            return rc;
        }
    }
    if (h->persistent) {
        const size_t pairs = 2ull * h->geom.tiles_y * 2 * h->steps * 32 * static_cast<size_t>(h->shape.lane_dwords);
        if ((e = hipMalloc(&h->xbuf, pairs * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMalloc(&h->slab_flags, h->geom.tiles_y * sizeof(unsigned long long))) != hipSuccess ||
            (e = hipMalloc(&h->slab_error, sizeof(unsigned int))) != hipSuccess ||
            (e = hipMemsetAsync(h->slab_flags, 0, h->geom.tiles_y * sizeof(unsigned long long), h->stream)) !=
                hipSuccess ||
            (e = hipMemsetAsync(h->slab_error, 0, sizeof(unsigned int), h->stream)) != hipSuccess) {
            rc = fail_hip(e, "gol_create (persistent slab state)");
            gol_destroy(h);
            return rc;
        }
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) {
        rc = fail_hip(e, "gol_create");
        gol_destroy(h);
        return rc;
    }
    *out = h;
    return GOL_OK;
}

// A new state: every tile is active in the next block, which (change
// tracking with an adapting width) starts on 64-cell tiles again.
void new_state(gol_handle* h) {
    h->flags_valid = false;
    if (h->kernel == GOL_KERNEL_ACTIVE && h->fine_sel != 0) {
        h->fine_sel = 0;
        h->fine = h->fine_geo[0];
        h->n_tiles = h->fine.tiles_x * h->fine.tiles_y;
    }
}

int gol_upload(gol_handle* h, const void* words, size_t n_words) {
    clear_error();
    if (!h || !words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    const size_t want = h->n_dwords * 32 / h->cfg.word_bits;
    if (n_words != want) return fail(GOL_ERR_INVALID_ARGUMENT, "expected %zu words, got %zu", want, n_words);
    DeviceGuard g(h->device);
    hipError_t e;
    if (h->cfg.layout != GOL_LAYOUT_ROWS && !h->native) {
        if ((e = hipMemcpyAsync(h->staging, words, h->n_dwords * 4, hipMemcpyHostToDevice, h->stream)) == hipSuccess)
            e = layout_to_rows(static_cast<int>(h->cfg.layout), static_cast<int>(h->cfg.word_bits), h->staging,
                               h->buf[h->cur], h->x_dw * 32, h->y_rows, h->stream);
    } else {
        e = hipMemcpyAsync(h->buf[h->cur], words, h->n_dwords * 4, hipMemcpyHostToDevice, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "gol_upload");
    new_state(h);
    return GOL_OK;
}

int gol_upload_device(gol_handle* h, const void* device_words, size_t n_words) {
    clear_error();
    if (!h || !device_words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    const size_t want = h->n_dwords * 32 / h->cfg.word_bits;
    if (n_words != want) return fail(GOL_ERR_INVALID_ARGUMENT, "expected %zu words, got %zu", want, n_words);
    DeviceGuard g(h->device);
    hipError_t e =
        h->cfg.layout != GOL_LAYOUT_ROWS && !h->native
            ? layout_to_rows(static_cast<int>(h->cfg.layout), static_cast<int>(h->cfg.word_bits), device_words,
                             h->buf[h->cur], h->x_dw * 32, h->y_rows, h->stream)
            : hipMemcpyAsync(h->buf[h->cur], device_words, h->n_dwords * 4, hipMemcpyDeviceToDevice, h->stream);
    if (e != hipSuccess) return fail_hip(e, "gol_upload_device");
    new_state(h);
    return GOL_OK;
}

int gol_step(gol_handle* h, uint64_t generations) {
    clear_error();
    if (!h) return fail(GOL_ERR_INVALID_ARGUMENT, "null handle");
    DeviceGuard g(h->device);
    return enqueue(h, generations);
}

int gol_step_timed(gol_handle* h, uint64_t generations, double* gpu_ms) {
    clear_error();
    if (!h || !gpu_ms) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard g(h->device);
    hipError_t e = hipEventRecord(h->ev_start, h->stream);
    if (e != hipSuccess) return fail_hip(e, "hipEventRecord");
    int rc = enqueue(h, generations);
    if (rc) return rc;
    if ((e = hipEventRecord(h->ev_stop, h->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = hipEventSynchronize(h->ev_stop)) != hipSuccess) return fail_hip(e, "hipEventSynchronize");
    if ((rc = check_slab_error(h))) return rc;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, h->ev_start, h->ev_stop)) != hipSuccess)
        return fail_hip(e, "hipEventElapsedTime");
    *gpu_ms = ms;
    return GOL_OK;
}

int gol_download(gol_handle* h, void* words, size_t n_words) {
    clear_error();
    if (!h || !words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    const size_t want = h->n_dwords * 32 / h->cfg.word_bits;
    if (n_words != want) return fail(GOL_ERR_INVALID_ARGUMENT, "expected %zu words, got %zu", want, n_words);
    DeviceGuard g(h->device);
    int rc = check_slab_error(h);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    const void* from = h->buf[h->cur];
    if (h->cfg.layout != GOL_LAYOUT_ROWS && !h->native) {
        e = rows_to_layout(static_cast<int>(h->cfg.layout), static_cast<int>(h->cfg.word_bits), h->buf[h->cur],
                           h->staging, h->x_dw * 32, h->y_rows, h->stream);
        from = h->staging;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(words, from, h->n_dwords * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "gol_download");
    return GOL_OK;
}

int gol_sync(gol_handle* h) {
    clear_error();
    if (!h) return fail(GOL_ERR_INVALID_ARGUMENT, "null handle");
    DeviceGuard g(h->device);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_hip(e, "gol_sync");
    return check_slab_error(h);
}

int gol_device_state(gol_handle* h, void** device_words) {
    clear_error();
    if (!h || !device_words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    *device_words = h->buf[h->cur];
    return GOL_OK;
}

int gol_get_stream(gol_handle* h, void** stream) {
    clear_error();
    if (!h || !stream) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    *stream = h->stream;
    return GOL_OK;
}

int gol_get_info(gol_handle* h, gol_info* info) {
    clear_error();
    if (!h || !info) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->kernel = h->kernel;
    if (h->native) {
        // workgroups (tiles) and wavefronts of one launch
        info->temporal_steps = static_cast<uint32_t>(h->steps);
        info->tiles = static_cast<uint32_t>(h->n_tiles);
        info->waves = static_cast<uint32_t>(h->native == GOL_LAYOUT_TILES ? kTile8Waves : 1);
        info->band_rows = static_cast<uint32_t>(h->native == GOL_LAYOUT_TILES ? kTile8Rows : kColDwords);
    } else if (h->kernel == GOL_KERNEL_PLAIN) {
        info->lane_dwords = h->plain_dw;
        info->temporal_steps = 1;
    } else {
        info->lane_dwords = static_cast<uint32_t>(h->kernel == GOL_KERNEL_ACTIVE ? h->fine.tile_dwords
                                                                                 : h->shape.lane_dwords);
        info->band_rows = static_cast<uint32_t>(h->shape.band_rows);
        info->waves = static_cast<uint32_t>(h->shape.waves);
        info->temporal_steps = static_cast<uint32_t>(h->steps);
        info->tiles = static_cast<uint32_t>(h->kernel == GOL_KERNEL_ACTIVE ? h->n_tiles
                                                                           : h->geom.tiles_x * h->geom.tiles_y);
        info->persistent = h->persistent ? 1u : 0u;
    }
    info->launches = h->launches;
    info->generations = h->generations;
    info->device_bytes = (h->staging ? 3 : 2) * h->n_dwords * 4;
    info->layout = h->cfg.layout;
    info->native_layout = h->native ? 1u : 0u;
    if (h->kernel == GOL_KERNEL_ACTIVE) {
        DeviceGuard g(h->device);
        std::vector<unsigned long long> per_wg(
            static_cast<size_t>(std::max(fine_workgroups(h->fine_geo[0]), fine_workgroups(h->fine_geo[1]))));
        hipError_t e = hipMemcpyAsync(per_wg.data(), h->wg_active, per_wg.size() * sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail_hip(e, "gol_get_info");
        unsigned long long total = 0;
        for (unsigned long long v : per_wg) total += v;
        info->active_tiles_total = total;
        info->tracked_blocks = h->blocks;
        info->tile_width_switches = static_cast<uint32_t>(h->fine_switches);
        info->workgroups = static_cast<uint32_t>(per_wg.size());
        // as allocated: each buffer for the larger of the two geometries
        const FineGeo& g0 = h->fine_geo[0];
        const FineGeo& g1 = h->fine_geo[1];
        const uint64_t nt = std::max<uint64_t>(static_cast<uint64_t>(g0.tiles_x) * g0.tiles_y,
                                               static_cast<uint64_t>(g1.tiles_x) * g1.tiles_y);
        info->device_bytes += nt * 3 * sizeof(uint32_t) +
                              3ull * kFineLists * std::max(fine_list_cap(g0), fine_list_cap(g1)) * sizeof(int) +
                              2ull * std::max(fine_words(g0), fine_words(g1)) * sizeof(uint32_t) +
                              3ull * kFineCounts * sizeof(int) + per_wg.size() * sizeof(unsigned long long);
    }
    return GOL_OK;
}

void gol_destroy(gol_handle* h) {
    if (!h) return;
    {
        DeviceGuard g(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (h->buf[0]) (void)hipFree(h->buf[0]);
        if (h->buf[1]) (void)hipFree(h->buf[1]);
        if (h->staging) (void)hipFree(h->staging);
        if (h->marks) (void)hipFree(h->marks);
        if (h->lists) (void)hipFree(h->lists);
        if (h->counts) (void)hipFree(h->counts);
        if (h->wg_active) (void)hipFree(h->wg_active);
        if (h->fbuf[0]) (void)hipFree(h->fbuf[0]);
        if (h->fbuf[1]) (void)hipFree(h->fbuf[1]);
        if (h->fine_sample) (void)hipHostFree(h->fine_sample);
        if (h->xbuf) (void)hipFree(h->xbuf);
        if (h->slab_flags) (void)hipFree(h->slab_flags);
        if (h->slab_error) (void)hipFree(h->slab_error);
        if (h->ev_start) (void)hipEventDestroy(h->ev_start);
        if (h->ev_stop) (void)hipEventDestroy(h->ev_stop);
        if (h->stream) (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

const char* gol_last_error(void) { return last_error(); }

const char* gol_version(void) { return "golhip 0.1 (gfx950)"; }

}  // extern "C"
