// Batched execution of many independent grids (no reference counterpart: the
// reference runs one grid per process, src/infrastructure/experiment_manager.hpp:235-270).
//
// Every grid gets a tile shape from its size class (a measured choice from
// gol_batch_autotune, else the built-in table and cost model) and its own
// ping-pong buffers in one
// device arena. Work is cut into rounds; in each round every unfinished grid
// contributes one work item per tile, advancing that tile by the grid's
// per-round generation count:
//   * grids split into several tiles advance `round_generations` (the tile
//     halo) per round, so neighbouring tiles exchange halos through the
//     grid's buffers between rounds;
//   * grids that fit one tile need no halo; they advance a chunk sized so the
//     item costs about as much as a large tile's item, which keeps the items
//     of a round balanced. Grids at most 512 cells wide and 512 rows tall run
//     on the narrow kernel (one row per lane, the whole grid in registers).
// A round is one launch per launch key present (a tile shape, or a row width
// of narrow grids), one workgroup per item, items ordered largest first. The
// launch keys are spread over four streams (lanes) by total work, largest
// first onto the least loaded lane; a key's launches stay on its lane, so the
// lanes run their rounds independently between one fork and one join, and a
// small shape fills the CUs a large one leaves idle. All rounds are recorded
// once into a hipGraph and replayed, so a whole batch is one graph launch.
//
// With GOL_BATCH_PERSISTENT the same items run as one persistent launch per
// launch key instead (plan_queues below): its workgroups drain the key's
// queue, and a split grid's round waits on the device for the grid's previous
// round. The results are the same bit for bit; on the batched workload the
// per-round graph is faster, so it stays the default.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/golhip.h"
#include "gol_capi_internal.hpp"
#include "gol_kernels.hpp"

using namespace golhip;

namespace golhip {

// Row budget of a round's work item, in tile rows x generations: a split
// grid's item is (tile rows) x round_generations; one-tile grids get as many
// generations per round as make the same product.
// A narrow item's wavefronts each advance their rows' lane_dwords dwords plus
// about two dwords' worth of edge exchange per generation, in parallel; it is
// charged as a tile of 16 rows per dword of that sequential work, so its
// latency stays near a large tile's item (one grid of few wavefronts must not
// hold up a round).
static int chunk_generations(const TileShape& s, int round_generations, int big_tile_rows) {
    const int rows = is_narrow(s) ? 16 * (s.lane_dwords + 2) : s.band_rows * s.waves;
    const int64_t budget = static_cast<int64_t>(round_generations) * big_tile_rows;
    return static_cast<int>(std::max<int64_t>(round_generations, budget / rows));
}

TileShape batch_default_shape(int64_t x_dw, int64_t y_rows) {
    if (x_dw <= 64) {
        if (y_rows <= 64) return TileShape{1, 8, 8};
        if (y_rows <= 128) return TileShape{1, 16, 8};
        return TileShape{1, 32, 16};
    }
    return default_shape(x_dw, y_rows);
}

// Size classes of the measured autotuner: power-of-two classes of the row
// width in dwords and of the row count.
struct SizeClass {
    int x_class, y_class;
    bool operator<(const SizeClass& o) const { return x_class != o.x_class ? x_class < o.x_class : y_class < o.y_class; }
};
static int ceil_log2(int64_t v) {
    int c = 0;
    while ((int64_t{1} << c) < v) ++c;
    return c;
}
static SizeClass size_class(int64_t x_dw, int64_t y_rows) { return SizeClass{ceil_log2(x_dw), ceil_log2(y_rows)}; }

// A batch choice: tile shape and, for grids split into tiles, generations per
// round (the halo depth; one-tile grids advance a chunk instead).
struct BatchChoice {
    TileShape shape;
    int steps;
};

static std::mutex g_tune_mu;
static std::map<SizeClass, BatchChoice> g_tuned;  // process-wide measured choices
// A choice the autotuner is measuring; overrides everything in this thread.
static thread_local const BatchChoice* t_forced = nullptr;
// Set while the autotuner measures the built-in choice: measured choices are ignored.
static thread_local bool t_untuned = false;

static bool tuned_choice(int64_t x_dw, int64_t y_rows, BatchChoice* c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(size_class(x_dw, y_rows));
    if (it == g_tuned.end()) return false;
    *c = it->second;
    return true;
}

TileShape batch_shape(int64_t x_dw, int64_t y_rows) {
    BatchChoice c;
    if (!t_untuned && tuned_choice(x_dw, y_rows, &c)) return c.shape;
    return batch_default_shape(x_dw, y_rows);
}

// Cost per generation of advancing a whole grid split into tiles, in CU
// microseconds, when many grids share the GPU (no wave quantization): the
// compute of every tile row (about 12.7 + 8 / band_rows VALU ops per dword at
// 64 lanes per clock, 2.4 GHz) and, once per round, loading each tile with its
// halo and storing its interior (about 5 TB/s shared by 256 CUs). With
// `occupancy` of two or more workgroups per CU, one tile's loads and stores
// overlap another's compute (1: priced on top of compute).
static double split_cost_per_generation(const TileShape& s, const TileGeometry& g, int steps, int occupancy) {
    const double tiles = static_cast<double>(g.tiles_x) * g.tiles_y;
    const double width = 64.0 * s.lane_dwords;
    const double compute = tiles * g.tile_rows * width * (12.7 + 8.0 / s.band_rows) / 64.0 / 2400.0;
    const double memory =
        tiles * (static_cast<double>(g.tile_rows) + g.stride_rows) * width * 4.0 * 256.0 / 5.0e6 / steps;
    return occupancy >= 2 ? std::max(compute, memory) : compute + memory;
}

// The given choice, if it applies to the grid: one-tile shapes advance a
// chunk per round, split shapes `steps` (<= round_generations) generations.
static bool plan_choice(const BatchChoice& c, int64_t x_dw, int64_t y_rows, int round_generations, int big_tile_rows,
                        TileShape* shape, TileGeometry* geom, int* per_round) {
    const char* why = nullptr;
    const TileShape& s = c.shape;
    if (is_narrow(s)) {
        // a measured narrow choice means the narrow kernel, sized for this grid
        TileShape n;
        if (!narrow_shape_for(x_dw, y_rows, &n)) return false;
        *shape = n;
        *geom = narrow_geometry(n, x_dw, y_rows);
        *per_round = chunk_generations(n, round_generations, big_tile_rows);
        return true;
    }
    if (!shape_compiled(s)) return false;
    if (x_dw <= 64 * s.lane_dwords && y_rows <= s.band_rows * s.waves) {
        if (!plan_tiles(s, x_dw, y_rows, 1, geom, &why)) return false;
        *shape = s;
        *per_round = chunk_generations(s, round_generations, big_tile_rows);
        return true;
    }
    const int steps = std::min(c.steps > 0 ? c.steps : round_generations, round_generations);
    if (!plan_tiles(s, x_dw, y_rows, steps, geom, &why)) return false;
    *shape = s;
    *per_round = steps;
    return true;
}

int plan_batch_grid(int64_t x_dw, int64_t y_rows, int round_generations, Planner planner, TileShape* shape,
                    TileGeometry* geom, int* per_round) {
    const bool use_tuned = planner == Planner::kBatch;
    constexpr int kBigTileRows = 512;
    const char* why = nullptr;
    // a choice being measured, then a measured choice for the size class
    if (t_forced) {
        if (plan_choice(*t_forced, x_dw, y_rows, round_generations, kBigTileRows, shape, geom, per_round))
            return GOL_OK;
        return fail(GOL_ERR_RUNTIME, "tile shape under test does not fit the grid");
    }
    const bool tuned_ok = use_tuned && !t_untuned;
    BatchChoice tuned;
    if (tuned_ok && tuned_choice(x_dw, y_rows, &tuned) &&
        plan_choice(tuned, x_dw, y_rows, round_generations, kBigTileRows, shape, geom, per_round))
        return GOL_OK;
    // narrow grids: one row per lane, no halo
    TileShape narrow;
    if (narrow_shape_for(x_dw, y_rows, &narrow)) {
        *shape = narrow;
        *geom = narrow_geometry(narrow, x_dw, y_rows);
        *per_round = chunk_generations(narrow, round_generations, kBigTileRows);
        return GOL_OK;
    }
    // grids that fit one tile of their size class need no halo
    const TileShape table = tuned_ok ? batch_shape(x_dw, y_rows) : batch_default_shape(x_dw, y_rows);
    if (x_dw <= 64 * table.lane_dwords && y_rows <= table.band_rows * table.waves) {
        *shape = table;
        *per_round = chunk_generations(table, round_generations, kBigTileRows);
        if (!plan_tiles(table, x_dw, y_rows, 1, geom, &why))
            return fail(GOL_ERR_RUNTIME, "grid %lldx%lld: %s", static_cast<long long>(x_dw * 32),
                        static_cast<long long>(y_rows), why);
        return GOL_OK;
    }
    // split grids: the compiled shape and halo depth (<= round_generations)
    // with the lowest cost per generation
    double best = 1e300;
    for (int i = 0; i < kNumTileShapes; ++i) {
        const TileShape& s = kTileShapes[i];
        const int occupancy = tile_occupancy(s);
        // a server keeps rounds at least half as deep as round_generations
        const int min_steps = planner == Planner::kServer ? std::max(1, round_generations / 2) : 1;
        for (int steps = round_generations; steps >= min_steps; --steps) {
            TileGeometry g;
            if (!plan_tiles(s, x_dw, y_rows, steps, &g, &why)) continue;
            const double c = split_cost_per_generation(s, g, steps, occupancy);
            if (c < best) {
                best = c;
                *shape = s;
                *geom = g;
                *per_round = steps;
            }
        }
    }
    if (best >= 1e300)
        return fail(GOL_ERR_RUNTIME, "grid %lldx%lld: no compiled tile shape fits", static_cast<long long>(x_dw * 32),
                    static_cast<long long>(y_rows));
    return GOL_OK;
}

}  // namespace golhip

struct BatchEntry {
    int64_t x_dw, y_rows;
    uint32_t word_bits;
    uint64_t generations;
    std::vector<uint32_t> host;  // initial state until uploaded
    size_t offset = 0;           // dword offset of buf0 in the arena (buf1, init follow)
    TileShape shape{1, 16, 8};
    TileGeometry geom{};
    int chunk = 0;               // generations per round
    int final_parity = 0;        // buffer holding the result after a run of the current plan
    int result_parity = 0;       // buffer holding the result of the last completed run
};

struct gol_batch {
    gol_batch_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;  // lane 0
    static constexpr int kLanes = 4;
    hipStream_t side[kLanes - 1] = {};  // lanes 1..3
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kLanes - 1] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<BatchEntry> grids;
    bool built = false;
    uint32_t* arena = nullptr;
    size_t arena_dwords = 0;
    BatchGrid* d_grids = nullptr;
    WorkItem* d_items = nullptr;
    struct Launch {
        TileShape shape;
        size_t first;  // index into the item array
        int count;
        int lane;
        int workgroups = 0;  // persistent queue launch: its workgroups (0: one workgroup per item)
    };
    std::vector<Launch> launches;
    // persistent queue launches (GOL_BATCH_PERSISTENT): one per launch key;
    // qstate holds each launch's ticket counter, then each grid's completed
    // items, then the error word
    bool queue = false;
    unsigned int* d_qstate = nullptr;
    int qstate_words = 0;
    unsigned long long wait_ticks = 0;  // persistent queue: device wall-clock ticks a wait may last
    bool dirty = false;                // a run changed the buffers since the initial states were restored
    bool result_parity_valid = false;  // a run completed since then: BatchEntry::result_parity holds
    // GOL_BATCH_SELECT_PLAN: the two plans' timed runs and which one was kept
    double plan_tuned_ms = 0.0, plan_builtin_ms = 0.0;
    bool plan_builtin = false;
    uint32_t rounds = 0;
    uint64_t items = 0;
    uint64_t cell_updates = 0;
    uint64_t runs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

struct Guard {
    int prev = -1;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~Guard() {
        int now = -1;
        if (prev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev) (void)hipSetDevice(prev);
    }
};

// Generations one item of a one-tile grid advances in a persistent queue
// launch: such a grid needs no halo, so one item runs all its generations
// (longer runs are split into items that wait for each other).
constexpr uint64_t kQueueItemGenerations = uint64_t{1} << 20;

// Persistent queue launches (GOL_BATCH_PERSISTENT): one per launch key. A grid
// that one tile spans is one item; a split grid is rounds of one item per
// tile, round r of the grid waiting for the grid's r x tiles earlier items.
// Each queue holds its items round by round, so an item always follows those
// it waits for, the most costly first within a round. The launches go onto the
// lanes smallest total work first, each with as many workgroups as can be
// resident, capped at the most items of its queue that can run at once (the
// largest round).
int plan_queues(gol_batch* b, const std::map<LaunchKey, int>& lane_of, std::vector<int>& parity,
                std::vector<WorkItem>* all) {
    struct Entry {
        int round;
        double cost;
        WorkItem item;
    };
    struct Queue {
        std::vector<Entry> entries;
        std::map<int, int> per_round;  // items of each round
        int waves = 0;                 // narrow grids: the tallest grid's wavefronts
        double work = 0.0;
    };
    std::map<LaunchKey, Queue> queues;
    for (size_t i = 0; i < b->grids.size(); ++i) {
        auto& g = b->grids[i];
        const LaunchKey key = launch_key(g.shape);
        Queue& q = queues[key];
        q.waves = std::max(q.waves, g.shape.waves);
        const int tiles = g.geom.tiles_x * g.geom.tiles_y;
        const uint64_t step = tiles == 1 ? kQueueItemGenerations : static_cast<uint64_t>(g.chunk);
        const uint64_t cells = static_cast<uint64_t>(g.x_dw * 32 * g.y_rows);
        uint64_t left = g.generations;
        for (int round = 0; left > 0; ++round) {
            const int gens = static_cast<int>(std::min(left, step));
            for (int t = 0; t < tiles; ++t)
                q.entries.push_back(Entry{round, static_cast<double>(cells) / tiles * gens,
                                          WorkItem{static_cast<int>(i), t, gens, parity[i], round * tiles}});
            q.per_round[round] += tiles;
            parity[i] ^= 1;
            left -= static_cast<uint64_t>(gens);
            b->cell_updates += cells * static_cast<uint64_t>(gens);
            b->rounds = std::max(b->rounds, static_cast<uint32_t>(round + 1));
        }
        q.work += static_cast<double>(cells) * static_cast<double>(g.generations);
    }
    std::vector<std::pair<double, LaunchKey>> order;
    for (const auto& kv : queues) order.push_back({kv.second.work, kv.first});
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, LaunchKey>& a, const std::pair<double, LaunchKey>& c) {
                         return a.first < c.first;
                     });
    for (const auto& o : order) {
        Queue& q = queues[o.second];
        if (q.entries.empty()) continue;  // every grid of this key has 0 generations
        std::stable_sort(q.entries.begin(), q.entries.end(), [](const Entry& a, const Entry& c) {
            return a.round != c.round ? a.round < c.round : a.cost > c.cost;
        });
        const LaunchKey& key = o.second;
        gol_batch::Launch L;
        L.shape = TileShape{key.lane_dwords, key.band_rows, key.band_rows ? key.waves : q.waves};
        L.first = all->size();
        L.count = static_cast<int>(q.entries.size());
        L.lane = lane_of.at(key);
        int parallel = 0;
        for (const auto& pr : q.per_round) parallel = std::max(parallel, pr.second);
        const int capacity = queue_capacity(L.shape);
        if (capacity <= 0)
            return fail(GOL_ERR_RUNTIME, "the persistent batch kernel for shape (%d, %d, %d) cannot be resident",
                        L.shape.lane_dwords, L.shape.band_rows, L.shape.waves);
        L.workgroups = std::min(capacity, parallel);
        for (const auto& en : q.entries) all->push_back(en.item);
        b->launches.push_back(L);
    }
    return GOL_OK;
}

// Allocate the arena (every grid's two buffers and its initial copy) and
// upload the initial states; the host copies are released afterwards.
int build_arena(gol_batch* b) {
    size_t off = 0;
    for (auto& g : b->grids) {
        g.offset = off;
        const size_t n = static_cast<size_t>(g.x_dw * g.y_rows);
        off += 3 * ((n + 63) & ~size_t(63));  // buf0, buf1, initial copy; 256-byte aligned
    }
    b->arena_dwords = off;
    hipError_t e;
    if (off > 0 && (e = hipMalloc(&b->arena, off * 4)) != hipSuccess) return fail_hip(e, "batch arena");
    for (auto& g : b->grids) {
        const size_t n = static_cast<size_t>(g.x_dw * g.y_rows);
        const size_t stride = (n + 63) & ~size_t(63);
        if ((e = hipMemcpyAsync(b->arena + g.offset + 2 * stride, g.host.data(), n * 4, hipMemcpyHostToDevice,
                                b->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(b->arena + g.offset, g.host.data(), n * 4, hipMemcpyHostToDevice, b->stream)) !=
                hipSuccess)
            return fail_hip(e, "batch upload");
    }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail_hip(e, "batch upload");
    for (auto& g : b->grids) {
        g.host.clear();
        g.host.shrink_to_fit();
    }
    return GOL_OK;
}

// Release the current plan: descriptors, items, queue state and graph.
void free_plan(gol_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_grids) (void)hipFree(b->d_grids);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_qstate) (void)hipFree(b->d_qstate);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->d_grids = nullptr;
    b->d_items = nullptr;
    b->d_qstate = nullptr;
    b->launches.clear();
}

// Plan shapes, rounds and items (the measured choices unless t_untuned is set)
// and upload descriptors and items.
int build_plan(gol_batch* b) {
    const int T = b->cfg.round_generations ? static_cast<int>(b->cfg.round_generations) : 32;
    for (auto& g : b->grids) {
        int rc = plan_batch_grid(g.x_dw, g.y_rows, T, Planner::kBatch, &g.shape, &g.geom, &g.chunk);
        if (rc) return rc;
    }
    hipError_t e;
    std::vector<BatchGrid> descs(b->grids.size());
    for (size_t i = 0; i < b->grids.size(); ++i) {
        auto& g = b->grids[i];
        const size_t n = static_cast<size_t>(g.x_dw * g.y_rows);
        const size_t stride = (n + 63) & ~size_t(63);
        descs[i].buf[0] = b->arena + g.offset;
        descs[i].buf[1] = b->arena + g.offset + stride;
        descs[i].geo = to_device_geo(g.geom, g.x_dw, g.y_rows);
    }

    // rounds of work items, grouped by shape, largest items first
    std::vector<WorkItem> all;
    std::vector<uint64_t> left(b->grids.size());
    std::vector<int> parity(b->grids.size(), 0);
    for (size_t i = 0; i < b->grids.size(); ++i) left[i] = b->grids[i].generations;
    b->launches.clear();
    b->rounds = 0;
    b->cell_updates = 0;
    // lanes: launch keys by total work, largest first onto the least loaded lane
    std::map<LaunchKey, double> key_work;
    for (const auto& g : b->grids)
        key_work[launch_key(g.shape)] += static_cast<double>(g.x_dw * g.y_rows) * static_cast<double>(g.generations);
    std::vector<std::pair<double, LaunchKey>> by_work;
    for (const auto& kv : key_work) by_work.push_back({kv.second, kv.first});
    std::stable_sort(by_work.begin(), by_work.end(),
                     [](const std::pair<double, LaunchKey>& a, const std::pair<double, LaunchKey>& c) {
                         return a.first > c.first;
                     });
    std::map<LaunchKey, int> lane_of;
    double lane_work[gol_batch::kLanes] = {};
    for (const auto& kw : by_work) {
        int best = 0;
        for (int l = 1; l < gol_batch::kLanes; ++l)
            if (lane_work[l] < lane_work[best]) best = l;
        lane_of[kw.second] = best;
        lane_work[best] += kw.first;
    }
    struct Group {
        std::vector<WorkItem> items;
        int waves = 0;  // narrow grids: the tallest grid's wavefronts
    };
    if (b->queue) {
        int rc = plan_queues(b, lane_of, parity, &all);
        if (rc) return rc;
        left.assign(left.size(), 0);  // every generation is in the queues
    }
    for (;;) {
        std::map<LaunchKey, Group> by_shape;
        bool any = false;
        for (size_t i = 0; i < b->grids.size(); ++i) {
            if (left[i] == 0) continue;
            any = true;
            auto& g = b->grids[i];
            const int gens = static_cast<int>(std::min<uint64_t>(left[i], static_cast<uint64_t>(g.chunk)));
            Group& grp = by_shape[launch_key(g.shape)];
            grp.waves = std::max(grp.waves, g.shape.waves);
            auto& v = grp.items;
            const int tiles = g.geom.tiles_x * g.geom.tiles_y;
            for (int t = 0; t < tiles; ++t) v.push_back(WorkItem{static_cast<int>(i), t, gens, parity[i]});
            parity[i] ^= 1;
            left[i] -= static_cast<uint64_t>(gens);
            b->cell_updates += static_cast<uint64_t>(g.x_dw * 32 * g.y_rows) * static_cast<uint64_t>(gens);
        }
        if (!any) break;
        for (auto& kv : by_shape) {
            auto& v = kv.second.items;
            std::stable_sort(v.begin(), v.end(), [](const WorkItem& a, const WorkItem& c) { return a.gens > c.gens; });
            gol_batch::Launch L;
            const LaunchKey& key = kv.first;
            L.shape = TileShape{key.lane_dwords, key.band_rows, key.band_rows ? key.waves : kv.second.waves};
            L.first = all.size();
            L.count = static_cast<int>(v.size());
            L.lane = lane_of[key];
            all.insert(all.end(), v.begin(), v.end());
            b->launches.push_back(L);
        }
        b->rounds += 1;
    }
    for (size_t i = 0; i < b->grids.size(); ++i) b->grids[i].final_parity = parity[i];
    b->items = all.size();

    if (!descs.empty()) {
        if ((e = hipMalloc(&b->d_grids, descs.size() * sizeof(BatchGrid))) != hipSuccess)
            return fail_hip(e, "batch descriptors");
        if ((e = hipMemcpyAsync(b->d_grids, descs.data(), descs.size() * sizeof(BatchGrid), hipMemcpyHostToDevice,
                                b->stream)) != hipSuccess)
            return fail_hip(e, "batch descriptors upload");
    }
    if (!all.empty()) {
        if ((e = hipMalloc(&b->d_items, all.size() * sizeof(WorkItem))) != hipSuccess)
            return fail_hip(e, "batch items");
        if ((e = hipMemcpyAsync(b->d_items, all.data(), all.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                                b->stream)) != hipSuccess)
            return fail_hip(e, "batch items upload");
    }
    if (b->queue) {
        const size_t words = b->launches.size() + b->grids.size() + 1;
        if (words > (1u << 30)) return fail(GOL_ERR_INVALID_ARGUMENT, "too many grids for a persistent batch");
        b->qstate_words = static_cast<int>(words);
        if ((e = hipMalloc(&b->d_qstate, words * sizeof(unsigned int))) != hipSuccess)
            return fail_hip(e, "batch queue state");
    }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail_hip(e, "batch build");
    return GOL_OK;
}

// Fork the side lanes from the batch stream, run every lane's rounds (or
// persistent queue launches), join.
int enqueue_rounds(gol_batch* b) {
    bool used[gol_batch::kLanes] = {};
    for (const auto& L : b->launches) used[L.lane] = true;
    hipError_t e;
    if (b->queue && (e = launch_zero(b->d_qstate, b->qstate_words, b->stream)) != hipSuccess)
        return fail_hip(e, "batch queue reset");
    bool forked = false;
    for (int l = 1; l < gol_batch::kLanes; ++l) {
        if (!used[l]) continue;
        if (!forked && (e = hipEventRecord(b->fork_ev, b->stream)) != hipSuccess) return fail_hip(e, "batch fork");
        forked = true;
        if ((e = hipStreamWaitEvent(b->side[l - 1], b->fork_ev, 0)) != hipSuccess) return fail_hip(e, "batch fork");
    }
    const bool probe = (b->cfg.flags & GOL_BATCH_WAIT_PROBE) != 0;
    const int n_launches = static_cast<int>(b->launches.size());
    for (int k = 0; k < n_launches; ++k) {
        const auto& L = b->launches[static_cast<size_t>(k)];
        hipStream_t st = L.lane == 0 ? b->stream : b->side[L.lane - 1];
        if (L.workgroups > 0) {
            QueueArgs q;
            q.grids = b->d_grids;
            q.items = b->d_items + L.first;
            q.n_items = L.count;
            q.spin_limit = probe ? 256 : 0;
            q.wait_ticks = b->wait_ticks;
            q.next = b->d_qstate + k;
            q.done = b->d_qstate + n_launches;
            q.error = b->d_qstate + b->qstate_words - 1;
            q.wait_ahead = probe ? (1u << 30) : 0u;
            if ((e = launch_queue(L.shape, q, L.workgroups, st)) != hipSuccess)
                return fail_hip(e, "persistent batch launch");
        } else if ((e = launch_items(L.shape, b->d_grids, b->d_items + L.first, L.count, st)) != hipSuccess) {
            return fail_hip(e, "item_kernel launch");
        }
    }
    for (int l = 1; l < gol_batch::kLanes; ++l) {
        if (!used[l]) continue;
        if ((e = hipEventRecord(b->join_ev[l - 1], b->side[l - 1])) != hipSuccess ||
            (e = hipStreamWaitEvent(b->stream, b->join_ev[l - 1], 0)) != hipSuccess)
            return fail_hip(e, "batch join");
    }
    return GOL_OK;
}

// One run of the current plan: restore the initial states if a run changed
// them, then every launch (replayed from the graph when use_graph).
int run_once(gol_batch* b, double* gpu_ms) {
    int rc;
    if (b->dirty && (rc = gol_batch_reset(b))) return rc;
    b->dirty = true;
    hipError_t e = hipEventRecord(b->ev0, b->stream);
    if (e != hipSuccess) return fail_hip(e, "hipEventRecord");
    if (b->cfg.use_graph) {
        if (!b->graph_exec) {
            if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return fail_hip(e, "hipStreamBeginCapture");
            rc = enqueue_rounds(b);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(b->stream, &graph);
            if (rc) return rc;
            if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
            b->graph = graph;
            if ((e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0)) != hipSuccess)
                return fail_hip(e, "hipGraphInstantiate");
        }
        if ((e = hipGraphLaunch(b->graph_exec, b->stream)) != hipSuccess) return fail_hip(e, "hipGraphLaunch");
    } else if ((rc = enqueue_rounds(b))) {
        return rc;
    }
    if ((e = hipEventRecord(b->ev1, b->stream)) != hipSuccess) return fail_hip(e, "hipEventRecord");
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return fail_hip(e, "gol_batch_run");
    if (b->queue && b->d_qstate) {
        unsigned int err = 0;
        if ((e = hipMemcpy(&err, b->d_qstate + b->qstate_words - 1, sizeof(err), hipMemcpyDeviceToHost)) !=
            hipSuccess)
            return fail_hip(e, "gol_batch_run (queue state)");
        if (err)
            return fail(GOL_ERR_RUNTIME, "persistent batch stopped: an item waited too long for its grid's earlier "
                        "items (GOL_BATCH_PERSISTENT off selects one launch per round)");
    }
    if (gpu_ms) {
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess) return fail_hip(e, "hipEventElapsedTime");
        *gpu_ms = ms;
    }
    b->runs += 1;
    b->result_parity_valid = true;
    for (auto& g : b->grids) g.result_parity = g.final_parity;
    return GOL_OK;
}

// GOL_BATCH_SELECT_PLAN: time the batch under the measured per-class choices
// and under the built-in model (one run each after a warm-up run that records
// the graph) and keep the faster plan.
int select_plan(gol_batch* b) {
    int rc;
    double tuned_ms = 0.0, builtin_ms = 0.0;
    if ((rc = run_once(b, nullptr)) || (rc = run_once(b, &tuned_ms))) return rc;
    free_plan(b);
    t_untuned = true;
    rc = build_plan(b);
    t_untuned = false;
    if (rc || (rc = run_once(b, nullptr)) || (rc = run_once(b, &builtin_ms))) return rc;
    b->plan_tuned_ms = tuned_ms;
    b->plan_builtin_ms = builtin_ms;
    b->plan_builtin = builtin_ms <= tuned_ms;
    if (!b->plan_builtin) {
        free_plan(b);
        if ((rc = build_plan(b))) return rc;
    }
    return GOL_OK;
}

}  // namespace

extern "C" {

int gol_batch_create(const gol_batch_config* config, gol_batch** out) {
    clear_error();
    if (!out) return fail(GOL_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    gol_batch_config cfg{};
    cfg.device = -1;
    cfg.use_graph = 1;
    if (config) cfg = *config;
    if (cfg.round_generations > 32)
        return fail(GOL_ERR_INVALID_ARGUMENT, "round_generations must be at most 32 (the one-lane x halo)");
    if (cfg.flags & ~(GOL_BATCH_PERSISTENT | GOL_BATCH_WAIT_PROBE | GOL_BATCH_SELECT_PLAN))
        return fail(GOL_ERR_INVALID_ARGUMENT, "unknown batch flags 0x%x", cfg.flags);
    if ((cfg.flags & GOL_BATCH_WAIT_PROBE) && !(cfg.flags & GOL_BATCH_PERSISTENT))
        return fail(GOL_ERR_INVALID_ARGUMENT, "GOL_BATCH_WAIT_PROBE applies to persistent batches only");
    int dev = cfg.device;
    hipError_t e;
    if (dev < 0 && (e = hipGetDevice(&dev)) != hipSuccess) return fail_hip(e, "hipGetDevice");
    gol_batch* b = new gol_batch();
    b->cfg = cfg;
    b->queue = (cfg.flags & GOL_BATCH_PERSISTENT) != 0;
    b->device = dev;
    Guard g(dev);
    // a queue wait gives up after 20 s of device time (the wall clock's rate
    // in kHz; 100 MHz if the device does not report it)
    int clock_khz = 0;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || clock_khz <= 0)
        clock_khz = 100000;
    b->wait_ticks = static_cast<unsigned long long>(clock_khz) * 1000ull * 20ull;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming)) != hipSuccess) {
        int rc = fail_hip(e, "gol_batch_create");
        gol_batch_destroy(b);
        return rc;
    }
    for (int l = 0; l < gol_batch::kLanes - 1; ++l) {
        if ((e = hipStreamCreateWithFlags(&b->side[l], hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&b->join_ev[l], hipEventDisableTiming)) != hipSuccess) {
            int rc = fail_hip(e, "gol_batch_create");
            gol_batch_destroy(b);
            return rc;
        }
    }
    *out = b;
    return GOL_OK;
}

int gol_batch_add(gol_batch* b, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t generations,
                  const void* words, uint32_t* grid_id) {
    clear_error();
    if (!b || !words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    if (b->built) return fail(GOL_ERR_RUNTIME, "grids cannot be added after the first gol_batch_run");
    if (word_bits != 32 && word_bits != 64) return fail(GOL_ERR_INVALID_ARGUMENT, "word_bits must be 32 or 64");
    if (x_cells == 0 || y_cells == 0 || x_cells % word_bits != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions X must be a multiple of %u", word_bits);
    if (x_cells / 32 > (1u << 20) || y_cells > (1u << 20) || generations > (1ull << 40))
        return fail(GOL_ERR_INVALID_ARGUMENT, "grid or generation count too large for a batch");
    BatchEntry g;
    g.x_dw = static_cast<int64_t>(x_cells / 32);
    g.y_rows = static_cast<int64_t>(y_cells);
    g.word_bits = word_bits;
    g.generations = generations;
    const size_t n = static_cast<size_t>(g.x_dw * g.y_rows);
    g.host.resize(n);
    std::memcpy(g.host.data(), words, n * 4);
    b->grids.push_back(std::move(g));
    if (grid_id) *grid_id = static_cast<uint32_t>(b->grids.size() - 1);
    return GOL_OK;
}

int gol_batch_reset(gol_batch* b) {
    clear_error();
    if (!b) return fail(GOL_ERR_INVALID_ARGUMENT, "null batch");
    if (!b->built) return GOL_OK;
    Guard g(b->device);
    for (auto& e : b->grids) {
        const size_t n = static_cast<size_t>(e.x_dw * e.y_rows);
        const size_t stride = (n + 63) & ~size_t(63);
        hipError_t r = hipMemcpyAsync(b->arena + e.offset, b->arena + e.offset + 2 * stride, n * 4,
                                      hipMemcpyDeviceToDevice, b->stream);
        if (r != hipSuccess) return fail_hip(r, "gol_batch_reset");
    }
    b->dirty = false;
    b->result_parity_valid = false;
    return GOL_OK;
}

int gol_batch_run(gol_batch* b, double* gpu_ms) {
    clear_error();
    if (!b) return fail(GOL_ERR_INVALID_ARGUMENT, "null batch");
    Guard g(b->device);
    int rc;
    if (!b->built) {
        if ((rc = build_arena(b))) return rc;
        b->built = true;  // from here on the arena holds the initial states
        if ((rc = build_plan(b))) return rc;
        if ((b->cfg.flags & GOL_BATCH_SELECT_PLAN) && (rc = select_plan(b))) return rc;
    }
    return run_once(b, gpu_ms);
}

int gol_batch_fetch(gol_batch* b, uint32_t grid_id, void* words, size_t n_words) {
    clear_error();
    if (!b || !words) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    if (grid_id >= b->grids.size()) return fail(GOL_ERR_NOT_FOUND, "no grid %u in the batch", grid_id);
    auto& g = b->grids[grid_id];
    const size_t n = static_cast<size_t>(g.x_dw * g.y_rows);
    if (n_words != n * 32 / g.word_bits)
        return fail(GOL_ERR_INVALID_ARGUMENT, "expected %zu words, got %zu", n * 32 / g.word_bits, n_words);
    if (!b->built) {
        std::memcpy(words, g.host.data(), n * 4);  // not run yet: the initial state
        return GOL_OK;
    }
    Guard gd(b->device);
    const size_t stride = (n + 63) & ~size_t(63);
    // the last completed run's result, else the initial state the arena keeps
    const size_t at = b->result_parity_valid ? (g.result_parity ? stride : 0) : 2 * stride;
    const uint32_t* src = b->arena + g.offset + at;
    hipError_t e = hipMemcpyAsync(words, src, n * 4, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return fail_hip(e, "gol_batch_fetch");
    return GOL_OK;
}

int gol_batch_get_info(gol_batch* b, gol_batch_info* info) {
    clear_error();
    if (!b || !info) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->grids = static_cast<uint32_t>(b->grids.size());
    info->rounds = b->rounds;
    info->items = b->items;
    info->launches = b->launches.size() + (b->queue && b->built ? 1 : 0);  // + the queue reset
    info->device_bytes = b->arena_dwords * 4 + b->items * sizeof(WorkItem) + b->grids.size() * sizeof(BatchGrid);
    info->cell_updates = b->cell_updates;
    info->plan_builtin = b->plan_builtin ? 1u : 0u;
    info->plan_tuned_ms = b->plan_tuned_ms;
    info->plan_builtin_ms = b->plan_builtin_ms;
    return GOL_OK;
}

void gol_batch_destroy(gol_batch* b) {
    if (!b) return;
    {
        Guard g(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        for (hipStream_t st : b->side)
            if (st) (void)hipStreamSynchronize(st);
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->arena) (void)hipFree(b->arena);
        if (b->d_grids) (void)hipFree(b->d_grids);
        if (b->d_items) (void)hipFree(b->d_items);
        if (b->d_qstate) (void)hipFree(b->d_qstate);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
        for (int l = 0; l < gol_batch::kLanes - 1; ++l) {
            if (b->join_ev[l]) (void)hipEventDestroy(b->join_ev[l]);
            if (b->side[l]) (void)hipStreamDestroy(b->side[l]);
        }
        if (b->stream) (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner: for each representative grid size, time every compiled tile
// shape (and, for grids split into tiles, generations per round R, 3R/4, R/2,
// 3R/8, R/4 and R/8)
// on a batch of synthetic grids of that size filling the device, and install
// the fastest for the size class. The built-in choice is timed too, for the
// report. Measurements use graph replays timed with HIP events.
// ---------------------------------------------------------------------------
namespace {

struct Measure {
    double ms = 0.0;
    uint64_t cell_updates = 0;
};

int measure_batch(int device, int round_generations, uint64_t x, uint64_t y, int count, uint64_t generations,
                  const std::vector<uint32_t>& words, Measure* m) {
    gol_batch_config bc{};
    bc.device = device;
    bc.round_generations = static_cast<uint32_t>(round_generations);
    bc.use_graph = 1;
    gol_batch* b = nullptr;
    int rc = gol_batch_create(&bc, &b);
    if (rc) return rc;
    for (int i = 0; i < count && !rc; ++i) rc = gol_batch_add(b, x, y, 32, generations, words.data(), nullptr);
    double ms = 0.0;
    if (!rc) rc = gol_batch_run(b, &ms);  // plans, uploads, records the graph
    if (!rc) rc = gol_batch_run(b, &ms);  // the timed replay
    gol_batch_info info{};
    if (!rc) rc = gol_batch_get_info(b, &info);
    gol_batch_destroy(b);
    if (rc) return rc;
    m->ms = ms;
    m->cell_updates = info.cell_updates;
    return GOL_OK;
}

}  // namespace

extern "C" {

int gol_batch_autotune(const gol_tune_config* config, gol_tune_entry* entries, uint32_t n_entries) {
    clear_error();
    if (n_entries > 0 && !entries) return fail(GOL_ERR_INVALID_ARGUMENT, "entries is null");
    gol_tune_config cfg{};
    cfg.device = -1;
    if (config) cfg = *config;
    const int R = cfg.round_generations ? static_cast<int>(cfg.round_generations) : 32;
    if (R > 32) return fail(GOL_ERR_INVALID_ARGUMENT, "round_generations must be at most 32 (the one-lane x halo)");
    const uint64_t target = cfg.cells_per_measure ? cfg.cells_per_measure : (uint64_t{1} << 28);
    const uint64_t gens = cfg.generations ? cfg.generations : 4ull * static_cast<uint64_t>(R);
    for (uint32_t k = 0; k < n_entries; ++k) {
        const uint64_t x = entries[k].x_cells, y = entries[k].y_cells;
        if (x == 0 || y == 0 || x % 32 != 0 || x / 32 > (1u << 20) || y > (1u << 20))
            return fail(GOL_ERR_INVALID_ARGUMENT, "entry %u: grid %llux%llu is not a valid size", k,
                        static_cast<unsigned long long>(x), static_cast<unsigned long long>(y));
    }
    int device = cfg.device;
    hipError_t e;
    if (device < 0 && (e = hipGetDevice(&device)) != hipSuccess) return fail_hip(e, "hipGetDevice");

    for (uint32_t k = 0; k < n_entries; ++k) {
        gol_tune_entry& out = entries[k];
        const uint64_t x = out.x_cells, y = out.y_cells;
        const int64_t x_dw = static_cast<int64_t>(x / 32), y_rows = static_cast<int64_t>(y);
        const int count = static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(4096, target / (x * y))));
        std::vector<uint32_t> words(static_cast<size_t>(x_dw * y_rows));
        int rc = gol_fill_bernoulli(words.data(), x, y, 32, 20240601ull + k, 0.35);
        if (rc) return rc;

        std::vector<BatchChoice> cands;
        TileShape narrow;
        if (narrow_shape_for(x_dw, y_rows, &narrow)) cands.push_back(BatchChoice{narrow, 1});
        for (int i = 0; i < kNumTileShapes; ++i) {
            const TileShape& s = kTileShapes[i];
            if (x_dw <= 64 * s.lane_dwords && y_rows <= s.band_rows * s.waves) {
                cands.push_back(BatchChoice{s, 1});
                continue;
            }
            int last = 0;
            for (int steps : {R, 3 * R / 4, R / 2, 3 * R / 8, R / 4, R / 8}) {
                if (steps < 1 || steps == last) continue;
                last = steps;
                TileGeometry g;
                const char* why = nullptr;
                if (plan_tiles(s, x_dw, y_rows, steps, &g, &why)) cands.push_back(BatchChoice{s, steps});
            }
        }

        // the built-in choice
        Measure base;
        t_untuned = true;
        rc = measure_batch(device, R, x, y, count, gens, words, &base);
        t_untuned = false;
        if (rc) return rc;
        TileShape base_shape;
        TileGeometry base_geom;
        int base_round = 0;
        t_untuned = true;
        rc = plan_batch_grid(x_dw, y_rows, R, Planner::kBatch, &base_shape, &base_geom, &base_round);
        t_untuned = false;
        if (rc) return rc;

        BatchChoice best{base_shape, base_round};
        double best_rate = base.cell_updates / (base.ms * 1e-3);
        uint32_t measured = 0;
        for (const auto& c : cands) {
            Measure m;
            t_forced = &c;
            rc = measure_batch(device, R, x, y, count, gens, words, &m);
            t_forced = nullptr;
            if (rc) {
                clear_error();
                continue;  // this shape cannot run the grid: not a candidate
            }
            ++measured;
            const double rate = m.cell_updates / (m.ms * 1e-3);
            if (rate > best_rate) {
                best_rate = rate;
                best = c;
            }
        }
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[size_class(x_dw, y_rows)] = BatchChoice{best.shape, std::max(best.steps, 1)};
        }
        out.lane_dwords = static_cast<uint32_t>(best.shape.lane_dwords);
        out.band_rows = static_cast<uint32_t>(best.shape.band_rows);
        out.waves = static_cast<uint32_t>(best.shape.waves);
        const bool one_tile = is_narrow(best.shape) ||
                              (x_dw <= 64 * best.shape.lane_dwords && y_rows <= best.shape.band_rows * best.shape.waves);
        out.generations_per_round = one_tile ? 0u : static_cast<uint32_t>(best.steps);
        out.candidates = measured;
        out.grids_measured = static_cast<uint32_t>(count);
        out.tuned_cell_updates_per_s = best_rate;
        out.builtin_cell_updates_per_s = base.cell_updates / (base.ms * 1e-3);
    }
    return GOL_OK;
}

int gol_batch_tuning_reset(void) {
    clear_error();
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
    return GOL_OK;
}

}  // extern "C"
