// Continuous-batching job server (no reference counterpart: the reference runs
// one grid per process, src/infrastructure/experiment_manager.hpp:235-270).
//
// A scheduler thread owns the streams and runs rounds of work items, like a
// batch (gol_batch.hip): every resident job contributes one item per tile of
// its grid, advancing it by its per-round generation count. Between rounds the
// scheduler
//   * admits queued jobs in arrival order while their packed state (two
//     buffers) fits the HBM budget: a first-fit arena inside one device
//     allocation, so admission never calls the device allocator;
//   * uploads the new jobs' states and descriptors ahead of the next round,
//     which waits for them: arrivals fold into the running batch;
//   * copies finished jobs' results to pinned host memory after the round that
//     finished them, and frees their arena space once that copy is done.
// Planning runs ahead of the GPU (a job's progress is known on the host), and
// up to two rounds are in flight. Each job's result is the same as running it
// alone: items of different jobs never share buffers.
// A round's launches (one per tile shape, one per row width of narrow grids)
// touch disjoint jobs, so they run side by side on four compute lanes: a
// stream each, with its own double-buffered item list. A launch key stays on
// the lane it was first given (the least loaded) for the server's lifetime, so
// each job's rounds are ordered on one stream and rounds need no cross-stream
// waits; small shapes fill the CUs that a large shape's last workgroups leave
// idle. Job transfers run on two copy streams beside the lanes: new jobs'
// uploads on one, which their lane waits for, and result downloads on the
// other, after the round that finished the job.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/golhip.h"
#include "gol_capi_internal.hpp"
#include "gol_kernels.hpp"

using namespace golhip;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

enum JobState { kQueued = 0, kResident = 1, kDraining = 2, kDone = 3, kFailed = 4 };

struct ServerJob {
    uint64_t id = 0;
    int64_t x_dw = 0, y_rows = 0;
    uint32_t word_bits = 64;
    uint64_t generations = 0, remaining = 0;
    int parity = 0;
    TileShape shape{1, 16, 8};
    TileGeometry geom{};
    int chunk = 0;
    int slot = -1;
    int lane = -1;       // compute lane of the job's launch key
    bool fresh = false;  // admitted in this pass: its lane waits for the upload
    size_t offset = 0, bytes = 0;  // arena block holding buf0 then buf1
    size_t dwords = 0;
    uint32_t* pinned_in = nullptr;
    uint32_t* pinned_out = nullptr;
    size_t pinned_off = 0, pinned_block = 0;  // block of the server's pinned region (0 bytes: own allocations)
    hipEvent_t done_ev = nullptr;
    JobState state = kQueued;
    int error = GOL_OK;
    uint32_t rounds = 0;
    Clock::time_point t_submit, t_admit, t_done;
};

// First-fit allocator over [0, size) in 256-byte units, coalescing on free.
class ArenaAllocator {
   public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

   private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

}  // namespace

struct gol_server {
    gol_server_config cfg{};
    int device = 0;
    int round_generations = 32;
    uint8_t* arena = nullptr;
    size_t arena_bytes = 0;
    ArenaAllocator alloc;
    // pinned host region for the jobs' input and result buffers (taken under
    // mu), so that a submit does not pin pages; a job that finds no room in
    // it pins its own
    uint8_t* pinned = nullptr;
    size_t pinned_bytes = 0;
    ArenaAllocator pinned_alloc;
    BatchGrid* d_desc = nullptr;
    BatchGrid* h_desc = nullptr;  // pinned staging, one entry per slot
    std::vector<int> free_slots;
    struct Lane {
        hipStream_t stream = nullptr;
        // a round's items, double-buffered, in pinned host memory that the
        // kernels read directly (coherent, mapped): no copy per round
        WorkItem* h_items[2] = {nullptr, nullptr};
        WorkItem* dev_items[2] = {nullptr, nullptr};  // the same memory as the device addresses it
        size_t cap = 0;
        hipEvent_t round_ev[2] = {nullptr, nullptr};
        bool used[2] = {false, false};
        double work = 0;  // dword-generations given so far, to place new launch keys
    };
    static constexpr int kLanes = 4;
    Lane lanes[kLanes];
    std::map<LaunchKey, int> lane_of;
    hipStream_t up = nullptr, down = nullptr;  // host-to-device and device-to-host copies
    hipEvent_t up_ev = nullptr;                 // the pass's job uploads are done
    hipEvent_t up_all_ev = nullptr;             // every upload so far is done (orders downloads)
    int next_buf = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerJob*> queue;
    std::vector<ServerJob*> resident;
    std::vector<ServerJob*> draining;
    std::unordered_map<uint64_t, ServerJob*> jobs;
    std::thread worker;
    bool stop = false;
    int worker_error = GOL_OK;
    uint64_t next_id = 1;
    gol_server_stats stats{};
};

namespace {

// Returns a job's block of the server's pinned region; the caller holds s->mu
// (or is the only thread left).
void release_pinned_block(gol_server* s, ServerJob* j) {
    if (!j->pinned_block) return;
    s->pinned_alloc.release(j->pinned_off, j->pinned_block);
    s->stats.pinned_region_used = s->pinned_alloc.used();
    j->pinned_block = 0;
    j->pinned_in = j->pinned_out = nullptr;
}

// Frees what else a job owns: host buffers it pinned itself (when the region
// had no room) and its event. Called without s->mu held: hipHostFree may wait
// for the device, which would stall the scheduler and every caller.
void destroy_job(ServerJob* j) {
    if (j->pinned_in) (void)hipHostFree(j->pinned_in);
    if (j->pinned_out) (void)hipHostFree(j->pinned_out);
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
    delete j;
}

int ensure_item_capacity(gol_server::Lane& L, size_t n) {
    if (n <= L.cap) return GOL_OK;
    hipError_t e;
    if ((e = hipStreamSynchronize(L.stream)) != hipSuccess) return fail_hip(e, "server lane");
    size_t cap = std::max<size_t>(n, 2 * L.cap);
    for (int k = 0; k < 2; ++k) {
        if (L.h_items[k]) (void)hipHostFree(L.h_items[k]);
        L.h_items[k] = nullptr;
        L.dev_items[k] = nullptr;
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&L.h_items[k]), cap * sizeof(WorkItem),
                               hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess ||
            (e = hipHostGetDevicePointer(reinterpret_cast<void**>(&L.dev_items[k]), L.h_items[k], 0)) != hipSuccess)
            return fail_hip(e, "server items (host)");
    }
    L.cap = cap;
    L.used[0] = L.used[1] = false;
    return GOL_OK;
}

// The lane of a launch key: fixed once given, new keys to the least loaded lane.
int lane_for(gol_server* s, const TileShape& sh) {
    const LaunchKey key = launch_key(sh);
    auto it = s->lane_of.find(key);
    if (it != s->lane_of.end()) return it->second;
    int best = 0;
    for (int i = 1; i < gol_server::kLanes; ++i)
        if (s->lanes[i].work < s->lanes[best].work) best = i;
    s->lane_of[key] = best;
    return best;
}

// One pass of the scheduler; returns false when the worker should exit.
bool schedule_once(gol_server* s) {
    std::vector<ServerJob*> admitted;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait(lk, [&] {
            return s->stop || !s->queue.empty() || !s->resident.empty() || !s->draining.empty();
        });
        if (s->stop) return false;
        while (!s->queue.empty()) {
            ServerJob* j = s->queue.front();
            if (j->bytes > s->arena_bytes) {
                s->queue.pop_front();
                j->state = kFailed;
                j->error = GOL_ERR_OUT_OF_MEMORY;
                j->t_done = Clock::now();
                s->stats.rejected += 1;
                s->cv_done.notify_all();
                continue;
            }
            size_t off = 0;
            if (s->free_slots.empty() || !s->alloc.alloc(j->bytes, &off)) break;  // head of line waits
            s->queue.pop_front();
            j->offset = off;
            j->slot = s->free_slots.back();
            s->free_slots.pop_back();
            j->state = kResident;
            j->fresh = true;
            j->t_admit = Clock::now();
            admitted.push_back(j);
            s->stats.admitted += 1;
            s->stats.resident_jobs += 1;
            s->stats.resident_bytes_peak = std::max<uint64_t>(s->stats.resident_bytes_peak, s->alloc.used());
        }
    }
    hipError_t e;
    for (ServerJob* j : admitted) {
        uint32_t* buf0 = reinterpret_cast<uint32_t*>(s->arena + j->offset);
        const size_t stride = ((j->dwords + 63) & ~size_t(63));
        s->h_desc[j->slot].buf[0] = buf0;
        s->h_desc[j->slot].buf[1] = buf0 + stride;
        s->h_desc[j->slot].geo = to_device_geo(j->geom, j->x_dw, j->y_rows);
        if ((e = hipMemcpyAsync(buf0, j->pinned_in, j->dwords * 4, hipMemcpyHostToDevice, s->up)) != hipSuccess ||
            (e = hipMemcpyAsync(s->d_desc + j->slot, s->h_desc + j->slot, sizeof(BatchGrid), hipMemcpyHostToDevice,
                                s->up)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server upload");
            return false;
        }
        s->resident.push_back(j);
    }

    // one round over the resident jobs: per lane, one launch per launch key
    struct Group {
        std::vector<WorkItem> items;
        int waves = 0;  // narrow grids: the tallest grid's wavefronts
    };
    std::map<LaunchKey, Group> groups[gol_server::kLanes];
    bool lane_waits_upload[gol_server::kLanes] = {};
    bool lane_finishes[gol_server::kLanes] = {};
    size_t n_items = 0;
    std::vector<ServerJob*> finished;
    uint64_t round_updates = 0;
    for (ServerJob* j : s->resident) {
        if (j->remaining == 0) {  // no generations: never joins a round
            finished.push_back(j);
            continue;
        }
        const int gens = static_cast<int>(std::min<uint64_t>(j->remaining, static_cast<uint64_t>(j->chunk)));
        j->lane = lane_for(s, j->shape);
        Group& g = groups[j->lane][launch_key(j->shape)];
        g.waves = std::max(g.waves, j->shape.waves);
        const int tiles = j->geom.tiles_x * j->geom.tiles_y;
        for (int t = 0; t < tiles; ++t) g.items.push_back(WorkItem{j->slot, t, gens, j->parity});
        n_items += static_cast<size_t>(tiles);
        if (j->fresh) lane_waits_upload[j->lane] = true;
        j->fresh = false;
        j->parity ^= 1;
        j->remaining -= static_cast<uint64_t>(gens);
        j->rounds += 1;
        const uint64_t updates = static_cast<uint64_t>(j->x_dw * 32 * j->y_rows) * static_cast<uint64_t>(gens);
        round_updates += updates;
        s->lanes[j->lane].work += static_cast<double>(updates);
        if (j->remaining == 0) {
            finished.push_back(j);
            lane_finishes[j->lane] = true;
        }
    }
    const int k = s->next_buf;
    if (n_items > 0) {
        if (!admitted.empty() && (e = hipEventRecord(s->up_ev, s->up)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server upload order");
            return false;
        }
        for (int ln = 0; ln < gol_server::kLanes; ++ln) {
            if (groups[ln].empty()) continue;
            gol_server::Lane& L = s->lanes[ln];
            size_t count = 0;
            for (auto& kv : groups[ln]) count += kv.second.items.size();
            // the lane's items buffer k was last read two passes ago
            if (L.used[k] && (e = hipEventSynchronize(L.round_ev[k])) != hipSuccess) {
                s->worker_error = fail_hip(e, "server round wait");
                return false;
            }
            int rc = ensure_item_capacity(L, count);
            if (rc) {
                s->worker_error = rc;
                return false;
            }
            size_t at = 0;
            for (auto& kv : groups[ln]) {
                std::memcpy(L.h_items[k] + at, kv.second.items.data(), kv.second.items.size() * sizeof(WorkItem));
                at += kv.second.items.size();
            }
            // the lane's round follows new jobs' uploads; its kernels read the
            // items from host memory, written above before the launches
            if (lane_waits_upload[ln] && (e = hipStreamWaitEvent(L.stream, s->up_ev, 0)) != hipSuccess) {
                s->worker_error = fail_hip(e, "server upload order");
                return false;
            }
            at = 0;
            for (auto& kv : groups[ln]) {
                const LaunchKey& key = kv.first;
                const TileShape sh{key.lane_dwords, key.band_rows, key.band_rows ? key.waves : kv.second.waves};
                const int n = static_cast<int>(kv.second.items.size());
                if ((e = launch_items(sh, s->d_desc, L.dev_items[k] + at, n, L.stream)) != hipSuccess) {
                    s->worker_error = fail_hip(e, "server round launch");
                    return false;
                }
                at += static_cast<size_t>(n);
            }
            if ((e = hipEventRecord(L.round_ev[k], L.stream)) != hipSuccess) {
                s->worker_error = fail_hip(e, "hipEventRecord");
                return false;
            }
            L.used[k] = true;
        }
        s->next_buf ^= 1;
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rounds += 1;
        s->stats.items += n_items;
        s->stats.cell_updates += round_updates;
    }
    // results of the jobs this round finished, ordered after their lanes'
    // round and after every upload (a job with no generations never joins a
    // round)
    if (!finished.empty()) {
        for (int ln = 0; ln < gol_server::kLanes; ++ln) {
            if (lane_finishes[ln] && (e = hipStreamWaitEvent(s->down, s->lanes[ln].round_ev[k], 0)) != hipSuccess) {
                s->worker_error = fail_hip(e, "server download order");
                return false;
            }
        }
        if ((e = hipEventRecord(s->up_all_ev, s->up)) != hipSuccess ||
            (e = hipStreamWaitEvent(s->down, s->up_all_ev, 0)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server download order");
            return false;
        }
    }
    for (ServerJob* j : finished) {
        const size_t stride = ((j->dwords + 63) & ~size_t(63));
        const uint32_t* src = reinterpret_cast<const uint32_t*>(s->arena + j->offset) + (j->parity ? stride : 0);
        if ((e = hipMemcpyAsync(j->pinned_out, src, j->dwords * 4, hipMemcpyDeviceToHost, s->down)) != hipSuccess ||
            (e = hipEventRecord(j->done_ev, s->down)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server download");
            return false;
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->state = kDraining;
        }
        s->resident.erase(std::find(s->resident.begin(), s->resident.end(), j));
        s->draining.push_back(j);
    }
    // retire drained jobs; block on the oldest one only when nothing was launched
    if (!s->draining.empty()) {
        if (n_items == 0 && (e = hipEventSynchronize(s->draining.front()->done_ev)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server drain");
            return false;
        }
        std::vector<ServerJob*> still;
        std::lock_guard<std::mutex> lk(s->mu);
        for (ServerJob* j : s->draining) {
            if (hipEventQuery(j->done_ev) == hipSuccess) {
                j->t_done = Clock::now();
                j->state = kDone;
                s->alloc.release(j->offset, j->bytes);
                s->free_slots.push_back(j->slot);
                s->stats.completed += 1;
                s->stats.resident_jobs -= 1;
            } else {
                still.push_back(j);
            }
        }
        s->draining.swap(still);
        s->cv_done.notify_all();
    }
    return true;
}

void sync_streams(gol_server* s) {
    for (hipStream_t st : {s->up, s->down})
        if (st) (void)hipStreamSynchronize(st);
    for (auto& L : s->lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
}

void worker_main(gol_server* s) {
    (void)hipSetDevice(s->device);
    while (schedule_once(s)) {
    }
    sync_streams(s);
    std::lock_guard<std::mutex> lk(s->mu);
    // anything left is failed (server stopping or a device error)
    const int code = s->worker_error ? s->worker_error : GOL_ERR_RUNTIME;
    for (auto& kv : s->jobs) {
        ServerJob* j = kv.second;
        if (j->state != kDone && j->state != kFailed) {
            j->state = kFailed;
            j->error = code;
            j->t_done = Clock::now();
        }
    }
    s->stop = true;
    s->cv_done.notify_all();
}

void destroy_resources(gol_server* s) {
    sync_streams(s);
    for (auto& L : s->lanes) {
        for (int k = 0; k < 2; ++k) {
            if (L.h_items[k]) (void)hipHostFree(L.h_items[k]);
            if (L.round_ev[k]) (void)hipEventDestroy(L.round_ev[k]);
        }
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (hipEvent_t ev : {s->up_ev, s->up_all_ev})
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t st : {s->up, s->down})
        if (st) (void)hipStreamDestroy(st);
    if (s->d_desc) (void)hipFree(s->d_desc);
    if (s->h_desc) (void)hipHostFree(s->h_desc);
    if (s->arena) (void)hipFree(s->arena);
    if (s->pinned) (void)hipHostFree(s->pinned);
}

}  // namespace

extern "C" {

int gol_server_create(const gol_server_config* config, gol_server** out) {
    clear_error();
    if (!out) return fail(GOL_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    gol_server_config cfg{};
    cfg.device = -1;
    if (config) cfg = *config;
    if (cfg.round_generations > 32)
        return fail(GOL_ERR_INVALID_ARGUMENT, "round_generations must be at most 32 (the one-lane x halo)");
    if (cfg.hbm_budget_bytes == 0) cfg.hbm_budget_bytes = 8ull << 30;
    if (cfg.max_jobs == 0) cfg.max_jobs = 4096;
    int dev = cfg.device;
    hipError_t e;
    if (dev < 0 && (e = hipGetDevice(&dev)) != hipSuccess) return fail_hip(e, "hipGetDevice");
    gol_server* s = new gol_server();
    s->cfg = cfg;
    s->device = dev;
    s->round_generations = cfg.round_generations ? static_cast<int>(cfg.round_generations) : 32;
    (void)hipSetDevice(dev);
    s->arena_bytes = (cfg.hbm_budget_bytes + 255) & ~uint64_t(255);
    if ((e = hipMalloc(&s->arena, s->arena_bytes)) != hipSuccess ||
        (e = hipMalloc(&s->d_desc, cfg.max_jobs * sizeof(BatchGrid))) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&s->h_desc), cfg.max_jobs * sizeof(BatchGrid),
                           hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->up_ev, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->up_all_ev, hipEventDisableTiming)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking)) != hipSuccess) {
        int rc = fail_hip(e, "gol_server_create");
        destroy_resources(s);
        delete s;
        return rc;
    }
    // a thirty-second of the HBM budget, at most 1 GiB; without it every job
    // pins its own buffers
    s->pinned_bytes = std::min<size_t>(size_t(1) << 30, (s->arena_bytes / 32) & ~size_t(255));
    if (s->pinned_bytes &&
        hipHostMalloc(reinterpret_cast<void**>(&s->pinned), s->pinned_bytes, hipHostMallocDefault) != hipSuccess) {
        s->pinned = nullptr;
        s->pinned_bytes = 0;
    }
    s->pinned_alloc.reset(s->pinned_bytes);
    s->stats.pinned_region_bytes = s->pinned_bytes;
    for (auto& L : s->lanes) {
        if ((e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&L.round_ev[0], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&L.round_ev[1], hipEventDisableTiming)) != hipSuccess) {
            int rc = fail_hip(e, "gol_server_create");
            destroy_resources(s);
            delete s;
            return rc;
        }
    }
    s->alloc.reset(s->arena_bytes);
    for (int i = static_cast<int>(cfg.max_jobs) - 1; i >= 0; --i) s->free_slots.push_back(i);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return GOL_OK;
}

int gol_server_submit(gol_server* s, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t generations,
                      const void* words, uint64_t* job_id) {
    clear_error();
    if (!s || !words || !job_id) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    if (word_bits != 32 && word_bits != 64) return fail(GOL_ERR_INVALID_ARGUMENT, "word_bits must be 32 or 64");
    if (x_cells == 0 || y_cells == 0 || x_cells % word_bits != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions X must be a multiple of %u", word_bits);
    if (x_cells / 32 > (1u << 20) || y_cells > (1u << 20) || generations > (1ull << 40))
        return fail(GOL_ERR_INVALID_ARGUMENT, "grid or generation count too large");
    ServerJob* j = new ServerJob();
    j->x_dw = static_cast<int64_t>(x_cells / 32);
    j->y_rows = static_cast<int64_t>(y_cells);
    j->word_bits = word_bits;
    j->generations = j->remaining = generations;
    j->dwords = static_cast<size_t>(j->x_dw * j->y_rows);
    j->bytes = 2 * ((j->dwords + 63) & ~size_t(63)) * 4;
    // the built-in choices and rounds at least half as deep as a batch may
    // take: a job's latency counts its rounds
    int rc = plan_batch_grid(j->x_dw, j->y_rows, s->round_generations, Planner::kServer, &j->shape, &j->geom,
                             &j->chunk);
    if (rc) {
        delete j;
        return rc;
    }
    hipError_t e = hipSuccess;
    const size_t half = (j->dwords * 4 + 255) & ~size_t(255);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        size_t off = 0;
        if (s->pinned && s->pinned_alloc.alloc(2 * half, &off)) {
            j->pinned_off = off;
            j->pinned_block = 2 * half;
            j->pinned_in = reinterpret_cast<uint32_t*>(s->pinned + off);
            j->pinned_out = reinterpret_cast<uint32_t*>(s->pinned + off + half);
            s->stats.pinned_region_used = s->pinned_alloc.used();
            s->stats.pinned_region_peak = std::max<uint64_t>(s->stats.pinned_region_peak, s->pinned_alloc.used());
        } else {
            s->stats.pinned_own_jobs += 1;
        }
    }
    if ((!j->pinned_block &&
         ((e = hipHostMalloc(reinterpret_cast<void**>(&j->pinned_in), j->dwords * 4, hipHostMallocDefault)) !=
              hipSuccess ||
          (e = hipHostMalloc(reinterpret_cast<void**>(&j->pinned_out), j->dwords * 4, hipHostMallocDefault)) !=
              hipSuccess)) ||
        (e = hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming)) != hipSuccess) {
        rc = fail_hip(e, "gol_server_submit");
        {
            std::lock_guard<std::mutex> lk(s->mu);
            release_pinned_block(s, j);
        }
        destroy_job(j);
        return rc;
    }
    std::memcpy(j->pinned_in, words, j->dwords * 4);
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop) {
        release_pinned_block(s, j);
        lk.unlock();
        destroy_job(j);
        return fail(GOL_ERR_RUNTIME, "server is stopped");
    }
    j->id = s->next_id++;
    j->t_submit = Clock::now();
    s->jobs[j->id] = j;
    s->queue.push_back(j);
    s->stats.submitted += 1;
    *job_id = j->id;
    s->cv_work.notify_one();
    return GOL_OK;
}

int gol_server_poll(gol_server* s, uint64_t job_id, int* state) {
    clear_error();
    if (!s || !state) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(GOL_ERR_NOT_FOUND, "no job %llu", static_cast<unsigned long long>(job_id));
    *state = it->second->state;
    return GOL_OK;
}

int gol_server_wait(gol_server* s, uint64_t job_id, void* words, size_t n_words, gol_job_stats* stats) {
    clear_error();
    if (!s) return fail(GOL_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(GOL_ERR_NOT_FOUND, "no job %llu", static_cast<unsigned long long>(job_id));
    ServerJob* j = it->second;
    s->cv_done.wait(lk, [&] { return j->state == kDone || j->state == kFailed; });
    s->jobs.erase(it);
    lk.unlock();
    int rc = GOL_OK;
    if (j->state == kFailed) {
        rc = fail(j->error ? j->error : GOL_ERR_RUNTIME, "job %llu failed%s", static_cast<unsigned long long>(job_id),
                  j->error == GOL_ERR_OUT_OF_MEMORY ? ": its state exceeds the HBM budget" : "");
    } else if (words) {
        const size_t want = j->dwords * 32 / j->word_bits;
        if (n_words != want)
            rc = fail(GOL_ERR_INVALID_ARGUMENT, "expected %zu words, got %zu", want, n_words);
        else
            std::memcpy(words, j->pinned_out, j->dwords * 4);
    }
    if (stats) {
        stats->queue_ms = j->state == kDone ? ms_between(j->t_submit, j->t_admit) : 0.0;
        stats->latency_ms = ms_between(j->t_submit, j->t_done);
        stats->rounds = j->rounds;
        stats->reserved = 0;
    }
    lk.lock();
    release_pinned_block(s, j);
    lk.unlock();
    destroy_job(j);
    return rc;
}

int gol_server_get_stats(gol_server* s, gol_server_stats* out) {
    clear_error();
    if (!s || !out) return fail(GOL_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->queued_jobs = static_cast<uint32_t>(s->queue.size());
    out->resident_bytes = s->alloc.used();
    return GOL_OK;
}

void gol_server_destroy(gol_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    destroy_resources(s);
    for (auto& kv : s->jobs) {
        release_pinned_block(s, kv.second);
        destroy_job(kv.second);
    }
    delete s;
}

}  // extern "C"
