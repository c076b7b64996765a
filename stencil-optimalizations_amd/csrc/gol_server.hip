// Continuous-batching job server (no reference counterpart: the reference runs
// one grid per process, src/infrastructure/experiment_manager.hpp:235-270).
//
// A scheduler thread owns one stream and runs rounds of work items, like a
// batch (gol_batch.hip): every resident job contributes one item per tile of
// its grid, advancing it by its per-round generation count. Between rounds the
// scheduler
//   * admits queued jobs in arrival order while their packed state (two
//     buffers) fits the HBM budget: a first-fit arena inside one device
//     allocation, so admission never calls the device allocator;
//   * uploads the new jobs' states and descriptors on the same stream, so they
//     join the very next round: arrivals fold into the running batch;
//   * copies finished jobs' results to pinned host memory after the round that
//     finished them, and frees their arena space once that copy is done.
// Planning runs ahead of the GPU (a job's progress is known on the host), and
// up to two rounds are in flight. Each job's result is the same as running it
// alone: items of different jobs never share buffers.

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
    size_t offset = 0, bytes = 0;  // arena block holding buf0 then buf1
    size_t dwords = 0;
    uint32_t* pinned_in = nullptr;
    uint32_t* pinned_out = nullptr;
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
    hipStream_t stream = nullptr;
    uint8_t* arena = nullptr;
    size_t arena_bytes = 0;
    ArenaAllocator alloc;
    BatchGrid* d_desc = nullptr;
    BatchGrid* h_desc = nullptr;  // pinned staging, one entry per slot
    std::vector<int> free_slots;
    WorkItem* d_items[2] = {nullptr, nullptr};
    WorkItem* h_items[2] = {nullptr, nullptr};
    size_t items_cap = 0;
    hipEvent_t round_ev[2] = {nullptr, nullptr};
    bool round_ev_used[2] = {false, false};
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

void free_job(ServerJob* j) {
    if (j->pinned_in) (void)hipHostFree(j->pinned_in);
    if (j->pinned_out) (void)hipHostFree(j->pinned_out);
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
    delete j;
}

int ensure_item_capacity(gol_server* s, size_t n) {
    if (n <= s->items_cap) return GOL_OK;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail_hip(e, "server stream");
    size_t cap = std::max<size_t>(n, 2 * s->items_cap);
    for (int k = 0; k < 2; ++k) {
        if (s->d_items[k]) (void)hipFree(s->d_items[k]);
        if (s->h_items[k]) (void)hipHostFree(s->h_items[k]);
        s->d_items[k] = nullptr;
        s->h_items[k] = nullptr;
        if ((e = hipMalloc(&s->d_items[k], cap * sizeof(WorkItem))) != hipSuccess)
            return fail_hip(e, "server items");
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->h_items[k]), cap * sizeof(WorkItem), hipHostMallocDefault)) !=
            hipSuccess)
            return fail_hip(e, "server items (host)");
    }
    s->items_cap = cap;
    s->round_ev_used[0] = s->round_ev_used[1] = false;
    return GOL_OK;
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
        if ((e = hipMemcpyAsync(buf0, j->pinned_in, j->dwords * 4, hipMemcpyHostToDevice, s->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(s->d_desc + j->slot, s->h_desc + j->slot, sizeof(BatchGrid), hipMemcpyHostToDevice,
                                s->stream)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server upload");
            return false;
        }
        s->resident.push_back(j);
    }

    // one round over the resident jobs, grouped by tile shape
    std::map<std::tuple<int, int, int>, std::vector<WorkItem>> by_shape;
    size_t n_items = 0;
    std::vector<ServerJob*> finished;
    uint64_t round_updates = 0;
    for (ServerJob* j : s->resident) {
        if (j->remaining == 0) {
            finished.push_back(j);
            continue;
        }
        const int gens = static_cast<int>(std::min<uint64_t>(j->remaining, static_cast<uint64_t>(j->chunk)));
        auto& v = by_shape[{j->shape.lane_dwords, j->shape.band_rows, j->shape.waves}];
        const int tiles = j->geom.tiles_x * j->geom.tiles_y;
        for (int t = 0; t < tiles; ++t) v.push_back(WorkItem{j->slot, t, gens, j->parity});
        n_items += static_cast<size_t>(tiles);
        j->parity ^= 1;
        j->remaining -= static_cast<uint64_t>(gens);
        j->rounds += 1;
        round_updates += static_cast<uint64_t>(j->x_dw * 32 * j->y_rows) * static_cast<uint64_t>(gens);
        if (j->remaining == 0) finished.push_back(j);
    }
    if (n_items > 0) {
        int rc = ensure_item_capacity(s, n_items);
        if (rc) {
            s->worker_error = rc;
            return false;
        }
        const int k = s->next_buf;
        if (s->round_ev_used[k] && (e = hipEventSynchronize(s->round_ev[k])) != hipSuccess) {
            s->worker_error = fail_hip(e, "server round wait");
            return false;
        }
        size_t at = 0;
        std::vector<std::pair<TileShape, std::pair<size_t, int>>> launches;
        for (auto& kv : by_shape) {
            auto& v = kv.second;
            std::memcpy(s->h_items[k] + at, v.data(), v.size() * sizeof(WorkItem));
            launches.push_back({TileShape{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)},
                                {at, static_cast<int>(v.size())}});
            at += v.size();
        }
        if ((e = hipMemcpyAsync(s->d_items[k], s->h_items[k], at * sizeof(WorkItem), hipMemcpyHostToDevice,
                                s->stream)) != hipSuccess) {
            s->worker_error = fail_hip(e, "server items upload");
            return false;
        }
        for (auto& L : launches) {
            if ((e = launch_items(L.first, s->d_desc, s->d_items[k] + L.second.first, L.second.second, s->stream)) !=
                hipSuccess) {
                s->worker_error = fail_hip(e, "server round launch");
                return false;
            }
        }
        if ((e = hipEventRecord(s->round_ev[k], s->stream)) != hipSuccess) {
            s->worker_error = fail_hip(e, "hipEventRecord");
            return false;
        }
        s->round_ev_used[k] = true;
        s->next_buf ^= 1;
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rounds += 1;
        s->stats.items += n_items;
        s->stats.cell_updates += round_updates;
    }
    // results of the jobs this round finished, ordered after the round
    for (ServerJob* j : finished) {
        const size_t stride = ((j->dwords + 63) & ~size_t(63));
        const uint32_t* src = reinterpret_cast<const uint32_t*>(s->arena + j->offset) + (j->parity ? stride : 0);
        if ((e = hipMemcpyAsync(j->pinned_out, src, j->dwords * 4, hipMemcpyDeviceToHost, s->stream)) != hipSuccess ||
            (e = hipEventRecord(j->done_ev, s->stream)) != hipSuccess) {
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

void worker_main(gol_server* s) {
    (void)hipSetDevice(s->device);
    while (schedule_once(s)) {
    }
    (void)hipStreamSynchronize(s->stream);
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
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int k = 0; k < 2; ++k) {
        if (s->d_items[k]) (void)hipFree(s->d_items[k]);
        if (s->h_items[k]) (void)hipHostFree(s->h_items[k]);
        if (s->round_ev[k]) (void)hipEventDestroy(s->round_ev[k]);
    }
    if (s->d_desc) (void)hipFree(s->d_desc);
    if (s->h_desc) (void)hipHostFree(s->h_desc);
    if (s->arena) (void)hipFree(s->arena);
    if (s->stream) (void)hipStreamDestroy(s->stream);
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
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&s->arena, s->arena_bytes)) != hipSuccess ||
        (e = hipMalloc(&s->d_desc, cfg.max_jobs * sizeof(BatchGrid))) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&s->h_desc), cfg.max_jobs * sizeof(BatchGrid),
                           hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->round_ev[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&s->round_ev[1], hipEventDisableTiming)) != hipSuccess) {
        int rc = fail_hip(e, "gol_server_create");
        destroy_resources(s);
        delete s;
        return rc;
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
    // the built-in choices, not the autotuner's: those maximise a batch's
    // throughput with shallower rounds, and a job's latency counts rounds
    int rc = plan_batch_grid(j->x_dw, j->y_rows, s->round_generations, false, &j->shape, &j->geom, &j->chunk);
    if (rc) {
        delete j;
        return rc;
    }
    hipError_t e;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&j->pinned_in), j->dwords * 4, hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&j->pinned_out), j->dwords * 4, hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming)) != hipSuccess) {
        rc = fail_hip(e, "gol_server_submit");
        free_job(j);
        return rc;
    }
    std::memcpy(j->pinned_in, words, j->dwords * 4);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        free_job(j);
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
    free_job(j);
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
    for (auto& kv : s->jobs) free_job(kv.second);
    delete s;
}

}  // extern "C"
