/*
 * golhip — C ABI of the MI355X Game-of-Life engine (B3/S23, cells outside
 * the grid dead).
 *
 * This is the boundary a host program binds to. It replaces the CUDA side of
 * the reference's algorithm plug-ins (the device allocation, copies and
 * per-generation launch loop inside initialize_data_structures / run /
 * finalize_data_structures of an infrastructure::Algorithm<2, cell_t>,
 * src/infrastructure/algorithm.hpp:37-69), so that a plug-in class compiled
 * with a plain C++ compiler can drive the GPU through plain pointers and sizes.
 *
 * Word layout ("rows"), as the reference's FullyPackedRows policy
 * (src/algorithms/_shared/bitwise/bit_modes.hpp:127-163,
 *  src/algorithms/_shared/bitwise/general_bit_grid.hpp:217-246):
 *   word (wx, y) is at index y * (X / W) + wx, bit k (value 1 << k) is cell
 *   (W * wx + k, y), for W = word_bits = 32 or 64. On a little-endian host the
 *   two layouts are the same bytes, so W only fixes the width granularity and
 *   the unit of the word counts below.
 *
 * Conventions:
 *   - every call returns GOL_OK (0) or an error code; gol_last_error() gives
 *     the calling thread's last message;
 *   - the caller owns host buffers, the library owns device buffers;
 *   - one handle is used by one host thread at a time;
 *   - gol_step only enqueues work on the handle's stream; gol_download and
 *     gol_sync wait for it.
 */
#ifndef GOLHIP_H
#define GOLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The C++ plug-in wrapper maps them onto the exceptions the
 * reference throws: */
enum gol_status {
    GOL_OK = 0,
    GOL_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument (general_bit_grid.hpp:208-215) */
    GOL_ERR_RUNTIME = 2,          /* std::runtime_error (temporal_rowed.cu:190-206) */
    GOL_ERR_HIP = 3,              /* CudaError raised by CUCH (_shared/cuda-helpers/cuch.hpp:42-52) */
    GOL_ERR_OUT_OF_MEMORY = 4,    /* cudaMalloc failure inside CUCH */
    GOL_ERR_NOT_FOUND = 5,        /* unknown job / handle */
    GOL_ERR_BUSY = 6              /* resource temporarily full (admission) */
};

/* Which kernel family advances the grid. */
enum gol_kernel {
    GOL_KERNEL_AUTO = 0,     /* temporal tiles, shape and depth from the built-in table */
    GOL_KERNEL_PLAIN = 1,    /* one generation per launch
                                (reference: cuda-naive-bitwise/cuda_naive_bitwise_kernel.cu:32-79) */
    GOL_KERNEL_TEMPORAL = 2, /* blocks of T generations on LDS-exchanged register tiles; when one
                                tile spans the grid's width and all tiles fit on the device at
                                once, a single persistent launch runs every block
                                (reference: cuda-temporal-rowed/temporal_rowed.cu:33-229) */
    GOL_KERNEL_ACTIVE = 3    /* change tracking over a device-side active-tile list: a tile of
                                32 * lane_dwords cells (lane_dwords 2, 4 or 8; 0: 2, moving to 4
                                while a step's blocks list many tiles, and back) x
                                (64 - 2 * temporal_steps) rows is advanced only when a cell
                                within temporal_steps cells of it changed over the last block of
                                temporal_steps (1..21, default 16) generations; band_rows and
                                waves do not apply
                                (reference: cuda-naive-local-one-cell/cuda_local_one_cell.cu:53-242) */
};

/* Bit layout of the words passed to gol_upload / gol_upload_device /
 * gol_download (reference: src/algorithms/_shared/bitwise/bit_modes.hpp). The
 * engine advances the row layout; other layouts are converted on the device
 * at upload and download. */
enum gol_layout {
    GOL_LAYOUT_ROWS = 0,    /* word (wx, y) at y*(X/W)+wx, bit k = cell (W*wx+k, y); W = 32 or 64
                               (FullyPackedRows, bit_modes.hpp:127-163) */
    GOL_LAYOUT_COLUMNS = 1, /* word (x, wy) at wy*X+x, bit k = cell (x, W*wy+k); W = 16, 32 or 64,
                               Y a multiple of W (BitColumns, bit_modes.hpp:10-32) */
    GOL_LAYOUT_TILES = 2    /* 4x4 / 8x4 / 8x8 cells in a 16 / 32 / 64-bit word (wx, wy) at
                               wy*(X/XB)+wx, cell (tx, ty) in bit W-1-(ty*XB+tx); Y a multiple of
                               the tile height (BitTile, bit_modes.hpp:37-125) */
};

typedef struct gol_config {
    uint64_t x_cells;        /* grid width in cells, a multiple of 32 and of word_bits (rows) */
    uint64_t y_cells;        /* grid height in cells (rows) */
    uint32_t word_bits;      /* 32 or 64 (rows); 16, 32 or 64 (columns, tiles) */
    uint32_t kernel;         /* enum gol_kernel */
    uint32_t temporal_steps; /* generations per temporal launch; 0 = chosen by the library
                                (reference: ExperimentParams::temporal_steps,
                                 src/infrastructure/experiment_params.hpp:62) */
    uint32_t lane_dwords;    /* dwords per lane 1/2/4/8; 0 = chosen by the library */
    uint32_t band_rows;      /* rows per wavefront band; 0 = chosen by the library */
    uint32_t waves;          /* wavefronts per workgroup; 0 = chosen by the library */
    int32_t device;          /* HIP device ordinal; -1 = current device */
    uint32_t flags;          /* GOL_FLAG_* */
    uint32_t layout;         /* enum gol_layout of the words crossing the boundary; X is a
                                multiple of 32 in every layout */
} gol_config;

/* gol_config.flags */
#define GOL_FLAG_NO_PERSISTENT 1u /* temporal kernel: one launch per block of generations, never the
                                     persistent slab launch */
#define GOL_FLAG_SLAB_WAIT_PROBE 2u /* testing the failure path: in the persistent launch every slab
                                       waits, with a short spin limit, for a block count no slab
                                       publishes, so each wait ends by its time-out and the launch
                                       stops; the next gol_sync / gol_download reports
                                       GOL_ERR_RUNTIME and the handle is usable after an upload */
#define GOL_FLAG_NATIVE_LAYOUT 4u /* advance the layout's own words, without converting to rows: 8x8 tiles
                                     (GOL_LAYOUT_TILES, 64-bit words) or columns (GOL_LAYOUT_COLUMNS, 32- or
                                     64-bit words), with GOL_KERNEL_PLAIN (one generation per launch;
                                     reference: cuda_naive_bitwise_kernel.cu:32-79 on BitTile / BitColumns)
                                     or GOL_KERNEL_TEMPORAL / AUTO (temporal_steps generations per launch,
                                     1..16, default 16 for tiles and 8 for columns; reference:
                                     cuda-temporal-simple-tiled/temporal_simple_tiled.cu:41-132). Without
                                     the flag those layouts are converted to rows on the device, which is
                                     faster: the row network is the shortest */

typedef struct gol_info {
    uint32_t kernel;          /* resolved enum gol_kernel */
    uint32_t lane_dwords;     /* resolved tile shape */
    uint32_t band_rows;
    uint32_t waves;
    uint32_t temporal_steps;  /* generations per launch (per block for the persistent launch) */
    uint32_t tiles;           /* workgroups per launch */
    uint32_t persistent;      /* 1: one cooperative launch per gol_step, slabs exchange halos in memory */
    uint32_t layout;          /* enum gol_layout of uploads and downloads */
    uint64_t launches;        /* kernel launches issued so far */
    uint64_t generations;     /* generations enqueued so far (actually_performed_iterations) */
    uint64_t active_tiles_total; /* GOL_KERNEL_ACTIVE: tiles computed, summed over blocks (syncs) */
    uint64_t tracked_blocks;     /* GOL_KERNEL_ACTIVE: blocks of temporal_steps generations run */
    uint64_t device_bytes;    /* device memory owned by the handle */
    uint32_t native_layout;   /* 1: the layout's own words are advanced (GOL_FLAG_NATIVE_LAYOUT) */
    uint32_t workgroups;      /* GOL_KERNEL_ACTIVE: workgroups per launch (one wavefront per tile) */
    uint32_t tile_width_switches; /* GOL_KERNEL_ACTIVE with lane_dwords 0: changes between 64- and
                                     128-cell tiles so far (lane_dwords reports the width in use) */
} gol_info;

typedef struct gol_handle gol_handle;

/* Allocate device state for one grid (reference: cudaMalloc in
 * initialize_data_structures, e.g. cuda-temporal-rowed/temporal_rowed.hpp:34-44). */
int gol_create(const gol_config* config, gol_handle** out);

/* Copy n_words host words (word_bits each) in as the current state
 * (reference: cudaMemcpy H2D in initialize_data_structures). */
int gol_upload(gol_handle* h, const void* words, size_t n_words);

/* Copy n_words words already in device memory (e.g. another handle's
 * gol_device_state) in as the current state, ordered on the handle's stream. */
int gol_upload_device(gol_handle* h, const void* device_words, size_t n_words);

/* Enqueue `generations` generations (0 = identity)
 * (reference: Algorithm::run(iterations), src/infrastructure/algorithm.hpp:49). */
int gol_step(gol_handle* h, uint64_t generations);

/* Same as gol_step, then wait and return the GPU time of the launches in ms,
 * measured with HIP events on the handle's stream. */
int gol_step_timed(gol_handle* h, uint64_t generations, double* gpu_ms);

/* Wait for the handle's stream, then copy the current state out
 * (reference: cudaMemcpy D2H in finalize_data_structures,
 *  e.g. cuda-temporal-rowed/temporal_rowed.hpp:53-62). */
int gol_download(gol_handle* h, void* words, size_t n_words);

/* Wait for all work enqueued on the handle
 * (reference: the cudaDeviceSynchronize of Timer::measure_ms, src/infrastructure/timer.hpp:15-26). */
int gol_sync(gol_handle* h);

/* Device pointer to the current state (valid until the next gol_step) and
 * the hipStream_t the handle launches on. */
int gol_device_state(gol_handle* h, void** device_words);
int gol_get_stream(gol_handle* h, void** stream);

int gol_get_info(gol_handle* h, gol_info* info);

/* Release the device state (reference: cudaFree in finalize_data_structures). */
void gol_destroy(gol_handle* h);

/* Advance a grid held in caller-owned device memory without allocating
 * anything: no hipMalloc, hipFree or allocating library routine. No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). It is gol_step's persistent slab launch (GOL_KERNEL_TEMPORAL)
 * with the launch's scratch, the exchanged halo rows, per-slab block
 * counters and an error word, in a file-scope device array of 16 MiB + 2 KiB
 * that each GPU the library runs on holds until the process ends.
 *   device_in  rows-layout words of the x_cells x y_cells grid (read only);
 *   device_out receives the grid after `generations` generations; it must not
 *              overlap device_in;
 *   stream     hipStream_t to run on (NULL: the null stream), on the current
 *              device; the call waits for it before returning.
 * Limits (GOL_ERR_INVALID_ARGUMENT beyond them): rows of at most 16384 cells,
 * grids the persistent slab launch takes on the device (at least two slabs,
 * every slab resident at once, at most 256), at most 65536 blocks of
 * generations per call (524288 at the usual 8 per block: call again with the
 * output as the input for more). Calls that overlap in time on one GPU share
 * its scratch and race: serialise them. A stopped launch (a slab's wait ran
 * out) returns GOL_ERR_RUNTIME. */
int gol_step_device_noalloc(const void* device_in, void* device_out, uint64_t x_cells, uint64_t y_cells,
                            uint64_t generations, void* stream);

/* ---- Batches: many independent grids of uneven size, each with its own
 * generation count (no reference counterpart: the reference runs one grid per
 * process, src/infrastructure/experiment_manager.hpp:235-270). Each grid's
 * result is bit-identical to running it alone through gol_step. ---- */
typedef struct gol_batch gol_batch;

/* gol_batch_config.flags */
#define GOL_BATCH_PERSISTENT 1u /* one persistent launch per kernel shape present, its workgroups draining a
                                   device work queue of (grid, tile, generations) items; a split grid's round
                                   waits on the device for the grid's previous round. Without it: one launch
                                   per round and kernel shape. */
#define GOL_BATCH_WAIT_PROBE 2u /* testing the failure path of GOL_BATCH_PERSISTENT: every item that continues
                                   a grid waits (256 polls at most) for a count no item reaches, so the run
                                   stops and gol_batch_run reports GOL_ERR_RUNTIME */
#define GOL_BATCH_SELECT_PLAN 4u /* the first gol_batch_run plans the batch twice, with the autotuner's
                                    measured per-size-class choices and with the built-in model, times one
                                    run of each and keeps the faster plan for every later run */

typedef struct gol_batch_config {
    int32_t device;             /* HIP device ordinal; -1 = current device */
    uint32_t round_generations; /* generations per round for grids split into tiles (1..32); 0 = 32 */
    uint32_t use_graph;         /* 1: record all launches once into a hipGraph and replay it */
    uint32_t flags;             /* GOL_BATCH_* */
} gol_batch_config;

typedef struct gol_batch_info {
    uint32_t grids;
    uint32_t rounds;            /* rounds of work items (persistent: the most rounds of one grid) */
    uint64_t items;             /* work items (grid tile x round) */
    uint64_t launches;          /* kernel launches per run (inside the graph when use_graph); persistent:
                                   one per kernel shape + one that clears the queue counters */
    uint64_t device_bytes;
    uint64_t cell_updates;      /* sum over grids of cells x generations */
    uint32_t plan_builtin;      /* GOL_BATCH_SELECT_PLAN: 1 if the built-in model's plan was kept */
    uint32_t reserved;
    double plan_tuned_ms;       /* GOL_BATCH_SELECT_PLAN: timed run under the measured choices */
    double plan_builtin_ms;     /* GOL_BATCH_SELECT_PLAN: timed run under the built-in model */
} gol_batch_info;

/* config may be NULL (current device, 32 generations per round, graph on). */
int gol_batch_create(const gol_batch_config* config, gol_batch** out);
/* Add a grid and its initial state (host words); returns its id. Only before
 * the first gol_batch_run. */
int gol_batch_add(gol_batch* b, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t generations,
                  const void* words, uint32_t* grid_id);
/* Advance every grid by its generation count from its initial state and wait;
 * gpu_ms (may be NULL) receives the GPU time of the run. Running again
 * restarts from the initial states. */
int gol_batch_run(gol_batch* b, double* gpu_ms);
/* Restore every grid's initial state (done by gol_batch_run before a re-run). */
int gol_batch_reset(gol_batch* b);
/* Copy grid `grid_id`'s state out (its final state after a run). */
int gol_batch_fetch(gol_batch* b, uint32_t grid_id, void* words, size_t n_words);
int gol_batch_get_info(gol_batch* b, gol_batch_info* info);
void gol_batch_destroy(gol_batch* b);

/* ---- Autotuner for batches and servers (no reference counterpart: the
 * reference picks kernel parameters by hand per run, ExperimentParams
 * thread_block_size / temporal_steps, experiment_params.hpp:21-65). For each
 * representative grid size it times every compiled tile shape (dwords per
 * lane 1/2/4/8, i.e. 32- to 256-bit lane words; rows per wavefront;
 * wavefronts per workgroup) and, for grids split into tiles, R, 3R/4, R/2,
 * 3R/8, R/4 and R/8 generations per round, plus, for grids at most 512 cells
 * wide and 512 rows tall, the narrow kernel (one grid row per lane), on a
 * batch of synthetic grids of that size, and
 * installs the fastest for the size class (power-of-two classes of the row
 * width and the row count). The choices apply to every batch planned
 * afterwards in the process, until gol_batch_tuning_reset. Servers keep the
 * built-in choices: the measured ones favour a batch's throughput with
 * shallower rounds, and a served job's latency grows with its rounds. ---- */
typedef struct gol_tune_config {
    int32_t device;              /* HIP device ordinal; -1 = current device */
    uint32_t round_generations;  /* R, as in gol_batch_config (1..32); 0 = 32 */
    uint64_t generations;        /* generations each measured grid advances; 0 = 4 R */
    uint64_t cells_per_measure;  /* cells in one measured batch (grids of the size); 0 = 2^28 */
} gol_tune_config;

typedef struct gol_tune_entry {
    uint64_t x_cells, y_cells;          /* in: a representative grid of the size class */
    uint32_t lane_dwords, band_rows, waves;  /* out: the chosen tile shape; band_rows 0 = the narrow
                                                kernel (lane_dwords dwords per row, waves x 64 rows) */
    uint32_t generations_per_round;     /* out: for split grids; 0 when one tile spans the grid */
    uint32_t candidates;                /* out: choices timed */
    uint32_t grids_measured;            /* out: grids per measured batch */
    double tuned_cell_updates_per_s;    /* out: rate of the chosen shape */
    double builtin_cell_updates_per_s;  /* out: rate of the built-in choice */
} gol_tune_entry;

/* Tune n_entries size classes (device memory: 3 x cells_per_measure / 8 bytes
 * at a time). */
int gol_batch_autotune(const gol_tune_config* config, gol_tune_entry* entries, uint32_t n_entries);
/* Forget every measured choice: batches use the built-in table and cost model. */
int gol_batch_tuning_reset(void);

/* ---- Server: continuous batching of simulation jobs (no reference
 * counterpart). A scheduler thread inside the library advances all resident
 * jobs in rounds; between rounds it admits queued jobs (arrival order) whose
 * packed state fits the HBM budget, so requests that arrive while others run
 * join the running batch. Results are bit-identical to running each job alone.
 * gol_server_submit / _poll / _wait / _get_stats may be called from any
 * thread; gol_server_destroy only after every wait has returned. ---- */
typedef struct gol_server gol_server;

typedef struct gol_server_config {
    int32_t device;             /* HIP device ordinal; -1 = current device */
    uint32_t round_generations; /* generations per round for grids split into tiles (1..32); 0 = 32 */
    uint64_t hbm_budget_bytes;  /* device bytes for resident job state (2 buffers per job); 0 = 8 GiB */
    uint32_t max_jobs;          /* resident jobs at most; 0 = 4096 */
    uint32_t reserved;
} gol_server_config;

enum gol_job_state {
    GOL_JOB_QUEUED = 0,   /* waiting for HBM budget */
    GOL_JOB_RESIDENT = 1, /* state on the device, advancing every round */
    GOL_JOB_DRAINING = 2, /* finished, result being copied out */
    GOL_JOB_DONE = 3,
    GOL_JOB_FAILED = 4    /* e.g. GOL_ERR_OUT_OF_MEMORY: larger than the whole budget */
};

typedef struct gol_job_stats {
    double queue_ms;   /* submit -> admitted */
    double latency_ms; /* submit -> result in host memory */
    uint32_t rounds;   /* rounds the job took part in */
    uint32_t reserved;
} gol_job_stats;

typedef struct gol_server_stats {
    uint64_t submitted, admitted, completed, rejected;
    uint64_t rounds, items;
    uint64_t cell_updates;        /* cells x generations enqueued so far */
    uint64_t resident_bytes;      /* arena bytes in use now */
    uint64_t resident_bytes_peak;
    uint32_t resident_jobs, queued_jobs;
    uint64_t pinned_region_bytes; /* the server's pinned host region for jobs' input and result buffers */
    uint64_t pinned_region_used;  /* bytes of it held by jobs now */
    uint64_t pinned_region_peak;
    uint64_t pinned_own_jobs;     /* jobs that found no room in the region and pinned their own buffers */
} gol_server_stats;

/* config may be NULL (defaults above). Allocates the budget up front. */
int gol_server_create(const gol_server_config* config, gol_server** out);
/* Queue a job: copies the input words; returns at once with its id. */
int gol_server_submit(gol_server* s, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t generations,
                      const void* words, uint64_t* job_id);
/* Current enum gol_job_state of a job not yet waited for. */
int gol_server_poll(gol_server* s, uint64_t job_id, int* state);
/* Block until the job is done, copy its final state into words (may be NULL)
 * and its timings into stats (may be NULL), and forget the job. Returns the
 * job's error code if it failed. */
int gol_server_wait(gol_server* s, uint64_t job_id, void* words, size_t n_words, gol_job_stats* stats);
int gol_server_get_stats(gol_server* s, gol_server_stats* stats);
/* Stop the scheduler (jobs not finished fail) and free everything. */
void gol_server_destroy(gol_server* s);

/* Host-side input generators (no device work), writing packed row words:
 *  gol_fill_random: the reference's RandomOnesZerosDataLoader
 *    (src/infrastructure/data_loader.hpp:57-75): std::mt19937(seed),
 *    uniform_int_distribution<int>(0,1), one draw per cell in y*X+x order;
 *  gol_fill_bernoulli: Bernoulli(p) cells, u = (mt19937_64() >> 11) * 2^-53 < p
 *    (no reference counterpart: the reference only draws p = 0.5);
 *  gol_fill_patches: empty grid with n_patches squares of side `patch`, corner
 *    (rng() % (X-patch+1), rng() % (Y-patch+1)) then Bernoulli(p) cells from the
 *    same mt19937_64(seed) stream, later patches overwriting earlier ones
 *    (no reference counterpart: the mostly-stable change-tracking workload);
 *  gol_fill_lexicon: the reference's pattern-expression loader
 *    (src/infrastructure/gol-lexicon/lexicon.hpp:92-186): an empty grid and
 *    "name[x, y]; ..." Life Lexicon patterns (block, blinker, beehive, glider,
 *    gosper-glider-gun, spacefiller), each pattern's rectangle written at its
 *    top-left corner (x, y), cells beyond the grid skipped. */
int gol_fill_random(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed);
int gol_fill_bernoulli(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed,
                       double p);
int gol_fill_patches(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed,
                     uint32_t n_patches, uint32_t patch, double p);
int gol_fill_lexicon(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, const char* expression);

/* Message of the calling thread's last failed call ("" if none). */
const char* gol_last_error(void);

/* Library version string. */
const char* gol_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GOLHIP_H */
