// Error plumbing and tile-shape defaults shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gol_kernels.hpp"

namespace golhip {

// Record a message for gol_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail_hip(hipError_t e, const char* what);
void clear_error();
const char* last_error();

TileShape default_shape(int64_t x_dw, int64_t y_rows);

// Batch size classes: the measured choice if gol_batch_autotune found one, else the table.
TileShape batch_default_shape(int64_t x_dw, int64_t y_rows);
TileShape batch_shape(int64_t x_dw, int64_t y_rows);

// Who plans a grid's rounds. Both price split grids with loads and stores
// hidden behind other tiles' compute where a shape holds two or more
// workgroups per CU, which favours shallow halos and many rounds. A batch
// replays its rounds from a graph, so a round costs little beyond its work:
// it takes the autotuner's measured choice for the size class when there is
// one, and any halo depth. A server plans every round on the host and a job's
// latency counts its rounds: it keeps the built-in choices and rounds at least
// half as deep as round_generations.
enum class Planner { kBatch, kServer };

// Tile shape, geometry and generations per round of one grid: grids split into
// tiles advance up to round_generations per round, and grids that fit one tile
// advance a chunk of about the same cost.
int plan_batch_grid(int64_t x_dw, int64_t y_rows, int round_generations, Planner planner, TileShape* shape,
                    TileGeometry* geom, int* per_round);
bool shape_compiled(const TileShape& s);

// Time model of a temporal launch (us per generation) and whether the
// persistent slab launch takes a shape and depth for a grid (gol_capi.hip).
double model_us_per_generation(const TileShape& s, const TileGeometry& g, int steps, int occupancy, bool persistent);
bool slab_fits(const TileShape& s, const TileGeometry& g, int64_t x_dw, int steps);

}  // namespace golhip
