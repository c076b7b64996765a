// gol_step_device_noalloc: the persistent slab launch over a grid in
// caller-owned device memory, with no device allocation. The slab launch's
// scratch (exchanged halo rows, per-slab block counters, the error word) lives
// in a file-scope device array instead of per-handle allocations; each device
// holds its own copy of it, whose address is looked up once per device.

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/golhip.h"
#include "gol_capi_internal.hpp"
#include "gol_kernels.hpp"

namespace golhip {

// Scratch bounds: rows of at most 16384 cells (the slab kernels' 8 dwords per
// lane), at most 256 slabs, at most kSlabMaxSteps (8) generations per block.
constexpr int64_t kNoAllocMaxXDwords = 512;
constexpr size_t kNoAllocMaxSlabs = 256;
// [2 parities][slabs][2 sides][T rows][32 * 8 dword pairs]: 16 MiB
constexpr size_t kNoAllocXbufPairs = 2 * kNoAllocMaxSlabs * 2 * kSlabMaxSteps * 32 * 8;
// then one block counter per slab and the error word (a 64-bit slot)
constexpr size_t kNoAllocWords = kNoAllocXbufPairs + kNoAllocMaxSlabs + 1;

static __device__ __align__(256) unsigned long long scratch[kNoAllocWords];

namespace {

// Device address of `scratch` per device ordinal (host code never uses the
// array's own name as a device pointer).
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch[kNoAllocMaxDevices] = {};

int device_scratch(int dev, unsigned long long** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(GOL_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch[dev];
    return GOL_OK;
}

// Among the shapes and depths the persistent slab launch takes for this grid,
// the one the engine's time model prices lowest.
bool choose_slab(int64_t x_dw, int64_t y_rows, TileShape* shape, int* steps, TileGeometry* geom) {
    double best = 1e300;
    for (int i = 0; i < kNumTileShapes; ++i) {
        const TileShape& s = kTileShapes[i];
        for (int t = 2; t <= kSlabMaxSteps; t += 2) {
            TileGeometry g;
            const char* why = nullptr;
            if (!plan_tiles(s, x_dw, y_rows, t, &g, &why) || !slab_fits(s, g, x_dw, t)) continue;
            const double us = model_us_per_generation(s, g, t, 1, true);
            if (us < best) {
                best = us;
                *shape = s;
                *steps = t;
                *geom = g;
            }
        }
    }
    return best < 1e300;
}

}  // namespace
}  // namespace golhip

using namespace golhip;

extern "C" int gol_step_device_noalloc(const void* device_in, void* device_out, uint64_t x_cells, uint64_t y_cells,
                                       uint64_t generations, void* stream) {
    clear_error();
    if (!device_in || !device_out) return fail(GOL_ERR_INVALID_ARGUMENT, "null device buffer");
    if (x_cells == 0 || y_cells == 0 || x_cells % 32 != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions X must be a positive multiple of 32");
    const int64_t x_dw = static_cast<int64_t>(x_cells / 32);
    if (x_dw > kNoAllocMaxXDwords)
        return fail(GOL_ERR_INVALID_ARGUMENT, "gol_step_device_noalloc takes rows of at most %lld cells, got %llu",
                    static_cast<long long>(32 * kNoAllocMaxXDwords), static_cast<unsigned long long>(x_cells));
    if (y_cells > (1u << 20)) return fail(GOL_ERR_INVALID_ARGUMENT, "grid too tall");
    const int64_t y_rows = static_cast<int64_t>(y_cells);
    const size_t bytes = static_cast<size_t>(x_dw) * static_cast<size_t>(y_rows) * 4;
    const char* in = static_cast<const char*>(device_in);
    const char* out = static_cast<const char*>(device_out);
    if (in < out + bytes && out < in + bytes)
        return fail(GOL_ERR_INVALID_ARGUMENT, "device_in and device_out overlap");

    TileShape shape{};
    TileGeometry geom{};
    int steps = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(GOL_ERR_INVALID_ARGUMENT, "device %d out of range", dev);
    if (!choose_slab(x_dw, y_rows, &shape, &steps, &geom))
        return fail(GOL_ERR_INVALID_ARGUMENT,
                    "a %llu x %llu grid is not one the persistent slab launch takes on this device (rows of at "
                    "most 16384 cells, at least two and at most as many slabs as the device holds at once)",
                    static_cast<unsigned long long>(x_cells), static_cast<unsigned long long>(y_cells));
    const size_t pairs = 2ull * static_cast<size_t>(geom.tiles_y) * 2 * static_cast<size_t>(steps) * 32 *
                         static_cast<size_t>(shape.lane_dwords);
    if (static_cast<size_t>(geom.tiles_y) > kNoAllocMaxSlabs || pairs > kNoAllocXbufPairs)
        return fail(GOL_ERR_INVALID_ARGUMENT,
                    "the grid needs %d slabs and %zu exchange pairs, beyond the scratch maximum of %zu slabs and "
                    "%zu pairs (16 MiB)",
                    geom.tiles_y, pairs, kNoAllocMaxSlabs, kNoAllocXbufPairs);
    const uint64_t max_generations = 65536ull * static_cast<uint64_t>(steps);
    if (generations > max_generations)
        return fail(GOL_ERR_INVALID_ARGUMENT,
                    "at most %llu generations per call for this grid (65536 blocks of %d); call again with the "
                    "output as the input",
                    static_cast<unsigned long long>(max_generations), steps);

    unsigned long long* base = nullptr;
    int rc = device_scratch(dev, &base);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (generations == 0) {
        if ((e = hipMemcpyAsync(device_out, device_in, bytes, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return fail_hip(e, "gol_step_device_noalloc");
        return GOL_OK;
    }
    unsigned long long* flags = base + kNoAllocXbufPairs;
    unsigned int* error = reinterpret_cast<unsigned int*>(flags + kNoAllocMaxSlabs);
    // counters and the error word start at zero (a fill kernel, as every call
    // shares them); the exchange rows are written before they are read
    e = launch_zero(reinterpret_cast<unsigned int*>(flags), static_cast<int>(2 * (kNoAllocMaxSlabs + 1)), st);
    if (e != hipSuccess) return fail_hip(e, "gol_step_device_noalloc: clearing the counters");
    const uint64_t T = static_cast<uint64_t>(steps);
    SlabArgs a;
    a.src = static_cast<const uint32_t*>(device_in);
    a.dst = static_cast<uint32_t*>(device_out);
    a.geo = to_device_geo(geom, x_dw, y_rows);
    a.steps = steps;
    a.blocks = static_cast<int>(generations / T);
    a.last_gens = static_cast<int>(generations % T);
    a.tiles_y = geom.tiles_y;
    a.xbuf = base;
    a.flags = flags;
    a.epoch = 0;
    a.error = error;
    a.spin_limit = 0;
    a.wait_ahead = 0;
    if ((e = launch_slab(shape, a, st)) != hipSuccess) return fail_hip(e, "gol_step_device_noalloc: slab launch");
    unsigned int err = 0;
    if ((e = hipMemcpyAsync(&err, error, sizeof(err), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return fail_hip(e, "gol_step_device_noalloc");
    if (err)
        return fail(GOL_ERR_RUNTIME, "persistent launch stopped: a slab waited too long for its neighbours; "
                                     "device_out is undefined");
    return GOL_OK;
}
