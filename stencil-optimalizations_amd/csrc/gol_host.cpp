// Host-side input generators, written straight into packed row words.
//
//   gol_fill_random     the reference's RandomOnesZerosDataLoader
//                       (src/infrastructure/data_loader.hpp:57-75): std::mt19937(seed)
//                       and uniform_int_distribution<int>(0,1), one draw per cell in
//                       y*X+x order; packed like general_bit_grid.hpp:217-246.
//   gol_fill_bernoulli  Bernoulli(p) cells for p != 0.5 (no reference counterpart):
//                       u = (mt19937_64() >> 11) * 2^-53, alive iff u < p, y*X+x order.
//   gol_fill_patches    the mostly-stable workload: an empty grid with n square
//                       patches of side `patch` at positions drawn from
//                       mt19937_64(seed) (px then py, each uniform in [0, X-patch] /
//                       [0, Y-patch] as rng() % (range+1)); each patch's cells are
//                       Bernoulli(p) drawn from the same stream in row-major order; later
//                       patches overwrite earlier ones.
//   gol_fill_lexicon    the reference's pattern-expression loader
//                       (src/infrastructure/gol-lexicon/lexicon.hpp:92-186), shared with
//                       the driver through host/include/stencils/lexicon.hpp.

#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/golhip.h"
#include "../host/include/stencils/lexicon.hpp"
#include "gol_capi_internal.hpp"

using namespace golhip;

namespace {

int check_dims(const void* words, uint64_t x, uint64_t y, uint32_t word_bits) {
    if (!words) return fail(GOL_ERR_INVALID_ARGUMENT, "words is null");
    if (word_bits != 32 && word_bits != 64) return fail(GOL_ERR_INVALID_ARGUMENT, "word_bits must be 32 or 64");
    if (x == 0 || y == 0 || x % word_bits != 0)
        return fail(GOL_ERR_INVALID_ARGUMENT, "Grid dimensions X must be a multiple of %u", word_bits);
    return GOL_OK;
}

inline void set_cell(void* words, uint64_t x_size, uint32_t word_bits, uint64_t x, uint64_t y, bool alive) {
    // dword view: identical bytes for 32- and 64-bit words on a little-endian host
    uint32_t* dw = static_cast<uint32_t*>(words);
    (void)word_bits;
    const uint64_t idx = y * (x_size / 32) + x / 32;
    const uint32_t bit = 1u << (x % 32);
    dw[idx] = alive ? (dw[idx] | bit) : (dw[idx] & ~bit);
}

inline double unit_double(std::mt19937_64& rng) {
    return static_cast<double>(rng() >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" {

int gol_fill_random(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed) {
    clear_error();
    int rc = check_dims(words, x_cells, y_cells, word_bits);
    if (rc) return rc;
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_int_distribution<int> dist(0, 1);
    uint32_t* dw = static_cast<uint32_t*>(words);
    const uint64_t n_dw = x_cells / 32 * y_cells;
    for (uint64_t i = 0; i < n_dw; ++i) {
        uint32_t v = 0;
        for (int k = 0; k < 32; ++k) v |= static_cast<uint32_t>(dist(rng) != 0) << k;
        dw[i] = v;
    }
    return GOL_OK;
}

int gol_fill_bernoulli(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed,
                       double p) {
    clear_error();
    int rc = check_dims(words, x_cells, y_cells, word_bits);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    uint32_t* dw = static_cast<uint32_t*>(words);
    const uint64_t n_dw = x_cells / 32 * y_cells;
    for (uint64_t i = 0; i < n_dw; ++i) {
        uint32_t v = 0;
        for (int k = 0; k < 32; ++k) v |= static_cast<uint32_t>(unit_double(rng) < p) << k;
        dw[i] = v;
    }
    return GOL_OK;
}

int gol_fill_patches(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, uint64_t seed,
                     uint32_t n_patches, uint32_t patch, double p) {
    clear_error();
    int rc = check_dims(words, x_cells, y_cells, word_bits);
    if (rc) return rc;
    if (patch == 0 || patch > x_cells || patch > y_cells)
        return fail(GOL_ERR_INVALID_ARGUMENT, "patch side must be in [1, min(X, Y)]");
    std::memset(words, 0, x_cells / 8 * y_cells);
    std::mt19937_64 rng(seed);
    for (uint32_t i = 0; i < n_patches; ++i) {
        const uint64_t px = rng() % (x_cells - patch + 1);
        const uint64_t py = rng() % (y_cells - patch + 1);
        for (uint32_t dy = 0; dy < patch; ++dy) {
            for (uint32_t dx = 0; dx < patch; ++dx) {
                set_cell(words, x_cells, word_bits, px + dx, py + dy, unit_double(rng) < p);
            }
        }
    }
    return GOL_OK;
}

int gol_fill_lexicon(void* words, uint64_t x_cells, uint64_t y_cells, uint32_t word_bits, const char* expression) {
    clear_error();
    int rc = check_dims(words, x_cells, y_cells, word_bits);
    if (rc) return rc;
    if (!expression) return fail(GOL_ERR_INVALID_ARGUMENT, "expression is null");
    std::memset(words, 0, x_cells / 8 * y_cells);
    try {
        static const infrastructure::PatternTable table;
        infrastructure::place_patterns(expression, x_cells, y_cells, table, [&](std::size_t x, std::size_t y, bool alive) {
            set_cell(words, x_cells, word_bits, x, y, alive);
        });
    } catch (const std::exception& e) {
        return fail(GOL_ERR_INVALID_ARGUMENT, "%s", e.what());
    }
    return GOL_OK;
}

}  // extern "C"
