// CPU ORACLE — test infrastructure only.
//
// A plain restatement of the reference's Game-of-Life path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker for
// the HIP engine. Nothing in the product links, loads or calls this file.
//
// Functions and the reference code they restate:
//   oracle_load_random      infrastructure/data_loader.hpp:57-75
//                           (RandomOnesZerosDataLoader: std::mt19937(seed),
//                           uniform_int_distribution<int>(0,1), cell i = y*X+x)
//   oracle_load_bernoulli   no reference counterpart (the reference only has
//                           p = 0.5); Bernoulli(p) = (mt19937_64() >> 11) * 2^-53 < p
//   oracle_load_glibc_rand  AN5D/GOL_AN5D.c:35-43 (srand(1985), rand() % 2,
//                           row-major interior) — input of the AN5D alive-count KATs
//   oracle_naive_run_boundary  the naive loop with a boundary switch: dead (the
//                           reference's) or periodic (only for the AN5D KATs,
//                           whose recorded counts come from a wrap-around grid)
//   oracle_naive_run        algorithms/cpu-naive/gol_cpu_naive.hpp:26-78 (loop
//                           order, B3/S23 rule, centre alive == 1) and :92-117
//                           (8-neighbour count, cells outside [0,X)x[0,Y) dead,
//                           neighbour alive = value > 0)
//   oracle_pack_rows /      algorithms/_shared/bitwise/general_bit_grid.hpp:171-246
//   oracle_unpack_rows      with the FullyPackedRows policy, bit_modes.hpp:127-163
//                           (word (wx,y) at y*(X/W)+wx, bit k = cell W*wx+k)
//   oracle_pack_layout /    general_bit_grid.hpp:171-246 for any of the bit
//   oracle_unpack_layout    policies of bit_modes.hpp: rows (FullyPackedRows,
//                           :127-163), columns (BitColumns, :10-32: X_BITS 1,
//                           Y_BITS W, bit y % W) and tiles (BitTile, :37-125:
//                           4x4 / 8x4 / 8x8 cells in 16 / 32 / 64 bits, first
//                           cell in the most significant bit); word (wx, wy) at
//                           wy * (X / X_BITS) + wx
//   oracle_bitpar_run       a second, independent CPU engine (uint64 rows,
//                           3x3-word neighbourhood, like cpu-bitwise-general
//                           algorithms/cpu-bitwise-general/gol_cpu_bitwise_general.hpp:32-76
//                           with the adder family of
//                           _shared/bitwise/bitwise-ops/adder.cuh:17-51);
//                           pinned against oracle_naive_run by the tests and used
//                           only where the naive loop is too slow (16384^2 x 10^4).
//
// Each function is pinned by tests/test_oracle.py: the AN5D alive-count KATs
// (AN5D/GOL_AN5D.c:92-97), the C++ standard's mt19937 and mt19937_64
// 10000th-output values, the loaders' cells against pure-Python restatements
// of both engines and of the two cell mappings,
// exhaustive 3x3 neighbourhoods against the B3/S23 rule table, still lifes,
// oscillators, glider translation, and the D4 symmetries of the grid.

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// infrastructure/data_loader.hpp:60-74
void oracle_load_random(uint8_t* cells, uint64_t x_size, uint64_t y_size, uint64_t seed) {
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_int_distribution<int> dist(0, 1);
    const uint64_t n = x_size * y_size;
    for (uint64_t i = 0; i < n; ++i) {
        cells[i] = static_cast<uint8_t>(dist(rng));
    }
}

uint32_t oracle_mt19937_nth(uint64_t seed, uint64_t n) {
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    uint32_t v = 0;
    for (uint64_t i = 0; i < n; ++i) v = rng();
    return v;
}

// The n-th output of std::mt19937_64(seed), the engine of oracle_load_bernoulli.
uint64_t oracle_mt19937_64_nth(uint64_t seed, uint64_t n) {
    std::mt19937_64 rng(seed);
    uint64_t v = 0;
    for (uint64_t i = 0; i < n; ++i) v = rng();
    return v;
}

void oracle_load_bernoulli(uint8_t* cells, uint64_t n, uint64_t seed, double p) {
    std::mt19937_64 rng(seed);
    for (uint64_t i = 0; i < n; ++i) {
        double u = static_cast<double>(rng() >> 11) * (1.0 / 9007199254740992.0);
        cells[i] = u < p ? 1 : 0;
    }
}

// AN5D/GOL_AN5D.c:35-43: srand(seed); interior cell (i,j) in row-major order = rand() % 2
void oracle_load_glibc_rand(uint8_t* cells, uint64_t n, unsigned seed) {
    srand(seed);
    for (uint64_t i = 0; i < n; ++i) cells[i] = static_cast<uint8_t>(rand() % 2);
}

// algorithms/cpu-naive/gol_cpu_naive.hpp:92-117. With periodic != 0 the
// neighbour coordinates wrap around instead (a torus); the reference has no such
// mode — it exists only to reproduce the AN5D alive-count KATs
// (AN5D/GOL_AN5D.c:92-97), which were produced on a wrap-around grid.
static inline uint64_t count_alive_neighbours(const uint8_t* g, uint64_t x_size, uint64_t y_size,
                                              uint64_t x, uint64_t y, int periodic) {
    uint64_t alive = 0;
    for (int i = -1; i <= 1; ++i) {
        for (int j = -1; j <= 1; ++j) {
            if (i == 0 && j == 0) continue;
            uint64_t xn = x + static_cast<uint64_t>(static_cast<int64_t>(i));
            uint64_t yn = y + static_cast<uint64_t>(static_cast<int64_t>(j));
            if (periodic) {
                xn = (x + x_size + static_cast<uint64_t>(static_cast<int64_t>(i))) % x_size;
                yn = (y + y_size + static_cast<uint64_t>(static_cast<int64_t>(j))) % y_size;
            }
            // unsigned wrap-around of -1 lands far above the size: one test covers both sides
            if (xn >= x_size || yn >= y_size) continue;
            alive += g[yn * x_size + xn] > 0 ? 1 : 0;
        }
    }
    return alive;
}

// algorithms/cpu-naive/gol_cpu_naive.hpp:26-78 (0 generations = identity)
void oracle_naive_run_boundary(uint8_t* cells, uint64_t x_size, uint64_t y_size,
                               uint64_t generations, int periodic) {
    std::vector<uint8_t> other(x_size * y_size);
    uint8_t* source = cells;
    uint8_t* target = other.data();
    for (uint64_t it = 0; it < generations; ++it) {
        for (uint64_t y = 0; y < y_size; ++y) {
            for (uint64_t x = 0; x < x_size; ++x) {
                uint64_t n = count_alive_neighbours(source, x_size, y_size, x, y, periodic);
                uint8_t& out = target[y * x_size + x];
                if (source[y * x_size + x] == 1) {
                    out = (n < 2 || n > 3) ? 0 : 1;
                } else {
                    out = (n == 3) ? 1 : 0;
                }
            }
        }
        uint8_t* t = source;
        source = target;
        target = t;
    }
    if (source != cells) std::memcpy(cells, source, x_size * y_size);
}

// the reference's convention: cells outside the grid are dead
void oracle_naive_run(uint8_t* cells, uint64_t x_size, uint64_t y_size, uint64_t generations) {
    oracle_naive_run_boundary(cells, x_size, y_size, generations, 0);
}

// general_bit_grid.hpp:217-246 (fill_grid) with FullyPackedRows, word_bits 32 or 64
int oracle_pack_rows(const uint8_t* cells, uint64_t x_size, uint64_t y_size, uint32_t word_bits,
                     void* words) {
    if ((word_bits != 32 && word_bits != 64) || x_size % word_bits != 0) return 1;
    const uint64_t wpr = x_size / word_bits;
    for (uint64_t y = 0; y < y_size; ++y) {
        for (uint64_t wx = 0; wx < wpr; ++wx) {
            uint64_t word = 0;
            for (uint32_t k = 0; k < word_bits; ++k) {
                if (cells[y * x_size + wx * word_bits + k]) word |= (uint64_t{1} << k);
            }
            if (word_bits == 64)
                static_cast<uint64_t*>(words)[y * wpr + wx] = word;
            else
                static_cast<uint32_t*>(words)[y * wpr + wx] = static_cast<uint32_t>(word);
        }
    }
    return 0;
}

// general_bit_grid.hpp:171-200 (to_grid)
int oracle_unpack_rows(const void* words, uint64_t x_size, uint64_t y_size, uint32_t word_bits,
                       uint8_t* cells) {
    if ((word_bits != 32 && word_bits != 64) || x_size % word_bits != 0) return 1;
    const uint64_t wpr = x_size / word_bits;
    for (uint64_t y = 0; y < y_size; ++y) {
        for (uint64_t wx = 0; wx < wpr; ++wx) {
            uint64_t word = (word_bits == 64) ? static_cast<const uint64_t*>(words)[y * wpr + wx]
                                              : static_cast<const uint32_t*>(words)[y * wpr + wx];
            for (uint32_t k = 0; k < word_bits; ++k) {
                cells[y * x_size + wx * word_bits + k] = (word >> k) & 1;
            }
        }
    }
    return 0;
}

// Bit policy of a layout (0 rows, 1 columns, 2 tiles) and word width:
// cells per word in x and y, and whether the first cell takes the most
// significant bit (BitTileCommon, bit_modes.hpp:37-56) or the least.
static bool layout_policy(uint32_t layout, uint32_t w, uint32_t* xb, uint32_t* yb, bool* msb_first) {
    if (w != 16 && w != 32 && w != 64) return false;
    if (layout == 0) {
        *xb = w, *yb = 1, *msb_first = false;
    } else if (layout == 1) {
        *xb = 1, *yb = w, *msb_first = false;
    } else if (layout == 2) {
        *xb = w == 16 ? 4 : 8, *yb = w == 64 ? 8 : 4, *msb_first = true;
    } else {
        return false;
    }
    return true;
}

static void store_word(void* words, uint64_t i, uint32_t w, uint64_t v) {
    if (w == 64) static_cast<uint64_t*>(words)[i] = v;
    else if (w == 32) static_cast<uint32_t*>(words)[i] = static_cast<uint32_t>(v);
    else static_cast<uint16_t*>(words)[i] = static_cast<uint16_t>(v);
}

static uint64_t load_word(const void* words, uint64_t i, uint32_t w) {
    if (w == 64) return static_cast<const uint64_t*>(words)[i];
    if (w == 32) return static_cast<const uint32_t*>(words)[i];
    return static_cast<const uint16_t*>(words)[i];
}

// general_bit_grid.hpp:217-246 (fill_grid): per word, the cells in y_bit, x_bit
// order take successive bits from the policy's first mask.
int oracle_pack_layout(const uint8_t* cells, uint64_t x_size, uint64_t y_size, uint32_t layout, uint32_t word_bits,
                       void* words) {
    uint32_t xb, yb;
    bool msb;
    if (!layout_policy(layout, word_bits, &xb, &yb, &msb) || x_size % xb || y_size % yb) return 1;
    const uint64_t wx_n = x_size / xb;
    for (uint64_t wy = 0; wy < y_size / yb; ++wy) {
        for (uint64_t wx = 0; wx < wx_n; ++wx) {
            uint64_t word = 0;
            uint32_t bit = msb ? word_bits - 1 : 0;
            for (uint32_t y_bit = 0; y_bit < yb; ++y_bit) {
                for (uint32_t x_bit = 0; x_bit < xb; ++x_bit) {
                    if (cells[(wy * yb + y_bit) * x_size + wx * xb + x_bit]) word |= uint64_t{1} << bit;
                    bit = msb ? bit - 1 : bit + 1;
                }
            }
            store_word(words, wy * wx_n + wx, word_bits, word);
        }
    }
    return 0;
}

// general_bit_grid.hpp:171-200 (to_grid)
int oracle_unpack_layout(const void* words, uint64_t x_size, uint64_t y_size, uint32_t layout, uint32_t word_bits,
                         uint8_t* cells) {
    uint32_t xb, yb;
    bool msb;
    if (!layout_policy(layout, word_bits, &xb, &yb, &msb) || x_size % xb || y_size % yb) return 1;
    const uint64_t wx_n = x_size / xb;
    for (uint64_t wy = 0; wy < y_size / yb; ++wy) {
        for (uint64_t wx = 0; wx < wx_n; ++wx) {
            const uint64_t word = load_word(words, wy * wx_n + wx, word_bits);
            uint32_t bit = msb ? word_bits - 1 : 0;
            for (uint32_t y_bit = 0; y_bit < yb; ++y_bit) {
                for (uint32_t x_bit = 0; x_bit < xb; ++x_bit) {
                    cells[(wy * yb + y_bit) * x_size + wx * xb + x_bit] = (word >> bit) & 1;
                    bit = msb ? bit - 1 : bit + 1;
                }
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Bit-parallel engine on rows-64 words. One uint64 holds 64 cells of a row,
// bit k = cell 64*wx + k. Each output word is computed from its 3x3 word
// neighbourhood (zero outside the grid) by a carry-save count of the eight
// neighbour planes, then the B3/S23 rule on the 4-bit count.
// ---------------------------------------------------------------------------

static inline void csa(uint64_t a, uint64_t b, uint64_t c, uint64_t& sum, uint64_t& carry) {
    uint64_t t = a ^ b;
    sum = t ^ c;
    carry = (a & b) | (t & c);
}

static inline uint64_t next_word(uint64_t lt, uint64_t ct, uint64_t rt, uint64_t lc, uint64_t cc,
                                 uint64_t rc, uint64_t lb, uint64_t cb, uint64_t rb) {
    // neighbour planes: value of the neighbour cell at every bit position
    uint64_t n_w = (cc << 1) | (lc >> 63);
    uint64_t n_e = (cc >> 1) | (rc << 63);
    uint64_t n_n = ct;
    uint64_t n_s = cb;
    uint64_t n_nw = (ct << 1) | (lt >> 63);
    uint64_t n_ne = (ct >> 1) | (rt << 63);
    uint64_t n_sw = (cb << 1) | (lb >> 63);
    uint64_t n_se = (cb >> 1) | (rb << 63);
    // 8 one-bit inputs -> 4-bit count (c8 c4 c2 c1)
    uint64_t s1, k1, s2, k2, s3, k3, s4, k4, s5, k5, s6, k6;
    csa(n_w, n_e, n_n, s1, k1);
    csa(n_s, n_nw, n_ne, s2, k2);
    uint64_t s_a = n_sw ^ n_se, k_a = n_sw & n_se;
    csa(s1, s2, s_a, s3, k3);  // s3 = bit 1 of count, k3 carries weight 2
    csa(k1, k2, k_a, s4, k4);  // weight-2 inputs
    csa(s4, k3, 0, s5, k5);
    (void)s6;
    (void)k6;
    uint64_t c1 = s3;
    uint64_t c2 = s5;
    uint64_t c4 = k4 ^ k5;
    uint64_t c8 = k4 & k5;
    // count == 3 -> alive; count == 2 -> keep
    uint64_t two_or_three = c2 & ~c4 & ~c8;
    return two_or_three & (c1 | cc);
}

void oracle_bitpar_run(uint64_t* words, uint64_t x_size, uint64_t y_size, uint64_t generations,
                       int threads) {
    const uint64_t wpr = x_size / 64;
    const uint64_t n = wpr * y_size;
    std::vector<uint64_t> other(n);
    uint64_t* src = words;
    uint64_t* dst = other.data();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    std::vector<uint64_t> zero_row(wpr + 2, 0);
    for (uint64_t g = 0; g < generations; ++g) {
#pragma omp parallel
        {
            // each thread pads the three rows it reads with one zero word per side
            std::vector<uint64_t> pad(3 * (wpr + 2), 0);
#pragma omp for schedule(static)
            for (int64_t yi = 0; yi < static_cast<int64_t>(y_size); ++yi) {
                const uint64_t y = static_cast<uint64_t>(yi);
                uint64_t* pt = pad.data();
                uint64_t* pc = pt + (wpr + 2);
                uint64_t* pb = pc + (wpr + 2);
                std::memcpy(pt + 1, (y > 0) ? src + (y - 1) * wpr : zero_row.data(), wpr * 8);
                std::memcpy(pc + 1, src + y * wpr, wpr * 8);
                std::memcpy(pb + 1, (y + 1 < y_size) ? src + (y + 1) * wpr : zero_row.data(), wpr * 8);
                uint64_t* out = dst + y * wpr;
                for (uint64_t wx = 1; wx <= wpr; ++wx) {
                    out[wx - 1] = next_word(pt[wx - 1], pt[wx], pt[wx + 1], pc[wx - 1], pc[wx], pc[wx + 1],
                                            pb[wx - 1], pb[wx], pb[wx + 1]);
                }
            }
        }
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != words) std::memcpy(words, src, n * sizeof(uint64_t));
}

}  // extern "C"
