// Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// pattern-expression parser, the flag parser, the loaders, the layout packers
// of the plug-ins, and the CPU oracle the tests check against. Every input is
// seeded, so a failure reproduces; malformed inputs must end in the documented
// exceptions, never in an out-of-bounds access or undefined behaviour.
// Host-only: nothing here touches the GPU engine.
#include <cstdint>
#include <cstring>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "stencils/data_loader.hpp"
#include "stencils/experiment_params.hpp"
#include "stencils/gol_hip_algorithms.hpp"
#include "stencils/lexicon.hpp"

extern "C" {
void oracle_load_random(uint8_t* cells, uint64_t x, uint64_t y, uint64_t seed);
void oracle_load_bernoulli(uint8_t* cells, uint64_t n, uint64_t seed, double p);
void oracle_naive_run_boundary(uint8_t* cells, uint64_t x, uint64_t y, uint64_t generations, int periodic);
int oracle_pack_rows(const uint8_t* cells, uint64_t x, uint64_t y, uint32_t word_bits, void* words);
int oracle_unpack_rows(const void* words, uint64_t x, uint64_t y, uint32_t word_bits, uint8_t* cells);
int oracle_pack_layout(const uint8_t* cells, uint64_t x, uint64_t y, uint32_t layout, uint32_t word_bits, void* words);
int oracle_unpack_layout(const void* words, uint64_t x, uint64_t y, uint32_t layout, uint32_t word_bits,
                         uint8_t* cells);
void oracle_bitpar_run(uint64_t* words, uint64_t x, uint64_t y, uint64_t generations, int threads);
}

static int failures = 0;
#define CHECK(cond)                                                                                   \
    do {                                                                                              \
        if (!(cond)) {                                                                                \
            std::cerr << __FILE__ << ":" << __LINE__ << ": CHECK failed: " #cond << std::endl;        \
            ++failures;                                                                               \
        }                                                                                             \
    } while (0)

using Grid = infrastructure::Grid<2, char>;

// Pattern expressions: well-formed ones place cells inside the grid only;
// random strings over the expression alphabet either parse or throw
// std::invalid_argument / std::out_of_range (std::stoul).
static void pattern_expressions() {
    const infrastructure::PatternTable table;
    std::size_t placed = 0;
    infrastructure::place_patterns("spacefiller[0, 0]; glider[30, 30]; gosper-glider-gun[20, 2]", 40, 33, table,
                                   [&](std::size_t x, std::size_t y, bool) {
                                       CHECK(x < 40 && y < 33);
                                       ++placed;
                                   });
    CHECK(placed > 0);
    std::mt19937 rng(7);
    const std::string alphabet = "bgilderockspacefinh-[],; 0123456789-";
    const char* names[] = {"block", "blinker", "glider", "beehive", "spacefiller", "gosper-glider-gun", "none"};
    int parsed = 0, rejected = 0;
    for (int t = 0; t < 4000; ++t) {
        std::string e;
        if (t % 2 == 0) {  // near-valid: names and numbers, sometimes cut short
            const int n = 1 + static_cast<int>(rng() % 4);
            for (int k = 0; k < n; ++k) {
                e += names[rng() % 7];
                e += "[" + std::to_string(rng() % 90) + ", " + std::to_string(rng() % 90) + "]";
                if (k + 1 < n) e += "; ";
            }
            if (rng() % 3 == 0) e.resize(rng() % (e.size() + 1));
        } else {
            const int n = static_cast<int>(rng() % 40);
            for (int k = 0; k < n; ++k) e += alphabet[rng() % alphabet.size()];
        }
        const std::size_t X = 1 + rng() % 70, Y = 1 + rng() % 70;
        try {
            Grid g(X, Y);
            infrastructure::insert_patterns(g, e, table);
            ++parsed;
        } catch (const std::invalid_argument&) {
            ++rejected;
        } catch (const std::out_of_range&) {
            ++rejected;
        }
    }
    CHECK(parsed > 0 && rejected > 0);
}

// Flags: random argument vectors over the known flag names either parse or
// throw std::invalid_argument.
static void flags() {
    const char* keys[] = {"--algorithm", "--grid-dimensions-x", "--grid-dimensions-y", "--iterations",
                          "--max-runtime-seconds", "--validate", "--colorful", "--random-seed", "--data-loader",
                          "--pattern-expression", "--temporal-steps", "--thread-block-size", "--bogus", "x"};
    const char* values[] = {"64", "0", "-3", "1e3", "true", "false", "gol-hip-rows-64", "", "18446744073709551616",
                            "glider[1,1]"};
    std::mt19937 rng(11);
    int parsed = 0, rejected = 0;
    for (int t = 0; t < 4000; ++t) {
        std::vector<std::string> args = {"stencils"};
        const int n = static_cast<int>(rng() % 8);
        for (int k = 0; k < n; ++k) {
            std::string a = keys[rng() % 14];
            switch (rng() % 3) {
                case 0: a += std::string("=") + values[rng() % 10]; break;
                case 1: args.push_back(a); a = values[rng() % 10]; break;
                default: break;
            }
            args.push_back(a);
        }
        std::vector<char*> argv;
        for (auto& s : args) argv.push_back(&s[0]);
        argv.push_back(nullptr);
        try {
            infrastructure::ParamsParser::parse(static_cast<int>(args.size()), argv.data());
            ++parsed;
        } catch (const std::invalid_argument&) {
            ++rejected;
        }
    }
    CHECK(rejected > 0);
    (void)parsed;
}

static void loaders() {
    for (const char* name : {"random-ones-zeros", "lexicon", "always-changing", "zeros"}) {
        infrastructure::ExperimentParams p;
        p.grid_dimensions = {37, 23};
        p.pattern_expression = "glider[35, 21]; block[0, 0]";
        auto g = infrastructure::make_loader<2, char>(name)->load_data(p);
        CHECK(g.size() == 37 * 23);
    }
}

// The plug-ins' packers against the oracle's, every layout and word size, on
// grids of several word multiples; unpacking returns the input.
template <uint32_t Layout, std::size_t Bits>
static void packer(std::size_t X, std::size_t Y, uint64_t seed) {
    using W = algorithms::LayoutWords<Layout, Bits>;
    Grid g(X, Y);
    std::vector<uint8_t> cells(X * Y);
    oracle_load_bernoulli(cells.data(), X * Y, seed, 0.5);
    for (std::size_t i = 0; i < X * Y; ++i) g.data()[i] = static_cast<char>(cells[i]);
    const auto words = W::template pack<char>(g);
    std::vector<typename W::word_type> want(words.size());
    CHECK(oracle_pack_layout(cells.data(), X, Y, Layout, static_cast<uint32_t>(Bits), want.data()) == 0);
    CHECK(words == want);
    const Grid back = W::template unpack<char>(words, X, Y);
    CHECK(std::memcmp(back.data(), g.data(), X * Y) == 0);
    std::vector<uint8_t> cells_back(X * Y);
    CHECK(oracle_unpack_layout(want.data(), X, Y, Layout, static_cast<uint32_t>(Bits), cells_back.data()) == 0);
    CHECK(cells_back == cells);
}

static void packers() {
    uint64_t seed = 1;
    for (std::size_t m : {1, 2, 3}) {
        packer<GOL_LAYOUT_ROWS, 32>(32 * m, 5 * m, seed++);
        packer<GOL_LAYOUT_ROWS, 64>(64 * m, 3 * m, seed++);
        packer<GOL_LAYOUT_COLUMNS, 16>(7 * m, 16 * m, seed++);
        packer<GOL_LAYOUT_COLUMNS, 32>(5 * m, 32 * m, seed++);
        packer<GOL_LAYOUT_COLUMNS, 64>(3 * m, 64 * m, seed++);
        packer<GOL_LAYOUT_TILES, 16>(4 * m + 8, 4 * m, seed++);
        packer<GOL_LAYOUT_TILES, 32>(8 * m, 4 * m + 4, seed++);
        packer<GOL_LAYOUT_TILES, 64>(8 * m + 8, 8 * m, seed++);
    }
    // sizes off the word grid are rejected, not read past
    bool threw = false;
    try {
        Grid g(33, 4);
        algorithms::LayoutWords<GOL_LAYOUT_ROWS, 32>::pack<char>(g);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    std::vector<uint8_t> c(33 * 4);
    std::vector<uint32_t> w(8);
    CHECK(oracle_pack_rows(c.data(), 33, 4, 32, w.data()) != 0);
}

// The oracle's loops on small and ragged grids: the bit-parallel engine equals
// the naive loop, including the one-word and one-row edge cases.
static void oracle_loops() {
    for (auto [X, Y] : std::vector<std::pair<uint64_t, uint64_t>>{{64, 1}, {64, 2}, {128, 3}, {64, 65}, {192, 40}}) {
        std::vector<uint8_t> cells(X * Y);
        oracle_load_random(cells.data(), X, Y, X * 131 + Y);
        std::vector<uint64_t> words(X / 64 * Y);
        CHECK(oracle_pack_rows(cells.data(), X, Y, 64, words.data()) == 0);
        oracle_naive_run_boundary(cells.data(), X, Y, 37, 0);
        oracle_bitpar_run(words.data(), X, Y, 37, 2);
        std::vector<uint8_t> got(X * Y);
        CHECK(oracle_unpack_rows(words.data(), X, Y, 64, got.data()) == 0);
        CHECK(got == cells);
    }
}

int main() {
    pattern_expressions();
    flags();
    loaders();
    packers();
    oracle_loops();
    if (failures) {
        std::cerr << failures << " check(s) failed" << std::endl;
        return 1;
    }
    std::cout << "host code under ASan/UBSan: ok" << std::endl;
    return 0;
}
