// Tests of the C++ plug-in boundary (stencil-optimalizations_amd/host/include/stencils),
// written like the reference's own use of it (experiment_manager.hpp:365-394):
// each plug-in goes through TimedAlgorithm's lifecycle and its result is
// compared with the CPU oracle on the same input.
//   test_plugins cpu   host-only checks (grid layout, parser, packing, loaders)
//   test_plugins gpu   every registered HIP plug-in against the oracle
#include <cstdint>
#include <cstring>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "stencils/algorithm.hpp"
#include "stencils/algorithm_repository.hpp"
#include "stencils/data_loader.hpp"
#include "stencils/experiment_params.hpp"
#include "stencils/gol_hip_algorithms.hpp"

extern "C" {
void oracle_load_random(uint8_t* cells, uint64_t x, uint64_t y, uint64_t seed);
void oracle_naive_run(uint8_t* cells, uint64_t x, uint64_t y, uint64_t generations);
int oracle_pack_layout(const uint8_t* cells, uint64_t x, uint64_t y, uint32_t layout, uint32_t word_bits, void* words);
}

static int failures = 0;
#define CHECK(cond)                                                                \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::cerr << __FILE__ << ":" << __LINE__ << ": CHECK failed: " #cond << std::endl; \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

template <typename E>
static bool throws(const std::function<void()>& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

using Grid = infrastructure::Grid<2, char>;

static std::vector<uint8_t> oracle_cells(const Grid& g) {
    std::vector<uint8_t> c(g.size());
    for (std::size_t i = 0; i < g.size(); ++i) c[i] = g.data()[i] > 0;
    return c;
}

static infrastructure::ExperimentParams params_for(std::size_t x, std::size_t y, std::size_t iters) {
    infrastructure::ExperimentParams p;
    p.grid_dimensions = {x, y};
    p.iterations = iters;
    return p;
}

static void repository_collection_tests() {
    // a 2-D and a 3-D repository for one cell type, the 3-D one empty
    infrastructure::AlgorithmReposCollection<infrastructure::AlgRepoParams<2, char>,
                                             infrastructure::AlgRepoParams<3, char>>
        repos;
    auto* r2 = repos.get_repository<2, char>();
    auto* r3 = repos.get_repository<3, char>();
    CHECK(r2 != nullptr && r3 != nullptr);
    CHECK((repos.get_repository<2, std::int32_t>() == nullptr));
    algorithms::register_hip_algorithms(*r2);
    CHECK(r2->has_algorithm("gol-hip-temporal-64") && r3->names().empty());
    int visited = 0, holding = 0;
    repos.for_each([&](auto& repo) {
        ++visited;
        holding += repo.has_algorithm("gol-hip-temporal-64") ? 1 : 0;
    });
    CHECK(visited == 2 && holding == 1);
}

static void cpu_tests() {
    repository_collection_tests();
    // x-fastest layout: grid[x][y] is element y * X + x
    Grid g(5, 3);
    g[4][2] = 1;
    g[1][0] = 1;
    CHECK(g.data()[2 * 5 + 4] == 1 && g.data()[1] == 1 && g.size() == 15);
    CHECK(g.size_in<0>() == 5 && g.size_in<1>() == 3);

    // loader equals the oracle's restatement of RandomOnesZerosDataLoader
    auto p = params_for(128, 17, 0);
    p.random_seed = 7;
    Grid r = infrastructure::RandomOnesZerosDataLoader<2, char>().load_data(p);
    std::vector<uint8_t> o(128 * 17);
    oracle_load_random(o.data(), 128, 17, 7);
    CHECK(std::memcmp(o.data(), r.data(), o.size()) == 0);

    // packing: bit k of word wx is cell 64 * wx + k, and back
    auto w = algorithms::RowWords<64>::pack<char>(r);
    CHECK(w.size() == 2 * 17);
    CHECK(((w[3] >> 5) & 1) == static_cast<uint64_t>(r[64 + 5][1]));
    CHECK(algorithms::RowWords<64>::unpack<char>(w, 128, 17).equals(r));
    CHECK(algorithms::RowWords<32>::unpack<char>(algorithms::RowWords<32>::pack<char>(r), 128, 17).equals(r));
    CHECK(throws<std::invalid_argument>([&] { algorithms::RowWords<64>::pack<char>(Grid(96, 2)); }));

    // parser: the reference's flag names and defaults
    const char* argv[] = {"stencils", "--algorithm=gol-hip-temporal-64", "--grid-dimensions-x", "256",
                          "--grid-dimensions-y=64", "--iterations=10", "--validate", "--temporal-steps=6",
                          "--base-grid-encoding=int"};
    auto q = infrastructure::ParamsParser::parse(9, const_cast<char**>(argv));
    CHECK(q.algorithm_name == "gol-hip-temporal-64" && q.grid_dimensions[0] == 256 && q.grid_dimensions[1] == 64);
    CHECK(q.iterations == 10 && q.validate && q.temporal_steps == 6 && q.base_grid_encoding == "int");
    CHECK(q.random_seed == 42 && q.measurement_rounds == 1 && q.warmup_rounds == 0);
    const char* bad[] = {"stencils", "--no-such-flag=1"};
    CHECK(throws<std::invalid_argument>([&] { infrastructure::ParamsParser::parse(2, const_cast<char**>(bad)); }));
    const char* bad2[] = {"stencils", "--algorithm=a", "--grid-dimensions-x=x", "--grid-dimensions-y=1",
                          "--iterations=1"};
    CHECK(throws<std::invalid_argument>([&] { infrastructure::ParamsParser::parse(5, const_cast<char**>(bad2)); }));

    // pattern expressions: the gun emits a glider every 30 generations
    auto lp = params_for(64, 64, 0);
    lp.pattern_expression = "gosper-glider-gun[1, 1]; blinker[60, 60]";
    Grid gun = infrastructure::LexiconLoader<2, char>().load_data(lp);
    std::vector<uint8_t> c = oracle_cells(gun);
    std::size_t pop0 = 0;
    for (auto v : c) pop0 += v;
    CHECK(pop0 == 36 + 3);
    oracle_naive_run(c.data(), 64, 64, 30);
    std::size_t pop30 = 0;
    for (auto v : c) pop30 += v;
    CHECK(pop30 == 36 + 5 + 3);

    // the plug-ins' packers against the oracle's, for every layout and word width
    {
        Grid g = infrastructure::RandomOnesZerosDataLoader<2, char>().load_data(params_for(128, 64, 0));
        const std::vector<uint8_t> cells = oracle_cells(g);
        auto check_layout = [&](auto words, uint32_t layout, uint32_t bits) {
            using W = typename decltype(words)::value_type;
            std::vector<W> want(words.size());
            CHECK(oracle_pack_layout(cells.data(), 128, 64, layout, bits, want.data()) == 0);
            CHECK(words == want);
        };
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_ROWS, 32>::pack(g), GOL_LAYOUT_ROWS, 32);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_ROWS, 64>::pack(g), GOL_LAYOUT_ROWS, 64);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_COLUMNS, 16>::pack(g), GOL_LAYOUT_COLUMNS, 16);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_COLUMNS, 32>::pack(g), GOL_LAYOUT_COLUMNS, 32);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_COLUMNS, 64>::pack(g), GOL_LAYOUT_COLUMNS, 64);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_TILES, 16>::pack(g), GOL_LAYOUT_TILES, 16);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_TILES, 32>::pack(g), GOL_LAYOUT_TILES, 32);
        check_layout(algorithms::LayoutWords<GOL_LAYOUT_TILES, 64>::pack(g), GOL_LAYOUT_TILES, 64);
        using T64 = algorithms::LayoutWords<GOL_LAYOUT_TILES, 64>;
        Grid back = T64::unpack<char>(T64::pack(g), 128, 64);
        CHECK(oracle_cells(back) == cells);
        Grid odd = infrastructure::ZerosLoader<2, char>().load_data(params_for(128, 20, 0));
        CHECK(throws<std::invalid_argument>([&] { algorithms::LayoutWords<GOL_LAYOUT_COLUMNS, 16>::pack(odd); }));
    }

    // always-changing: blinkers every 4 cells
    Grid ac = infrastructure::AlwaysChangingSpaceLoader<2, char>().load_data(params_for(8, 8, 0));
    CHECK(ac[0][0] == 1 && ac[2][0] == 1 && ac[3][0] == 0 && ac[4][4] == 1 && ac[0][1] == 0);
}

template <typename cell_t>
static void gpu_tests_for_encoding() {
    infrastructure::AlgorithmRepository<2, cell_t> repo;
    algorithms::register_hip_algorithms(repo);
    const std::vector<std::string> names = {"gol-hip-rows-32",  "gol-hip-rows-64",  "gol-hip-temporal-32",
                                            "gol-hip-temporal-64", "gol-hip-local-32", "gol-hip-local-64",
                                            "gol-hip-auto-64"};
    for (auto& name : names) {
        std::cout << "[case] plug-in " << name << " cell bytes " << sizeof(cell_t) << std::endl;
        auto* alg = repo.fetch_algorithm(name);
        CHECK(alg != nullptr);
        for (auto dims : {std::pair<std::size_t, std::size_t>{256, 96}, {128, 1}, {2048 + 128, 70}}) {
            auto p = params_for(dims.first, dims.second, 57);
            p.random_seed = dims.first + dims.second;
            auto data = infrastructure::RandomOnesZerosDataLoader<2, cell_t>().load_data(p);
            std::vector<uint8_t> ref(data.size());
            for (std::size_t i = 0; i < data.size(); ++i) ref[i] = data.data()[i] > 0;
            oracle_naive_run(ref.data(), dims.first, dims.second, p.iterations);
            // twice: the lifecycle is re-entrant (warm-up and measured rounds reuse the instance)
            for (int round = 0; round < 2; ++round) {
                infrastructure::TimedAlgorithm<2, cell_t> timed(alg);
                timed.set_params(p);
                timed.set_and_format_input_data(data);
                timed.initialize_data_structures();
                timed.run(p.iterations);
                timed.finalize_data_structures();
                auto out = timed.fetch_result();
                CHECK(timed.actually_performed_iterations() == p.iterations);
                bool same = out.size() == ref.size();
                for (std::size_t i = 0; same && i < ref.size(); ++i) same = (out.data()[i] != 0) == (ref[i] != 0);
                CHECK(same);
                auto rep = timed.get_time_report();
                CHECK(rep.run >= 0 && rep.actually_performed_iterations == p.iterations);
                CHECK(rep.pretty_print().find("Time report:") == 0);
            }
        }
        // a width that is not a multiple of the word size
        const std::size_t bits = name.find("-32") != std::string::npos ? 32 : 64;
        auto bad = params_for(bits + 32 == 64 ? 48 : 96, 8, 1);
        auto g = infrastructure::ZerosLoader<2, cell_t>().load_data(bad);
        CHECK(throws<std::invalid_argument>([&] { alg->set_and_format_input_data(g); }));
    }
    // the column and tile layouts: heights a multiple of 64 (a 64-bit column word)
    for (const std::string name : {"gol-hip-cols-16", "gol-hip-cols-32", "gol-hip-cols-64", "gol-hip-tiles-16",
                                   "gol-hip-tiles-32", "gol-hip-tiles-64", "gol-hip-tiles-plain-64",
                                   "gol-hip-tiles-temporal-64", "gol-hip-cols-plain-32", "gol-hip-cols-plain-64",
                                   "gol-hip-cols-temporal-32", "gol-hip-cols-temporal-64"}) {
        std::cout << "[case] plug-in " << name << " cell bytes " << sizeof(cell_t) << std::endl;
        auto* alg = repo.fetch_algorithm(name);
        CHECK(alg != nullptr);
        for (auto dims : {std::pair<std::size_t, std::size_t>{256, 128}, {128, 64}, {2048 + 128, 192}}) {
            auto p = params_for(dims.first, dims.second, 45);
            p.random_seed = dims.first + 3 * dims.second;
            auto data = infrastructure::RandomOnesZerosDataLoader<2, cell_t>().load_data(p);
            std::vector<uint8_t> ref(data.size());
            for (std::size_t i = 0; i < data.size(); ++i) ref[i] = data.data()[i] > 0;
            oracle_naive_run(ref.data(), dims.first, dims.second, p.iterations);
            infrastructure::TimedAlgorithm<2, cell_t> timed(alg);
            timed.set_params(p);
            timed.set_and_format_input_data(data);
            timed.initialize_data_structures();
            timed.run(p.iterations);
            timed.finalize_data_structures();
            auto out = timed.fetch_result();
            bool same = out.size() == ref.size();
            for (std::size_t i = 0; same && i < ref.size(); ++i) same = (out.data()[i] != 0) == (ref[i] != 0);
            CHECK(same);
        }
        // a height off the layout's word grid
        auto g = infrastructure::ZerosLoader<2, cell_t>().load_data(params_for(128, 6, 1));
        CHECK(throws<std::invalid_argument>([&] { alg->set_and_format_input_data(g); }));
    }
    // temporal_steps and a time budget
    auto p = params_for(1024, 512, 400);
    p.temporal_steps = 6;
    auto data = infrastructure::RandomOnesZerosDataLoader<2, cell_t>().load_data(p);
    auto* t = repo.fetch_algorithm("gol-hip-temporal-64");
    t->set_params(p);
    t->set_and_format_input_data(data);
    t->initialize_data_structures();
    t->run(p.iterations);
    t->finalize_data_structures();
    std::vector<uint8_t> ref(data.size());
    for (std::size_t i = 0; i < data.size(); ++i) ref[i] = data.data()[i] > 0;
    oracle_naive_run(ref.data(), 1024, 512, 400);
    auto out = t->fetch_result();
    bool same = true;
    for (std::size_t i = 0; same && i < ref.size(); ++i) same = (out.data()[i] != 0) == (ref[i] != 0);
    CHECK(same);
}

int main(int argc, char** argv) {
    const std::string what = argc > 1 ? argv[1] : "cpu";
    try {
        if (what == "cpu") cpu_tests();
        else if (what == "gpu") {
            gpu_tests_for_encoding<char>();
            gpu_tests_for_encoding<std::int32_t>();
        } else {
            std::cerr << "usage: test_plugins cpu|gpu" << std::endl;
            return 2;
        }
    } catch (const std::exception& e) {
        std::cerr << "exception: " << e.what() << std::endl;
        return 1;
    }
    std::cout << (failures ? "FAILED " : "OK ") << failures << " failures" << std::endl;
    return failures ? 1 : 0;
}
