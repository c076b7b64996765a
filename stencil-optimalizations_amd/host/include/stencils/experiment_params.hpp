// Experiment parameters and their command line (reference:
// src/infrastructure/experiment_params.hpp:21-316). Same fields, flag names and
// defaults; parsed by hand (the reference uses cxxopts). Flags take
// "--name=value" or "--name value"; a boolean flag alone means true.
#pragma once

#include <cstddef>
#include <cstdlib>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "timer.hpp"

namespace infrastructure {

enum class StreamingDirection {
    in_X = 0,
    in_Y = 1,
    NAIVE = 2,
};

class ExperimentParams {
   public:
    std::string algorithm_name;
    std::vector<std::size_t> grid_dimensions;
    std::size_t iterations = 0;
    std::size_t max_runtime_seconds = static_cast<std::size_t>(StopWatch::MAX_TIME);

    std::string base_grid_encoding = "char";  // char, int

    std::size_t warmup_rounds = 0;
    std::size_t measurement_rounds = 1;

    std::string data_loader_name = "random-ones-zeros";
    std::string pattern_expression;

    bool measure_speedup = false;
    std::string speedup_bench_algorithm_name = "gol-hip-rows-64";

    bool validate = false;
    bool print_validation_diff = false;
    bool use_contracted_diff = false;
    std::string validation_algorithm_name = "gol-hip-rows-64";

    bool animate_output = false;
    bool colorful = true;
    bool collect_touched_tiles_stats = false;

    std::size_t random_seed = 42;
    std::string tag;

    std::size_t state_bits_count = 64;
    std::size_t thread_block_size = 0;
    std::size_t warp_dims_x = 0;
    std::size_t warp_dims_y = 0;
    std::size_t warp_tile_dims_x = 0;
    std::size_t warp_tile_dims_y = 0;
    std::size_t temporal_steps = 1;

    StreamingDirection streaming_direction = StreamingDirection::NAIVE;

    std::string pretty_print() const {
        std::stringstream ss;
        ss << "Experiment Parameters:\n";
        ss << "  algorithm_name: " << algorithm_name << "\n";
        ss << "  grid_dimensions: " << (grid_dimensions.size() > 0 ? grid_dimensions[0] : 0) << "x"
           << (grid_dimensions.size() > 1 ? grid_dimensions[1] : 0) << "\n";
        ss << "  iterations: " << iterations << "\n";
        ss << "  base_grid_encoding: " << base_grid_encoding << "\n";
        ss << "  max_runtime_seconds: " << max_runtime_seconds << "\n";
        ss << "  warmup_rounds: " << warmup_rounds << "\n";
        ss << "  measurement_rounds: " << measurement_rounds << "\n";
        ss << "  data_loader_name: " << data_loader_name << "\n";
        ss << "  pattern_expression: " << pattern_expression << "\n";
        ss << "  measure_speedup: " << measure_speedup << "\n";
        ss << "  speedup_bench_algorithm_name: " << speedup_bench_algorithm_name << "\n";
        ss << "  validate: " << validate << "\n";
        ss << "  print_validation_diff: " << print_validation_diff << "\n";
        ss << "  use_contracted_diff: " << use_contracted_diff << "\n";
        ss << "  validation_algorithm_name: " << validation_algorithm_name << "\n";
        ss << "  animate_output: " << animate_output << "\n";
        ss << "  colorful: " << colorful << "\n";
        ss << "  random_seed: " << random_seed << "\n";
        ss << "  tag: " << tag << "\n";
        ss << "  state_bits_count: " << state_bits_count << "\n";
        ss << "  thread_block_size: " << thread_block_size << "\n";
        ss << "  warp_dims_x: " << warp_dims_x << "\n";
        ss << "  warp_dims_y: " << warp_dims_y << "\n";
        ss << "  warp_tile_dims_x: " << warp_tile_dims_x << "\n";
        ss << "  warp_tile_dims_y: " << warp_tile_dims_y << "\n";
        ss << "  temporal_steps: " << temporal_steps << "\n";
        ss << "  collect_touched_tiles_stats: " << collect_touched_tiles_stats << "\n";
        ss << "  streaming_direction: "
           << (streaming_direction == StreamingDirection::in_X
                   ? "in-x"
                   : (streaming_direction == StreamingDirection::in_Y ? "in-y" : "naive"))
           << "\n";
        return ss.str();
    }
};

namespace opts {
const std::string ALGORITHM_NAME = "algorithm";
const std::string GRID_DIMENSIONS_X = "grid-dimensions-x";
const std::string GRID_DIMENSIONS_Y = "grid-dimensions-y";
const std::string ITERATIONS = "iterations";
const std::string MAX_RUNTIME_SECONDS = "max-runtime-seconds";
const std::string WARMUP_ROUNDS = "warmup-rounds";
const std::string MEASUREMENT_ROUNDS = "measurement-rounds";
const std::string DATA_LOADER_NAME = "data-loader";
const std::string PATTERN_EXPRESSION = "pattern-expression";
const std::string MEASURE_SPEEDUP = "measure-speedup";
const std::string SPEEDUP_BENCH_ALGORITHM_NAME = "speedup-bench-algorithm";
const std::string VALIDATE = "validate";
const std::string PRINT_VALIDATION_DIFF = "print-validation-diff";
const std::string USE_CONTRACTED_DIFF = "use-contracted-diff";
const std::string VALIDATION_ALGORITHM_NAME = "validation-algorithm";
const std::string ANIMATE_OUTPUT = "animate-output";
const std::string COLORFUL = "colorful";
const std::string RANDOM_SEED = "random-seed";
const std::string THREAD_BLOCK_SIZE = "thread-block-size";
const std::string WARP_DIMS_X = "warp-dims-x";
const std::string WARP_DIMS_Y = "warp-dims-y";
const std::string WARP_TILE_DIMS_X = "warp-tile-dims-x";
const std::string WARP_TILE_DIMS_Y = "warp-tile-dims-y";
const std::string STREAMING_DIRECTION = "streaming-direction";
const std::string STATE_BITS_COUNT = "state-bits-count";
const std::string BASE_GRID_ENCODING = "base-grid-encoding";
const std::string TAG = "tag";
const std::string COLLECT_TOUCHED_TILES_STATS = "collect-touched-tiles-stats";
const std::string TEMPORAL_STEPS = "temporal-steps";
}  // namespace opts

class ParamsParser {
   public:
    // Throws std::invalid_argument for unknown flags, missing values and bad numbers.
    static ExperimentParams parse(int argc, char** argv) {
        std::map<std::string, std::string> v;
        const std::vector<std::string> bools = {opts::MEASURE_SPEEDUP, opts::VALIDATE, opts::PRINT_VALIDATION_DIFF,
                                                opts::USE_CONTRACTED_DIFF, opts::ANIMATE_OUTPUT, opts::COLORFUL,
                                                opts::COLLECT_TOUCHED_TILES_STATS};
        const std::vector<std::string> known = {
            opts::ALGORITHM_NAME, opts::GRID_DIMENSIONS_X, opts::GRID_DIMENSIONS_Y, opts::ITERATIONS,
            opts::MAX_RUNTIME_SECONDS, opts::WARMUP_ROUNDS, opts::MEASUREMENT_ROUNDS, opts::DATA_LOADER_NAME,
            opts::PATTERN_EXPRESSION, opts::SPEEDUP_BENCH_ALGORITHM_NAME, opts::VALIDATION_ALGORITHM_NAME,
            opts::RANDOM_SEED, opts::THREAD_BLOCK_SIZE, opts::WARP_DIMS_X, opts::WARP_DIMS_Y, opts::WARP_TILE_DIMS_X,
            opts::WARP_TILE_DIMS_Y, opts::STREAMING_DIRECTION, opts::STATE_BITS_COUNT, opts::BASE_GRID_ENCODING,
            opts::TAG, opts::TEMPORAL_STEPS, "help"};
        auto is_in = [](const std::vector<std::string>& list, const std::string& k) {
            for (auto& s : list)
                if (s == k) return true;
            return false;
        };
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) throw std::invalid_argument("unexpected argument: " + a);
            a = a.substr(2);
            std::string key = a, val;
            const auto eq = a.find('=');
            bool has_val = eq != std::string::npos;
            if (has_val) {
                key = a.substr(0, eq);
                val = a.substr(eq + 1);
            }
            if (!is_in(known, key) && !is_in(bools, key)) throw std::invalid_argument("unknown option --" + key);
            if (!has_val) {
                if (is_in(bools, key) || key == "help") {
                    val = "true";
                } else {
                    if (i + 1 >= argc) throw std::invalid_argument("option --" + key + " needs a value");
                    val = argv[++i];
                }
            }
            v[key] = val;
        }
        if (v.count("help")) {
            std::cout << "stencils --algorithm=<name> --grid-dimensions-x=X --grid-dimensions-y=Y "
                         "--iterations=N [options of src/infrastructure/experiment_params.hpp]\n";
            std::exit(0);
        }
        ExperimentParams p;
        auto num = [&](const std::string& k, std::size_t def) -> std::size_t {
            auto it = v.find(k);
            if (it == v.end()) return def;
            try {
                std::size_t used = 0;
                const unsigned long long n = std::stoull(it->second, &used);
                if (used != it->second.size()) throw std::invalid_argument(it->second);
                return static_cast<std::size_t>(n);
            } catch (const std::exception&) {
                throw std::invalid_argument("option --" + k + " expects a non-negative integer, got " + it->second);
            }
        };
        auto str = [&](const std::string& k, const std::string& def) {
            auto it = v.find(k);
            return it == v.end() ? def : it->second;
        };
        auto flag = [&](const std::string& k, bool def) {
            auto it = v.find(k);
            if (it == v.end()) return def;
            if (it->second == "true" || it->second == "1") return true;
            if (it->second == "false" || it->second == "0") return false;
            throw std::invalid_argument("option --" + k + " expects true or false");
        };
        for (const auto& req : {opts::ALGORITHM_NAME, opts::GRID_DIMENSIONS_X, opts::GRID_DIMENSIONS_Y, opts::ITERATIONS})
            if (!v.count(req)) throw std::invalid_argument("option --" + req + " is required");
        p.algorithm_name = str(opts::ALGORITHM_NAME, "");
        p.grid_dimensions = {num(opts::GRID_DIMENSIONS_X, 0), num(opts::GRID_DIMENSIONS_Y, 0)};
        p.iterations = num(opts::ITERATIONS, 0);
        p.max_runtime_seconds = num(opts::MAX_RUNTIME_SECONDS, p.max_runtime_seconds);
        p.warmup_rounds = num(opts::WARMUP_ROUNDS, 0);
        p.measurement_rounds = num(opts::MEASUREMENT_ROUNDS, 1);
        p.data_loader_name = str(opts::DATA_LOADER_NAME, "random-ones-zeros");
        p.pattern_expression = str(opts::PATTERN_EXPRESSION, "");
        p.measure_speedup = flag(opts::MEASURE_SPEEDUP, false);
        p.speedup_bench_algorithm_name = str(opts::SPEEDUP_BENCH_ALGORITHM_NAME, p.speedup_bench_algorithm_name);
        p.validate = flag(opts::VALIDATE, false);
        p.print_validation_diff = flag(opts::PRINT_VALIDATION_DIFF, false);
        p.use_contracted_diff = flag(opts::USE_CONTRACTED_DIFF, false);
        p.validation_algorithm_name = str(opts::VALIDATION_ALGORITHM_NAME, p.validation_algorithm_name);
        p.animate_output = flag(opts::ANIMATE_OUTPUT, false);
        p.colorful = flag(opts::COLORFUL, true);
        p.random_seed = num(opts::RANDOM_SEED, 42);
        p.state_bits_count = num(opts::STATE_BITS_COUNT, 64);
        p.thread_block_size = num(opts::THREAD_BLOCK_SIZE, 0);
        p.warp_dims_x = num(opts::WARP_DIMS_X, 0);
        p.warp_dims_y = num(opts::WARP_DIMS_Y, 0);
        p.warp_tile_dims_x = num(opts::WARP_TILE_DIMS_X, 0);
        p.warp_tile_dims_y = num(opts::WARP_TILE_DIMS_Y, 0);
        const std::string dir = str(opts::STREAMING_DIRECTION, "naive");
        if (dir == "in-x")
            p.streaming_direction = StreamingDirection::in_X;
        else if (dir == "in-y")
            p.streaming_direction = StreamingDirection::in_Y;
        else if (dir == "naive")
            p.streaming_direction = StreamingDirection::NAIVE;
        else
            throw std::runtime_error("Invalid streaming direction");
        p.base_grid_encoding = str(opts::BASE_GRID_ENCODING, "char");
        p.tag = str(opts::TAG, "");
        p.collect_touched_tiles_stats = flag(opts::COLLECT_TOUCHED_TILES_STATS, false);
        p.temporal_steps = num(opts::TEMPORAL_STEPS, 1);
        return p;
    }
};

}  // namespace infrastructure
