// stencils — experiment driver over the HIP plug-ins, with the reference's
// command line (src/main.cpp:22-45, src/infrastructure/experiment_manager.hpp:224-394):
// load data, warm up, measure rounds printing "Time report:" blocks, optionally
// measure a speedup against a bench algorithm and validate against another
// registered algorithm on the same input and iteration count.
//
// Validation compares with a second registered plug-in (default
// gol-hip-rows-64, the one-generation-per-launch kernel, whose results the
// test suite pins against the CPU oracle); the reference's default validator
// gol-cpu-naive is a CPU algorithm and is not part of this engine.

#include <cstdint>
#include <iostream>
#include <memory>
#include <string>

#include "stencils/algorithm.hpp"
#include "stencils/algorithm_repository.hpp"
#include "stencils/data_loader.hpp"
#include "stencils/experiment_params.hpp"
#include "stencils/gol_hip_algorithms.hpp"
#include "stencils/grid.hpp"

namespace {

using infrastructure::ExperimentParams;
using infrastructure::TimeReport;

template <typename cell_t>
struct Outcome {
    std::unique_ptr<infrastructure::Grid<2, cell_t>> result;
    TimeReport report;
};

template <typename cell_t>
Outcome<cell_t> perform(infrastructure::Algorithm<2, cell_t>& alg, const infrastructure::Grid<2, cell_t>& data,
                        const ExperimentParams& params, bool keep_result) {
    infrastructure::TimedAlgorithm<2, cell_t> timed(&alg);
    timed.set_params(params);
    timed.set_and_format_input_data(data);
    timed.initialize_data_structures();
    timed.run(params.iterations);
    timed.finalize_data_structures();
    Outcome<cell_t> out;
    if (keep_result) out.result = std::make_unique<infrastructure::Grid<2, cell_t>>(timed.fetch_result());
    out.report = timed.get_time_report();
    return out;
}

template <typename cell_t>
std::string first_differences(const infrastructure::Grid<2, cell_t>& a, const infrastructure::Grid<2, cell_t>& b) {
    std::string s;
    const std::size_t X = a.template size_in<0>();
    int shown = 0;
    for (std::size_t i = 0; i < a.size() && shown < 16; ++i) {
        if (a.data()[i] != b.data()[i]) {
            s += "  (" + std::to_string(i % X) + ", " + std::to_string(i / X) + "): expected " +
                 std::to_string(static_cast<int>(a.data()[i])) + ", got " + std::to_string(static_cast<int>(b.data()[i])) +
                 "\n";
            ++shown;
        }
    }
    return s;
}

template <typename cell_t>
int run_experiment(const ExperimentParams& params) {
    // a 2-D and a 3-D repository, as the reference's experiment manager holds;
    // its command line has no Z dimension, and the 3-D one stays empty
    infrastructure::AlgorithmReposCollection<infrastructure::AlgRepoParams<2, cell_t>,
                                             infrastructure::AlgRepoParams<3, cell_t>>
        repos;
    auto& repo = *repos.template get_repository<2, cell_t>();
    algorithms::register_hip_algorithms(repo);
    auto* alg = repo.fetch_algorithm(params.algorithm_name);
    if (!alg) {
        std::cerr << "unknown algorithm " << params.algorithm_name << "; registered:";
        for (auto& n : repo.names()) std::cerr << " " << n;
        std::cerr << std::endl;
        return 1;
    }
    auto loader = infrastructure::make_loader<2, cell_t>(params.data_loader_name);
    std::cout << "Loading data... " << params.data_loader_name << std::endl;
    infrastructure::Grid<2, cell_t> data;
    const double load_ms = infrastructure::Timer::measure_ms([&]() { data = loader->load_data(params); });
    std::cout << "  Data loaded in   " << load_ms << " ms" << std::endl << std::endl;

    TimeReport bench;
    if (params.measure_speedup) {
        auto* b = repo.fetch_algorithm(params.speedup_bench_algorithm_name);
        if (!b) throw std::invalid_argument("unknown bench algorithm " + params.speedup_bench_algorithm_name);
        std::cout << "Measuring bench algorithm... " << params.speedup_bench_algorithm_name << std::endl;
        for (std::size_t i = 0; i < params.warmup_rounds; ++i) perform(*b, data, params, false);
        bench = perform(*b, data, params, false).report;
    }
    for (std::size_t i = 0; i < params.warmup_rounds; ++i) {
        std::cout << "Warming up... " << i + 1 << " / " << params.warmup_rounds << std::endl;
        perform(*alg, data, params, false);
    }
    std::cout << "Running experiment...        " << params.algorithm_name << std::endl << std::endl;
    Outcome<cell_t> last;
    for (std::size_t i = 0; i < params.measurement_rounds; ++i) {
        std::cout << "Measurement round " << i + 1 << " / " << params.measurement_rounds << std::endl;
        last = perform(*alg, data, params, params.validate);
        std::cout << (params.measure_speedup ? last.report.pretty_print_speedup(bench) : last.report.pretty_print())
                  << std::endl;
    }
    if (params.validate && last.result) {
        std::cout << "Validating result... " << params.validation_algorithm_name << std::endl;
        auto* v = repo.fetch_algorithm(params.validation_algorithm_name);
        if (!v) throw std::invalid_argument("unknown validation algorithm " + params.validation_algorithm_name);
        ExperimentParams vp = params;
        vp.iterations = last.report.actually_performed_iterations;
        vp.max_runtime_seconds = static_cast<std::size_t>(infrastructure::StopWatch::MAX_TIME);
        auto expected = perform(*v, data, vp, true);
        if (expected.result->equals(*last.result)) {
            std::cout << "Validation successful" << std::endl;
        } else {
            std::cout << "Validation failed" << std::endl;
            if (params.print_validation_diff)
                std::cout << "Diff: \n" << first_differences(*expected.result, *last.result) << std::endl;
            return 2;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const ExperimentParams params = infrastructure::ParamsParser::parse(argc, argv);
        std::cout << params.pretty_print() << std::endl;
        if (params.base_grid_encoding == "char") return run_experiment<char>(params);
        if (params.base_grid_encoding == "int") return run_experiment<std::int32_t>(params);
        std::cerr << "Invalid base grid encoding: " << params.base_grid_encoding << std::endl;
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
