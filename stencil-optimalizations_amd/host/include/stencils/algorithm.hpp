// The algorithm plug-in interface (reference: src/infrastructure/algorithm.hpp:17-280):
// Algorithm<Dims, ElementType> with its lifecycle, TimeReport and TimedAlgorithm.
// Same virtual methods, same order of use, same time-report text (the
// reference's log parser, cluster-experiments/results_abstraction.py:4-79,
// reads "Time report:" blocks with "<phase>: <ms> ms" lines).
#pragma once

#include <cstddef>
#include <iomanip>
#include <sstream>
#include <string>

#include "experiment_params.hpp"
#include "grid.hpp"
#include "timer.hpp"

namespace infrastructure {

template <int Dims, typename ElementType>
class Algorithm {
   public:
    using size_type = std::size_t;
    using DataGrid = Grid<Dims, ElementType>;

    Algorithm() = default;
    virtual ~Algorithm() = default;
    Algorithm(const Algorithm&) = delete;
    Algorithm& operator=(const Algorithm&) = delete;
    Algorithm(Algorithm&&) = default;
    Algorithm& operator=(Algorithm&&) = default;

    // Takes the input grid and formats it (packing) on the host.
    virtual void set_and_format_input_data(const DataGrid& data) = 0;
    // Allocates device state and copies the input to the device.
    virtual void initialize_data_structures() = 0;
    // Advances the grid; may stop early at params.max_runtime_seconds.
    virtual void run(size_type iterations) = 0;
    // Copies the state back to the host and releases device state.
    virtual void finalize_data_structures() = 0;
    // Unpacks the result into a grid.
    virtual DataGrid fetch_result() = 0;
    // Generations actually performed by the last run.
    virtual std::size_t actually_performed_iterations() const = 0;

    virtual void set_params(const ExperimentParams& params) { this->params = params; }

    // Kept for interface parity; no AN5D algorithm exists here.
    virtual bool is_an5d_cuda_alg() const { return false; }

   protected:
    ExperimentParams params;
};

struct TimeReport {
    static constexpr double INVALID = -1;

    double set_and_format_input_data = INVALID;
    double initialize_data_structures = INVALID;
    double run = INVALID;
    double finalize_data_structures = INVALID;
    double fetch_result = INVALID;
    std::size_t actually_performed_iterations = 0;

    double runtime_per_iteration() const {
        if (run == INVALID || actually_performed_iterations == 0) return INVALID;
        return run / static_cast<double>(actually_performed_iterations);
    }

    std::string pretty_print() const {
        std::string out = "Time report:\n";
        out += line("  set_and_format_input_data:  ", set_and_format_input_data);
        out += line("  initialize_data_structures: ", initialize_data_structures);
        out += line("  run:                        ", run);
        out += "   performed iters:  " + std::to_string(actually_performed_iterations) + "\n";
        out += "   runtime per iter: " + std::to_string(runtime_per_iteration()) + " ms\n";
        out += line("  finalize_data_structures:   ", finalize_data_structures);
        out += line("  fetch_result:               ", fetch_result);
        return out;
    }

    std::string pretty_print_speedup(const TimeReport& bench) const {
        std::string out = "Time report:\n";
        out += speedup_line("  set_and_format_input_data:  ", set_and_format_input_data, bench.set_and_format_input_data);
        out += speedup_line("  initialize_data_structures: ", initialize_data_structures,
                            bench.initialize_data_structures);
        double bench_run = bench.run;
        if (bench.actually_performed_iterations && bench.actually_performed_iterations != actually_performed_iterations)
            bench_run = bench.run * actually_performed_iterations / bench.actually_performed_iterations;
        out += speedup_line("  run:                        ", run, bench_run);
        out += "   performed iters:  " + std::to_string(actually_performed_iterations) + "\n";
        out += "   runtime per iter: " + std::to_string(runtime_per_iteration()) + " ms ~ " +
               speedup(bench.runtime_per_iteration(), runtime_per_iteration()) + "\n";
        out += speedup_line("  finalize_data_structures:   ", finalize_data_structures, bench.finalize_data_structures);
        out += speedup_line("  fetch_result:               ", fetch_result, bench.fetch_result);
        return out;
    }

   private:
    static std::string line(const std::string& label, double ms) {
        if (ms == INVALID) return "";
        return label + std::to_string(ms) + " ms\n";
    }
    static std::string speedup(double bench, double time) {
        std::stringstream ss;
        ss << std::fixed << std::setprecision(2) << bench / time << " x";
        return ss.str();
    }
    static std::string speedup_line(const std::string& label, double ms, double bench) {
        if (ms == INVALID || bench == INVALID) return "";
        return label + std::to_string(ms) + " ms ~ " + speedup(bench, ms) + "\n";
    }
};

// Wraps an algorithm and records the wall time of each lifecycle phase.
template <int Dims, typename ElementType>
class TimedAlgorithm : public Algorithm<Dims, ElementType> {
   public:
    using size_type = std::size_t;
    using DataGrid = Grid<Dims, ElementType>;

    explicit TimedAlgorithm(Algorithm<Dims, ElementType>* algorithm) : algorithm_(algorithm) {}

    void set_and_format_input_data(const DataGrid& data) override {
        report_.set_and_format_input_data = Timer::measure_ms([&]() { algorithm_->set_and_format_input_data(data); });
    }
    void initialize_data_structures() override {
        report_.initialize_data_structures = Timer::measure_ms([&]() { algorithm_->initialize_data_structures(); });
    }
    void run(size_type iterations) override {
        report_.run = Timer::measure_ms([&]() { algorithm_->run(iterations); });
        report_.actually_performed_iterations = algorithm_->actually_performed_iterations();
    }
    void finalize_data_structures() override {
        report_.finalize_data_structures = Timer::measure_ms([&]() { algorithm_->finalize_data_structures(); });
    }
    DataGrid fetch_result() override {
        DataGrid result;
        report_.fetch_result = Timer::measure_ms([&]() { result = algorithm_->fetch_result(); });
        return result;
    }
    std::size_t actually_performed_iterations() const override {
        return algorithm_->actually_performed_iterations();
    }
    void set_params(const ExperimentParams& params) override { algorithm_->set_params(params); }

    TimeReport get_time_report() const { return report_; }

   private:
    TimeReport report_;
    Algorithm<Dims, ElementType>* algorithm_;
};

}  // namespace infrastructure
