// MI355X algorithm plug-ins for the reference's Algorithm<2, cell_t> interface
// (src/infrastructure/algorithm.hpp:17-87), driving the engine through the C ABI
// of include/golhip.h. Compiles with any C++17 compiler: no HIP headers.
//
// Registered names (register_hip_algorithms):
//   gol-hip-rows-32 / -64       one generation per launch
//                               (the role of cuda-naive-bitwise, cuda_naive_bitwise_kernel.cu:32-79)
//   gol-hip-temporal-32 / -64   temporal tiles, several generations per launch
//                               (the role of cuda-temporal-rowed, temporal_rowed.cu:33-229)
//   gol-hip-local-32 / -64      change tracking over active tiles
//                               (the role of cuda-naive-local-one-cell, cuda_local_one_cell.cu:53-242)
//   gol-hip-auto-64             shape and depth chosen by the library's time model
//   gol-hip-cols-16/32/64       the same engine fed in the reference's column
//   gol-hip-tiles-16/32/64      and tile bit layouts (BitColumns, BitTile), converted to rows
//   gol-hip-tiles-plain-64      kernels native to the 8x8-tile layout: one generation per launch
//   gol-hip-tiles-temporal-64   (cuda_naive_bitwise_kernel.cu:32-79 on BitTile, tiled-full-adder.cuh)
//                               and T per launch (temporal_simple_tiled.cu:41-132)
//   gol-hip-cols-plain-32/64    kernels native to the column layout, likewise
//   gol-hip-cols-temporal-32/64
//
// Parameters read from ExperimentParams: temporal_steps (> 1: generations per
// launch; 0 or 1: chosen by the library), thread_block_size (> 0: wavefronts
// per workgroup = thread_block_size / 64), max_runtime_seconds.
// Errors follow the reference: std::invalid_argument for a width that is not a
// multiple of the word size (general_bit_grid.hpp:208-215), std::runtime_error
// for unusable configurations, HipError (the CudaError of cuch.hpp:42-52) for
// device failures.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../../include/golhip.h"
#include "algorithm.hpp"
#include "algorithm_repository.hpp"
#include "grid.hpp"
#include "timer.hpp"

namespace algorithms {

class HipError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

inline void check_gol(int rc) {
    if (rc == GOL_OK) return;
    const std::string msg = gol_last_error();
    switch (rc) {
        case GOL_ERR_INVALID_ARGUMENT:
            throw std::invalid_argument(msg);
        case GOL_ERR_RUNTIME:
        case GOL_ERR_NOT_FOUND:
        case GOL_ERR_BUSY:
            throw std::runtime_error(msg);
        default:
            throw HipError(msg);
    }
}

// Packing under the reference's bit policies (GeneralBitGrid fill_grid /
// to_grid, general_bit_grid.hpp:171-246, with bit_modes.hpp):
//   GOL_LAYOUT_ROWS    FullyPackedRows (:127-163): word (wx, y) at y * (X / Bits) + wx,
//                      bit k = cell (Bits * wx + k, y);
//   GOL_LAYOUT_COLUMNS BitColumns (:10-32): word (x, wy) at wy * X + x, bit k = cell (x, Bits * wy + k);
//   GOL_LAYOUT_TILES   BitTile (:37-125): XB x YB cells (4x4, 8x4, 8x8 for 16, 32, 64 bits),
//                      word (wx, wy) at wy * (X / XB) + wx, cell (tx, ty) in bit Bits-1-(ty*XB+tx).
// Cells > 0 are alive.
template <uint32_t Layout, std::size_t Bits>
struct LayoutWords {
    using word_type = std::conditional_t<Bits == 64, std::uint64_t,
                                         std::conditional_t<Bits == 32, std::uint32_t, std::uint16_t>>;
    static constexpr std::size_t XB = Layout == GOL_LAYOUT_ROWS ? Bits
                                      : Layout == GOL_LAYOUT_COLUMNS ? 1
                                                                     : (Bits == 16 ? 4 : 8);
    static constexpr std::size_t YB = Layout == GOL_LAYOUT_ROWS ? 1
                                      : Layout == GOL_LAYOUT_COLUMNS ? Bits
                                                                     : (Bits == 64 ? 8 : 4);
    static constexpr bool kMsbFirst = Layout == GOL_LAYOUT_TILES;

    template <typename cell_t>
    static std::vector<word_type> pack(const infrastructure::Grid<2, cell_t>& grid) {
        const std::size_t X = grid.template size_in<0>(), Y = grid.template size_in<1>();
        if (Y % YB != 0) throw std::invalid_argument("Grid dimensions Y must be a multiple of " + std::to_string(YB));
        if (X % XB != 0) throw std::invalid_argument("Grid dimensions X must be a multiple of " + std::to_string(XB));
        std::vector<word_type> words(X / XB * (Y / YB), 0);
        const cell_t* d = grid.data();
        for (std::size_t wy = 0; wy < Y / YB; ++wy)
            for (std::size_t wx = 0; wx < X / XB; ++wx) {
                word_type w = 0;
                std::size_t bit = kMsbFirst ? Bits - 1 : 0;
                for (std::size_t yb = 0; yb < YB; ++yb)
                    for (std::size_t xb = 0; xb < XB; ++xb, bit = kMsbFirst ? bit - 1 : bit + 1)
                        if (d[(wy * YB + yb) * X + wx * XB + xb] > 0) w |= static_cast<word_type>(word_type{1} << bit);
                words[wy * (X / XB) + wx] = w;
            }
        return words;
    }

    template <typename cell_t>
    static infrastructure::Grid<2, cell_t> unpack(const std::vector<word_type>& words, std::size_t X, std::size_t Y) {
        infrastructure::Grid<2, cell_t> grid(X, Y);
        cell_t* d = grid.data();
        for (std::size_t wy = 0; wy < Y / YB; ++wy)
            for (std::size_t wx = 0; wx < X / XB; ++wx) {
                const word_type w = words[wy * (X / XB) + wx];
                std::size_t bit = kMsbFirst ? Bits - 1 : 0;
                for (std::size_t yb = 0; yb < YB; ++yb)
                    for (std::size_t xb = 0; xb < XB; ++xb, bit = kMsbFirst ? bit - 1 : bit + 1)
                        d[(wy * YB + yb) * X + wx * XB + xb] = static_cast<cell_t>((w >> bit) & 1);
            }
        return grid;
    }
};

template <std::size_t Bits>
using RowWords = LayoutWords<GOL_LAYOUT_ROWS, Bits>;

template <typename grid_cell_t, std::size_t Bits, uint32_t Kernel, uint32_t Layout = GOL_LAYOUT_ROWS,
          uint32_t Flags = 0>
class GoLHip : public infrastructure::Algorithm<2, grid_cell_t> {
   public:
    using size_type = std::size_t;
    using DataGrid = infrastructure::Grid<2, grid_cell_t>;
    using Words = LayoutWords<Layout, Bits>;

    GoLHip() = default;
    ~GoLHip() override { release(); }

    void set_and_format_input_data(const DataGrid& data) override {
        x_ = data.template size_in<0>();
        y_ = data.template size_in<1>();
        words_ = Words::template pack<grid_cell_t>(data);
    }

    void initialize_data_structures() override {
        release();
        gol_config c{};
        c.x_cells = x_;
        c.y_cells = y_;
        c.word_bits = static_cast<uint32_t>(Bits);
        c.kernel = Kernel;
        c.layout = Layout;
        c.temporal_steps = this->params.temporal_steps > 1 ? static_cast<uint32_t>(this->params.temporal_steps) : 0;
        c.waves = this->params.thread_block_size >= 64 ? static_cast<uint32_t>(this->params.thread_block_size / 64) : 0;
        c.device = -1;
        c.flags = Flags;
        check_gol(gol_create(&c, &handle_));
        check_gol(gol_upload(handle_, words_.data(), words_.size()));
    }

    void run(size_type iterations) override {
        performed_ = 0;
        if (Kernel == GOL_KERNEL_ACTIVE && this->params.collect_touched_tiles_stats) {
            run_with_tile_stats(iterations);
            return;
        }
        if (this->params.max_runtime_seconds >= static_cast<std::size_t>(infrastructure::StopWatch::MAX_TIME)) {
            check_gol(gol_step(handle_, iterations));
            check_gol(gol_sync(handle_));
            performed_ = iterations;
            return;
        }
        // time budget: advance in chunks and stop once it is spent
        infrastructure::StopWatch watch(this->params.max_runtime_seconds);
        gol_info info{};
        check_gol(gol_get_info(handle_, &info));
        const size_type chunk = std::max<size_type>(1, static_cast<size_type>(info.temporal_steps) * 16);
        while (performed_ < iterations && !watch.time_is_up()) {
            const size_type n = std::min(chunk, iterations - performed_);
            check_gol(gol_step(handle_, n));
            check_gol(gol_sync(handle_));
            performed_ += n;
        }
    }

    void finalize_data_structures() override {
        check_gol(gol_download(handle_, words_.data(), words_.size()));
        release();
    }

    DataGrid fetch_result() override { return Words::template unpack<grid_cell_t>(words_, x_, y_); }

    std::size_t actually_performed_iterations() const override { return performed_; }

   private:
    // --collect-touched-tiles-stats (the reference's cuda_local_one_cell.hpp:60-70,
    // 130-155, 200-270): one block of temporal_steps generations at a time, then
    // the tiles it advanced, printed per generation as the reference's three
    // lines. A tile here is 32 * lane_dwords cells x (64 - 2 * temporal_steps)
    // rows, advanced for a whole block; a skipped tile needs no copy (its state
    // from two blocks ago, still in the output buffer, equals its state now), so
    // Copied_tiles is 0 and Untouched_tiles counts every tile not advanced.
    void run_with_tile_stats(size_type iterations) {
        infrastructure::StopWatch watch(this->params.max_runtime_seconds);
        gol_info info{};
        check_gol(gol_get_info(handle_, &info));
        const size_type block = std::max<size_type>(1, info.temporal_steps);
        const size_type tiles = info.tiles;
        std::vector<size_type> computed;
        uint64_t before = info.active_tiles_total;
        while (performed_ < iterations && !watch.time_is_up()) {
            const size_type n = std::min(block, iterations - performed_);
            check_gol(gol_step(handle_, n));
            check_gol(gol_get_info(handle_, &info));  // waits for the block
            // a block shorter than temporal_steps advances every tile untracked
            const size_type advanced = n < block ? tiles : static_cast<size_type>(info.active_tiles_total - before);
            before = info.active_tiles_total;
            computed.insert(computed.end(), n, advanced);
            performed_ += n;
        }
        std::cout << "Changed_tiles: ";
        for (size_type c : computed) std::cout << c << ";";
        std::cout << std::endl << "Copied_tiles: ";
        for (size_type i = 0; i < computed.size(); ++i) std::cout << 0 << ";";
        std::cout << std::endl << "Untouched_tiles: ";
        for (size_type c : computed) std::cout << tiles - c << ";";
        std::cout << std::endl;
    }

    void release() {
        if (handle_) {
            gol_destroy(handle_);
            handle_ = nullptr;
        }
    }

    gol_handle* handle_ = nullptr;
    std::vector<typename Words::word_type> words_;
    size_type x_ = 0, y_ = 0;
    size_type performed_ = 0;
};

template <typename cell_t>
void register_hip_algorithms(infrastructure::AlgorithmRepository<2, cell_t>& repo) {
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_PLAIN>>("gol-hip-rows-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_PLAIN>>("gol-hip-rows-64");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_TEMPORAL>>("gol-hip-temporal-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_TEMPORAL>>("gol-hip-temporal-64");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_ACTIVE>>("gol-hip-local-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_ACTIVE>>("gol-hip-local-64");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_AUTO>>("gol-hip-auto-64");
    // the reference's column and 8x8-tile layouts at the boundary (its cols / tiles
    // algorithm families, experiment_manager.hpp:88-94, 134-140); the engine converts
    // them to rows on the device
    repo.template register_algorithm<GoLHip<cell_t, 16, GOL_KERNEL_AUTO, GOL_LAYOUT_COLUMNS>>("gol-hip-cols-16");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_AUTO, GOL_LAYOUT_COLUMNS>>("gol-hip-cols-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_AUTO, GOL_LAYOUT_COLUMNS>>("gol-hip-cols-64");
    repo.template register_algorithm<GoLHip<cell_t, 16, GOL_KERNEL_AUTO, GOL_LAYOUT_TILES>>("gol-hip-tiles-16");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_AUTO, GOL_LAYOUT_TILES>>("gol-hip-tiles-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_AUTO, GOL_LAYOUT_TILES>>("gol-hip-tiles-64");
    // kernels native to those layouts (no conversion): the 8x8-tile adder one
    // generation per launch (the reference's gol-cuda-naive-bitwise-tiles-64,
    // "Core Tiled 64-bit") and T per launch (cuda-temporal-simple-tiled), and
    // the column kernel likewise (gol-cuda-naive-bitwise-cols-32/64)
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_PLAIN, GOL_LAYOUT_TILES, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-tiles-plain-64");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_TEMPORAL, GOL_LAYOUT_TILES, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-tiles-temporal-64");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_PLAIN, GOL_LAYOUT_COLUMNS, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-cols-plain-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_PLAIN, GOL_LAYOUT_COLUMNS, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-cols-plain-64");
    repo.template register_algorithm<GoLHip<cell_t, 32, GOL_KERNEL_TEMPORAL, GOL_LAYOUT_COLUMNS, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-cols-temporal-32");
    repo.template register_algorithm<GoLHip<cell_t, 64, GOL_KERNEL_TEMPORAL, GOL_LAYOUT_COLUMNS, GOL_FLAG_NATIVE_LAYOUT>>(
        "gol-hip-cols-temporal-64");
}

}  // namespace algorithms
