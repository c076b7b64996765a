// Input loaders (reference: src/infrastructure/data_loader.hpp:14-152 and the
// pattern-expression loader src/infrastructure/gol-lexicon/lexicon.hpp:92-186).
#pragma once

#include <cctype>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "experiment_params.hpp"
#include "grid.hpp"
#include "lexicon.hpp"

namespace infrastructure {

template <int Dims, typename ElementType>
class Loader {
   public:
    virtual ~Loader() = default;
    virtual Grid<Dims, ElementType> load_data(const ExperimentParams& params) = 0;
};

// std::mt19937(seed), uniform_int_distribution<int>(0,1), one draw per cell in
// memory order (y * X + x) — data_loader.hpp:57-75.
template <int Dims, typename ElementType>
class RandomOnesZerosDataLoader : public Loader<Dims, ElementType> {
   public:
    Grid<Dims, ElementType> load_data(const ExperimentParams& params) override {
        Grid<Dims, ElementType> grid(params.grid_dimensions);
        std::mt19937 rng(static_cast<std::mt19937::result_type>(params.random_seed));
        std::uniform_int_distribution<int> dist(0, 1);
        ElementType* d = grid.data();
        for (std::size_t i = 0; i < grid.size(); ++i) d[i] = static_cast<ElementType>(dist(rng));
        return grid;
    }
};

template <int Dims, typename ElementType>
class ZerosLoader : public Loader<Dims, ElementType> {
   public:
    Grid<Dims, ElementType> load_data(const ExperimentParams& params) override {
        return Grid<Dims, ElementType>(params.grid_dimensions);
    }
};

// "name[x, y]; ..." expressions into a grid (lexicon.hpp).
template <typename ElementType>
void insert_patterns(Grid<2, ElementType>& grid, const std::string& expression, const PatternTable& table) {
    place_patterns(expression, grid.template size_in<0>(), grid.template size_in<1>(), table,
                   [&](std::size_t x, std::size_t y, bool alive) { grid[x][y] = static_cast<ElementType>(alive); });
}

template <int Dims, typename ElementType>
class LexiconLoader : public Loader<Dims, ElementType> {
   public:
    Grid<Dims, ElementType> load_data(const ExperimentParams& params) override {
        Grid<Dims, ElementType> grid(params.grid_dimensions);
        if constexpr (Dims == 2) insert_patterns(grid, params.pattern_expression, PatternTable());
        return grid;
    }
};

// Blinkers every 4 cells in x and y: every cell region changes every
// generation (data_loader.hpp:117-134).
template <int Dims, typename ElementType>
class AlwaysChangingSpaceLoader : public Loader<Dims, ElementType> {
   public:
    Grid<Dims, ElementType> load_data(const ExperimentParams& params) override {
        Grid<Dims, ElementType> grid(params.grid_dimensions);
        if constexpr (Dims == 2) {
            const std::size_t X = grid.template size_in<0>(), Y = grid.template size_in<1>();
            for (std::size_t x = 0; x < X; x += 4)
                for (std::size_t y = 0; y < Y; y += 4)
                    for (std::size_t k = 0; k < 3 && x + k < X; ++k) grid[x + k][y] = 1;
        }
        return grid;
    }
};

template <int Dims, typename ElementType>
std::unique_ptr<Loader<Dims, ElementType>> make_loader(const std::string& name) {
    if (name == "random-ones-zeros") return std::make_unique<RandomOnesZerosDataLoader<Dims, ElementType>>();
    if (name == "lexicon") return std::make_unique<LexiconLoader<Dims, ElementType>>();
    if (name == "always-changing") return std::make_unique<AlwaysChangingSpaceLoader<Dims, ElementType>>();
    if (name == "zeros") return std::make_unique<ZerosLoader<Dims, ElementType>>();
    throw std::invalid_argument("unknown data loader " + name);
}

}  // namespace infrastructure
