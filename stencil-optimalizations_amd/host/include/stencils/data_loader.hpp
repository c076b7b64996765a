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

// Named patterns in run-length form ("b" dead, "o" alive, "$" end of row, a
// count before a symbol repeats it). The shapes are those of the Life Lexicon
// entries the reference's workloads name.
class PatternTable {
   public:
    using Pattern = std::vector<std::vector<char>>;

    PatternTable() {
        add("block", "2o$2o!");
        add("blinker", "3o!");
        add("beehive", "b2o$o2bo$b2o!");
        add("glider", "3o$o$bo!");
        add("gosper-glider-gun",
            "24bo$22bobo$12b2o6b2o12b2o$11bo3bo4b2o12b2o$2o8bo5bo3b2o$2o8bo3bob2o4bobo$10bo5bo7bo$11bo3bo$12b2o!");
    }
    // Unknown names give an empty pattern, as the reference's lexicon does.
    const Pattern& get(const std::string& name) const {
        auto it = table_.find(name);
        return it == table_.end() ? empty_ : it->second;
    }
    bool has(const std::string& name) const { return table_.count(name) != 0; }

   private:
    void add(const std::string& name, const std::string& rle) {
        Pattern p(1);
        int count = 0;
        for (char ch : rle) {
            if (std::isdigit(static_cast<unsigned char>(ch))) {
                count = count * 10 + (ch - '0');
                continue;
            }
            const int n = count ? count : 1;
            count = 0;
            if (ch == 'b' || ch == 'o') {
                for (int i = 0; i < n; ++i) p.back().push_back(ch == 'o' ? 1 : 0);
            } else if (ch == '$') {
                for (int i = 0; i < n; ++i) p.emplace_back();
            } else if (ch == '!') {
                break;
            }
        }
        table_[name] = p;
    }
    std::map<std::string, Pattern> table_;
    Pattern empty_;
};

// "name[x, y]; name[x, y]; ..." — each pattern's top-left cell at (x, y);
// cells beyond the grid are skipped (the reference writes them unchecked).
template <typename ElementType>
void insert_patterns(Grid<2, ElementType>& grid, const std::string& expression, const PatternTable& table) {
    std::string e;
    for (char ch : expression)
        if (!std::isspace(static_cast<unsigned char>(ch))) e.push_back(ch);
    std::size_t i = 0;
    while (i < e.size()) {
        std::string name, xs, ys;
        while (i < e.size() && e[i] != '[') name.push_back(e[i++]);
        if (i >= e.size()) throw std::invalid_argument("pattern expression: missing '[' after " + name);
        ++i;
        while (i < e.size() && e[i] != ',') xs.push_back(e[i++]);
        ++i;
        while (i < e.size() && e[i] != ']') ys.push_back(e[i++]);
        if (i >= e.size()) throw std::invalid_argument("pattern expression: missing ']' after " + name);
        i += 1;
        if (i < e.size() && e[i] == ';') ++i;
        const std::size_t x0 = std::stoul(xs), y0 = std::stoul(ys);
        const auto& p = table.get(name);
        for (std::size_t y = 0; y < p.size(); ++y)
            for (std::size_t x = 0; x < p[y].size(); ++x)
                if (x0 + x < grid.template size_in<0>() && y0 + y < grid.template size_in<1>())
                    grid[x0 + x][y0 + y] = static_cast<ElementType>(p[y][x]);
    }
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
