// Pattern-expression inputs (reference: src/infrastructure/gol-lexicon/lexicon.hpp:92-186):
// a table of named Life Lexicon patterns and the "name[x, y]; ..." expression
// that places them. Standard C++ only, so the engine library and the driver
// share it.
#pragma once

#include <cctype>
#include <cstddef>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace infrastructure {

// Named patterns in run-length form ("b" dead, "o" alive, "$" end of row, a
// count before a symbol repeats it); rows are padded with dead cells to the
// pattern's width, and the whole rectangle is written when a pattern is
// placed. The shapes are those of the Life Lexicon entries the reference's
// workloads name (the spacefiller is the pattern of its published
// measurements, cluster-experiments/generate-final-measurements.py:25-33).
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
        add("spacefiller",
            "18bo$17b3o$12b3o4b2o$11bo2b3o2bob2o$10bo3bobo2bobo$10bo4bobobobob2o$12bo4bobo3b2o$4o5bobo4bo3bob3o$"
            "o3b2obob3ob2o9b2o$o5b2o5bo$bo2b2obo2bo2bob2o$7bobobobobobo5b4o$bo2b2obo2bo2bo2b2obob2o3bo$"
            "o5b2o3bobobo3b2o5bo$o3b2obob2o2bo2bo2bob2o2bo$4o5bobobobobobo$10b2obo2bo2bob2o2bo$13bo5b2o5bo$"
            "b2o9b2ob3obob2o3bo$2b3obo3bo4bobo5b4o$2b2o3bobo4bo$2b2obobobobo4bo$5bobo2bobo3bo$4b2obo2b3o2bo$"
            "6b2o4b3o$7b3o$8bo!");
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
        std::size_t width = 0;
        for (const auto& row : p) width = row.size() > width ? row.size() : width;
        for (auto& row : p) row.resize(width, 0);
        table_[name] = p;
    }
    std::map<std::string, Pattern> table_;
    Pattern empty_;
};

// "name[x, y]; name[x, y]; ..." — each pattern's top-left cell at (x, y); the
// pattern's whole rectangle is written, dead cells included (as the reference
// does), and cells beyond the x_size x y_size grid are skipped (the reference
// writes them unchecked). set(x, y, alive) receives every written cell.
template <typename SetCell>
void place_patterns(const std::string& expression, std::size_t x_size, std::size_t y_size, const PatternTable& table,
                    SetCell&& set) {
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
                if (x0 + x < x_size && y0 + y < y_size) set(x0 + x, y0 + y, p[y][x] != 0);
    }
}

}  // namespace infrastructure
