// Dense cell grid of the plug-in boundary (reference: src/infrastructure/grid.hpp:73-220).
// Layout x-fastest: grid[x][y] is element y * X + x; sizes given as (X, Y).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace infrastructure {

template <int DIMS, typename ElementType>
class Grid {
   public:
    using size_type = std::size_t;
    using element_t = ElementType;

    Grid() : dims_(DIMS, 0), strides_(DIMS + 1, 0) {}

    template <typename... Sizes, typename = std::enable_if_t<sizeof...(Sizes) == DIMS>>
    explicit Grid(Sizes... dims) : Grid(std::vector<size_type>{static_cast<size_type>(dims)...}) {}

    explicit Grid(const std::vector<size_type>& dims) : dims_(dims) {
        if (dims.size() != DIMS) throw std::invalid_argument("Dimension sizes must match the number of dimensions");
        size_type total = 1;
        strides_.clear();
        for (int i = 0; i < DIMS; ++i) {
            strides_.push_back(total);
            total *= dims[i];
        }
        strides_.push_back(total);
        elements_.assign(total, ElementType{});
    }

    // grid[x][y] (2-D): a column proxy for x, then the element at row y
    class Column {
       public:
        Column(ElementType* base, size_type stride) : base_(base), stride_(stride) {}
        ElementType& operator[](size_type y) const { return base_[y * stride_]; }

       private:
        ElementType* base_;
        size_type stride_;
    };
    class ConstColumn {
       public:
        ConstColumn(const ElementType* base, size_type stride) : base_(base), stride_(stride) {}
        const ElementType& operator[](size_type y) const { return base_[y * stride_]; }

       private:
        const ElementType* base_;
        size_type stride_;
    };
    Column operator[](size_type x) {
        static_assert(DIMS == 2, "indexing proxies are provided for 2-D grids");
        return Column(elements_.data() + x, strides_[1]);
    }
    ConstColumn operator[](size_type x) const {
        static_assert(DIMS == 2, "indexing proxies are provided for 2-D grids");
        return ConstColumn(elements_.data() + x, strides_[1]);
    }

    static constexpr int dimensions() { return DIMS; }
    ElementType* data() { return elements_.data(); }
    const ElementType* data() const { return elements_.data(); }
    std::vector<ElementType>* data_as_vector() { return &elements_; }
    const std::vector<ElementType>* data_as_vector() const { return &elements_; }
    size_type size() const { return elements_.size(); }
    template <int Dim>
    size_type size_in() const {
        return dims_[Dim];
    }
    size_type size_in(size_t dim) const { return dims_[dim]; }

    bool equals(const Grid& other) const { return dims_ == other.dims_ && elements_ == other.elements_; }

   private:
    std::vector<ElementType> elements_;
    std::vector<size_type> dims_;
    std::vector<size_type> strides_;
};

}  // namespace infrastructure
