// Name -> plug-in registries (reference: src/infrastructure/algorithm_repository.hpp:11-100).
#pragma once

#include <cstddef>
#include <memory>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "algorithm.hpp"

namespace infrastructure {

template <int Dims, typename ElementType>
class AlgorithmRepository {
   public:
    using AlgType = Algorithm<Dims, ElementType>;

    void register_algorithm(const std::string& name, std::unique_ptr<AlgType> algorithm) {
        algorithms_.insert_or_assign(name, std::move(algorithm));
    }
    template <typename Alg>
    void register_algorithm(const std::string& name) {
        algorithms_.insert_or_assign(name, std::make_unique<Alg>());
    }
    AlgType* fetch_algorithm(const std::string& name) {
        auto it = algorithms_.find(name);
        return it == algorithms_.end() ? nullptr : it->second.get();
    }
    const AlgType* fetch_algorithm(const std::string& name) const {
        auto it = algorithms_.find(name);
        return it == algorithms_.end() ? nullptr : it->second.get();
    }
    bool has_algorithm(const std::string& name) const { return algorithms_.count(name) != 0; }
    std::vector<std::string> names() const {
        std::vector<std::string> out;
        for (auto& kv : algorithms_) out.push_back(kv.first);
        return out;
    }

   private:
    std::unordered_map<std::string, std::unique_ptr<AlgType>> algorithms_;
};

// Names one repository of a collection: grids of Dims dimensions with cells of
// ElementType (reference: src/infrastructure/algorithm_repository.hpp:53-59).
template <int Dims, typename ElementType>
struct AlgRepoParams {
    static constexpr int DIMS = Dims;
    using ElementT = ElementType;
};

// One repository per (dimensions, cell type) pair, so that an experiment runs
// in the repository that holds the requested algorithm (reference:
// algorithm_repository.hpp:61-100; the experiment manager holds a 2-D and a
// 3-D repository for one cell type, experiment_manager.hpp:224-231, 442, and
// registers algorithms in the 2-D one only).
template <typename... Descriptors>
class AlgorithmReposCollection {
   public:
    // The repository for Dims and ElementType, or null when the collection has none.
    template <int Dims, typename ElementType>
    AlgorithmRepository<Dims, ElementType>* get_repository() {
        return find<AlgorithmRepository<Dims, ElementType>, 0>();
    }
    // Calls func(repository) for every repository, in declaration order.
    template <typename Func>
    void for_each(Func&& func) {
        std::apply([&](auto&... repo) { (func(repo), ...); }, repos_);
    }

   private:
    using Repos = std::tuple<AlgorithmRepository<Descriptors::DIMS, typename Descriptors::ElementT>...>;
    Repos repos_;

    template <typename Repo, std::size_t I>
    Repo* find() {
        if constexpr (I == std::tuple_size_v<Repos>) {
            return nullptr;
        } else if constexpr (std::is_same_v<std::tuple_element_t<I, Repos>, Repo>) {
            return &std::get<I>(repos_);
        } else {
            return find<Repo, I + 1>();
        }
    }
};

}  // namespace infrastructure
