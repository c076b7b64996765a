// Name -> plug-in registry (reference: src/infrastructure/algorithm_repository.hpp:11-51).
#pragma once

#include <memory>
#include <string>
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

}  // namespace infrastructure
