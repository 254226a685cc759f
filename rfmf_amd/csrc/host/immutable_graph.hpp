// immutable_graph.hpp -- frozen CSR snapshot every chain reads, API of the
// reference's ImmutableGraph / AnnealCsr (src/immutable_graph.hpp:23-157).
//
// Dense ids are ranks of the original ids in ascending order; each row is
// sorted by dense neighbour id, which fixes the summation order of every
// neighbour loop (initial fields, corr) independently of hash order.  Every
// undirected edge is stored twice (a self loop once).
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "graph.hpp"

namespace anneallib {

template <class IdT = std::int32_t, class WeightT = std::int32_t, class OffsetT = std::uint32_t>
class ImmutableGraph {
public:
    using id_type = IdT;
    using weight_type = WeightT;
    using offset_type = OffsetT;

    struct NeighborSpan {
        const IdT* to;
        const WeightT* weight;
        OffsetT n;
        OffsetT size() const { return n; }
    };

    ImmutableGraph() : offsets_(1, 0) {}

    explicit ImmutableGraph(const Graph& g) {
        newToOld_.assign(g.getVertices().begin(), g.getVertices().end());
        std::sort(newToOld_.begin(), newToOld_.end());
        const std::size_t n = newToOld_.size();
        auto dense = [&](std::int64_t old) -> IdT {
            const auto it = std::lower_bound(newToOld_.begin(), newToOld_.end(), old);
            if (it == newToOld_.end() || *it != old)
                throw std::runtime_error("ImmutableGraph: neighbour outside the vertex set");
            return static_cast<IdT>(it - newToOld_.begin());
        };
        offsets_.assign(n + 1, 0);
        for (std::size_t v = 0; v < n; ++v)
            offsets_[v + 1] = static_cast<OffsetT>(offsets_[v] + g.degree(newToOld_[v]));
        edgeTo_.resize(offsets_[n]);
        edgeWeight_.resize(offsets_[n]);
        std::vector<std::pair<IdT, WeightT>> row;
        for (std::size_t v = 0; v < n; ++v) {
            row.clear();
            for (const auto& nb : g.getEdges(newToOld_[v]))
                row.emplace_back(dense(nb.first), static_cast<WeightT>(nb.second));
            std::sort(row.begin(), row.end(),
                      [](const auto& a, const auto& b) { return a.first < b.first; });
            OffsetT pos = offsets_[v];
            for (const auto& e : row) {
                edgeTo_[pos] = e.first;
                edgeWeight_[pos] = e.second;
                ++pos;
            }
        }
    }

    // Adopt already-built dense CSR arrays (rows must be sorted).
    ImmutableGraph(std::vector<OffsetT> offsets, std::vector<IdT> to, std::vector<WeightT> w,
                   std::vector<std::int64_t> newToOld)
        : offsets_(std::move(offsets)), edgeTo_(std::move(to)), edgeWeight_(std::move(w)),
          newToOld_(std::move(newToOld)) {}

    std::uint64_t numVertices() const { return newToOld_.size(); }
    std::uint64_t numEdges() const { return edgeTo_.size() / 2; }

    NeighborSpan neighbors(std::int64_t v) const {
        if (v < 0 || static_cast<std::uint64_t>(v) >= newToOld_.size())
            throw std::out_of_range("ImmutableGraph::neighbors: invalid vertex");
        return NeighborSpan{edgeTo_.data() + offsets_[v], edgeWeight_.data() + offsets_[v],
                            static_cast<OffsetT>(offsets_[v + 1] - offsets_[v])};
    }

    // Raw arrays, for handing the snapshot to the device runtime.
    const std::vector<OffsetT>& offsets() const { return offsets_; }
    const std::vector<IdT>& edgeTo() const { return edgeTo_; }
    const std::vector<WeightT>& edgeWeight() const { return edgeWeight_; }
    const std::vector<std::int64_t>& newToOld() const { return newToOld_; }

private:
    std::vector<OffsetT> offsets_;
    std::vector<IdT> edgeTo_;
    std::vector<WeightT> edgeWeight_;
    std::vector<std::int64_t> newToOld_;
};

using AnnealCsr = ImmutableGraph<std::int32_t, std::int32_t, std::uint32_t>;

}  // namespace anneallib
