// graph.hpp -- mutable undirected integer-weighted graph used on the load
// path (file -> Graph -> AnnealCsr), API of the reference's anneallib::Graph
// (src/graph.hpp:13-40): adding an existing edge keeps the first weight.
#pragma once

#include <cstdint>
#include <unordered_map>
#include <unordered_set>

namespace anneallib {

class Graph {
public:
    Graph() = default;

    // No-op if the vertex exists.
    void addVertex(int64_t vertexId);
    // Adds missing endpoints; a repeated edge (either orientation) keeps the
    // weight it was first added with.
    void addEdge(int64_t u, int64_t v, int64_t weight);
    // Removes the vertex and its incident edges; no-op if absent.
    void removeVertex(int64_t vertexId);

    uint64_t degree(int64_t vertexId) const;   // throws if absent
    uint64_t numVertices() const { return vertices_.size(); }
    const std::unordered_set<int64_t>& getVertices() const { return vertices_; }
    // Neighbour -> weight map of a vertex; throws if absent.
    const std::unordered_map<int64_t, int64_t>& getEdges(int64_t vertexId) const;

private:
    std::unordered_set<int64_t> vertices_;
    std::unordered_map<int64_t, std::unordered_map<int64_t, int64_t>> adj_;
};

}  // namespace anneallib
