#include "graph.hpp"

#include <stdexcept>

namespace anneallib {

void Graph::addVertex(int64_t vertexId) {
    if (vertices_.insert(vertexId).second) adj_[vertexId];
}

void Graph::addEdge(int64_t u, int64_t v, int64_t weight) {
    addVertex(u);
    addVertex(v);
    auto& row = adj_[u];
    if (row.find(v) != row.end()) return;   // first weight wins
    row.emplace(v, weight);
    adj_[v].emplace(u, weight);
}

void Graph::removeVertex(int64_t vertexId) {
    const auto it = adj_.find(vertexId);
    if (it == adj_.end()) return;
    for (const auto& nb : it->second)
        if (nb.first != vertexId) adj_[nb.first].erase(vertexId);
    adj_.erase(it);
    vertices_.erase(vertexId);
}

uint64_t Graph::degree(int64_t vertexId) const {
    const auto it = adj_.find(vertexId);
    if (it == adj_.end()) throw std::runtime_error("Graph::degree: vertex not found");
    return it->second.size();
}

const std::unordered_map<int64_t, int64_t>& Graph::getEdges(int64_t vertexId) const {
    const auto it = adj_.find(vertexId);
    if (it == adj_.end()) throw std::runtime_error("Graph::getEdges: vertex not found");
    return it->second;
}

}  // namespace anneallib
