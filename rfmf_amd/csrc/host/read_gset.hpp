// read_gset.hpp -- Gset / MQLib instance reader, API of the reference's
// read_gset.hpp (src/read_gset.hpp:11-34).
//
// Format: blank lines and lines whose first non-blank character is '#' are
// skipped (CRLF tolerated); a header "<numVertices> <numEdges>"; then one
// "<u> <v> <weight>" line per edge.  Integral weights are kept as is.  If any
// weight is non-integral, every weight is multiplied by the largest power of
// ten (10^12 .. 10^-12) that keeps the max weighted degree within
// INT32_MAX/4 (else INT64_MAX/8) and every weight within INT32_MAX/2, then
// rounded; forcedScale > 0 overrides the choice.  Edges whose integer weight
// is 0 are dropped and counted.  A repeated edge keeps its first weight.
#pragma once

#include <cstdint>
#include <istream>
#include <string>

#include "graph.hpp"

namespace anneallib {

struct ReadStats {
    int64_t numVertices = 0;           // header N
    int64_t numEdges = 0;              // header M
    int64_t zeroDropped = 0;           // edges whose integer weight was 0
    bool floatWeights = false;         // a non-integral weight was seen
    double scale = 1.0;                // w_int = llround(scale * w)
    int64_t maxAbsWeightedDegree = 0;  // max_v sum |w_int| after dedup
};

bool read_instance_from_stream(std::istream& in, Graph& g, ReadStats* stats = nullptr,
                               double forcedScale = 0.0);
bool read_instance_from_file(const std::string& filename, Graph& g, ReadStats* stats = nullptr,
                             double forcedScale = 0.0);

}  // namespace anneallib
