// read_gset.cpp -- behaviour of the reference loader (src/read_gset.cpp:
// 17-214): the scale rule, rounding, zero-drop and dedup order must match it
// exactly, because they decide the integer instance every result is about.
#include "read_gset.hpp"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <unordered_map>
#include <vector>

namespace anneallib {
namespace {

bool isBlankOrComment(const std::string& line) {
    for (char c : line) {
        if (c == ' ' || c == '\t' || c == '\r') continue;
        return c == '#';
    }
    return true;
}

bool readDataLine(std::istream& in, std::string& line) {
    while (std::getline(in, line))
        if (!isBlankOrComment(line)) return true;
    return false;
}

// Exactly representable as an int64 through a double.
bool integralWeight(double w) { return std::floor(w) == w && std::fabs(w) < 9007199254740992.0; }

struct RawEdge {
    int64_t u, v;
    double w;
};

// Largest 10^k, k = 12 .. -12, with maxB*s + maxDeg/2 <= budgetB and
// maxW*s <= INT32_MAX/2; 0 if none.
double pickScale(double maxB, double maxW, double maxDeg, double budgetB) {
    const double budgetW = static_cast<double>(INT32_MAX) / 2;
    for (int k = 12; k >= -12; --k) {
        const double s = std::pow(10.0, k);
        if (maxB * s + maxDeg / 2 <= budgetB && maxW * s <= budgetW) return s;
    }
    return 0.0;
}

// "u v w": two integer endpoints and a weight (integer or decimal), each
// field required; anything after the weight is ignored.
bool parseEdgeLine(const std::string& line, RawEdge& e) {
    const char* p = line.c_str();
    char* next = nullptr;
    e.u = std::strtoll(p, &next, 10);
    bool ok = next != p;
    p = next;
    e.v = std::strtoll(p, &next, 10);
    ok = ok && next != p;
    p = next;
    e.w = std::strtod(p, &next);
    return ok && next != p;
}

}  // namespace

bool read_instance_from_stream(std::istream& in, Graph& g, ReadStats* stats, double forcedScale) {
    std::string line;
    if (!readDataLine(in, line)) {
        std::cerr << "read_gset: missing header line '<vertices> <edges>'\n";
        return false;
    }
    int64_t headerN = 0, headerM = 0;
    {
        std::istringstream hs(line);
        if (!(hs >> headerN >> headerM) || headerN < 0 || headerM < 0) {
            std::cerr << "read_gset: header is not two non-negative integers: '" << line << "'\n";
            return false;
        }
    }

    std::vector<RawEdge> raw;
    raw.reserve(static_cast<size_t>(headerM));
    bool anyFloat = false;
    for (int64_t i = 0; i < headerM; ++i) {
        if (!readDataLine(in, line)) {
            std::cerr << "read_gset: the file ends after " << i << " of the header's " << headerM << " edges\n";
            return false;
        }
        RawEdge e{};
        if (!parseEdgeLine(line, e)) {
            std::cerr << "read_gset: edge " << i << " is not '<u> <v> <weight>': '" << line << "'\n";
            return false;
        }
        if (!integralWeight(e.w)) anyFloat = true;
        raw.push_back(e);
    }

    // Pre-scale bounds over the raw (un-deduped) edge list, per-vertex sums
    // accumulated in file order.
    double maxB = 0.0, maxW = 0.0, maxDeg = 0.0;
    {
        std::unordered_map<int64_t, std::pair<double, int64_t>> acc;
        for (const RawEdge& e : raw) {
            const double aw = std::fabs(e.w);
            if (aw > maxW) maxW = aw;
            auto& a = acc[e.u];
            a.first += aw; a.second += 1;
            auto& b = acc[e.v];
            b.first += aw; b.second += 1;
        }
        for (const auto& kv : acc) {
            if (kv.second.first > maxB) maxB = kv.second.first;
            if (static_cast<double>(kv.second.second) > maxDeg) maxDeg = static_cast<double>(kv.second.second);
        }
    }

    double scale = 1.0;
    if (forcedScale > 0.0) {
        scale = forcedScale;
    } else if (anyFloat) {
        scale = pickScale(maxB, maxW, maxDeg, static_cast<double>(INT32_MAX) / 4);
        if (scale == 0.0) scale = pickScale(maxB, maxW, maxDeg, static_cast<double>(INT64_MAX) / 8);
        if (scale == 0.0) {
            std::cerr << "read_gset: weights up to " << maxW << " (weighted degree up to " << maxB
                      << ") fit no power-of-ten scale with int32 weights and int64 fields\n";
            return false;
        }
        if (maxW * scale < 0.5) {
            std::cerr << "read_gset: scale " << scale << " rounds every weight (largest |w| " << maxW
                      << ") to zero\n";
            return false;
        }
    }

    int64_t dropped = 0;
    for (const RawEdge& e : raw) {
        const int64_t wi = (scale == 1.0 && integralWeight(e.w)) ? static_cast<int64_t>(e.w)
                                                                 : std::llround(e.w * scale);
        if (wi == 0) { ++dropped; continue; }
        g.addEdge(e.u, e.v, wi);
    }

    int64_t maxAbsB = 0;
    for (int64_t v : g.getVertices()) {
        int64_t s = 0;
        for (const auto& nb : g.getEdges(v)) s += std::llabs(static_cast<long long>(nb.second));
        if (s > maxAbsB) maxAbsB = s;
    }
    if (stats) {
        stats->numVertices = headerN;
        stats->numEdges = headerM;
        stats->zeroDropped = dropped;
        stats->floatWeights = anyFloat;
        stats->scale = scale;
        stats->maxAbsWeightedDegree = maxAbsB;
    }
    return true;
}

bool read_instance_from_file(const std::string& filename, Graph& g, ReadStats* stats,
                             double forcedScale) {
    std::ifstream f(filename);
    if (!f.is_open()) {
        std::cerr << "read_gset: cannot open " << filename << "\n";
        return false;
    }
    return read_instance_from_stream(f, g, stats, forcedScale);
}

}  // namespace anneallib
