// ref_driver.cpp -- test infrastructure: a small command-line driver over the
// reference checkout's own host functions (compiled from its sources by
// oracle/Makefile), used to write golden vectors and to cross-check the
// oracle.  It contains no reference code; it only calls the reference API.
//
//   ref_driver csr <gset>                 n, m2, B, scale, then offsets / to / w
//   ref_driver spins <raw seed> <n>       initial spins as a 0/1 string
//   ref_driver beta <total> <t0> [t0..]   run tables: "end val" per run, chain offsets
//   ref_driver gate <total> <t0> <B>      annealScoreSafeInt32 verdict
//   ref_driver betafix <step> <t0>        betaFixedForStep
//   ref_driver cut <gset> <bits>          cutValueCsr of a partition string
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "host_common.h"
#include "philox.cuh"
#include "rng.hpp"

namespace {

anneallib::AnnealCsr load(const std::string& path, anneallib::ReadStats* st) {
    anneallib::Graph g;
    if (!anneallib::read_instance_from_file(path, g, st, 0.0)) std::exit(2);
    removeDegreeZeroVertices(g);
    return anneallib::AnnealCsr(g);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "csr" && argc == 3) {
        anneallib::ReadStats st;
        const auto csr = load(argv[2], &st);
        const int64_t n = static_cast<int64_t>(csr.numVertices());
        std::cout << n << " " << csr.numEdges() * 2 << " " << maxWeightedDegree(csr) << " " << st.scale << " "
                  << st.zeroDropped << " " << st.maxAbsWeightedDegree << "\n";
        std::vector<int32_t> to, w;
        std::cout << 0;
        uint64_t off = 0;
        for (int64_t v = 0; v < n; ++v) {
            const auto nb = csr.neighbors(v);
            off += nb.size();
            std::cout << " " << off;
            for (uint32_t i = 0; i < nb.size(); ++i) { to.push_back(nb.to[i]); w.push_back(nb.weight[i]); }
        }
        std::cout << "\n";
        for (size_t i = 0; i < to.size(); ++i) std::cout << (i ? " " : "") << to[i];
        std::cout << "\n";
        for (size_t i = 0; i < w.size(); ++i) std::cout << (i ? " " : "") << w[i];
        std::cout << "\n";
        return 0;
    }
    if (cmd == "spins" && argc == 4) {
        anneallib::Rng rng(std::stoull(argv[2]));
        const uint64_t n = std::stoull(argv[3]);
        std::string s(n, '0');
        for (uint64_t i = 0; i < n; ++i) s[i] = rng.bits(1) ? '1' : '0';
        std::cout << s << "\n";
        return 0;
    }
    if (cmd == "beta" && argc >= 4) {
        const int64_t total = std::stoll(argv[2]);
        std::vector<double> t0s;
        for (int i = 3; i < argc; ++i) t0s.push_back(std::stod(argv[i]));
        const BetaRunTables t = buildBetaRunTables(t0s, total);
        std::cout << t.runEnd.size() << " " << t.nTables << "\n";
        for (size_t c = 0; c < t.chainRunOff.size(); ++c) std::cout << (c ? " " : "") << t.chainRunOff[c];
        std::cout << "\n";
        for (size_t r = 0; r < t.runEnd.size(); ++r) std::cout << t.runEnd[r] << " " << t.runVal[r] << "\n";
        return 0;
    }
    if (cmd == "gate" && argc == 5) {
        std::cout << (annealScoreSafeInt32(std::stoll(argv[2]), std::stod(argv[3]), std::stoll(argv[4])) ? 1 : 0)
                  << "\n";
        return 0;
    }
    if (cmd == "betafix" && argc == 4) {
        std::cout << betaFixedForStep(std::stoull(argv[2]), std::stod(argv[3])) << "\n";
        return 0;
    }
    if (cmd == "philox" && argc == 8) {
        gpu3::Philox4 c;
        gpu3::Philox2 k;
        for (int i = 0; i < 4; ++i) c.v[i] = static_cast<uint32_t>(std::stoul(argv[2 + i]));
        for (int i = 0; i < 2; ++i) k.v[i] = static_cast<uint32_t>(std::stoul(argv[6 + i]));
        const gpu3::Philox4 o = gpu3::philox4x32_10(c, k);
        std::cout << o.v[0] << " " << o.v[1] << " " << o.v[2] << " " << o.v[3] << "\n";
        return 0;
    }
    if (cmd == "gumbels" && argc == 5) {
        // Gumbel of vertices [0, n) in 1-based sweep `step` of the chain keyed by `solverSeed`.
        const uint64_t seed = std::stoull(argv[2]);
        const uint64_t step = std::stoull(argv[3]);
        const uint64_t n = std::stoull(argv[4]);
        const gpu3::Philox2 key = gpu3::philoxKeyFromSeed(seed);
        for (uint64_t v = 0; v < n; ++v) {
            const gpu3::Philox4 blk = gpu3::philox4x32_10(gpu3::gumbelBlockCounter(step, v >> 3), key);
            std::cout << (v ? " " : "")
                      << gpu3::gumbelFromBitsG(gpu3::gumbelBitsFromBlock(blk, static_cast<uint32_t>(v & 7)));
        }
        std::cout << "\n";
        return 0;
    }
    if (cmd == "gumbelbits") {
        for (uint32_t r = 0; r < 2048; ++r) std::cout << (r ? " " : "") << gpu3::gumbelFromBitsG(r);
        std::cout << "\n";
        return 0;
    }
    if (cmd == "pack" && argc == 4) {
        std::cout << gpu3::packScoreId(static_cast<int32_t>(std::stol(argv[2])), static_cast<int32_t>(std::stol(argv[3])))
                  << "\n";
        return 0;
    }
    if (cmd == "cut" && argc == 4) {
        const auto csr = load(argv[2], nullptr);
        const std::string bits = argv[3];
        if (bits.size() != csr.numVertices()) return 3;
        std::vector<int8_t> spin(bits.size());
        for (size_t i = 0; i < bits.size(); ++i) spin[i] = bits[i] == '1' ? 1 : -1;
        std::cout << anneallib::solver::cutValueCsr(csr, spin) << "\n";
        return 0;
    }
    return 2;
}
