// rfmf_main -- command-line driver with the reference binary's interface
// (main.cu:20-103; printUsage, parseArgs and setupSolver at
// src/host_common.cpp:73-181,327-387, printChainResults and printSummary at
// src/host_common.cpp:389-443): the same flags, the
// same stdout blocks (device line, two partition lines per chain, summary)
// and the same stderr lines the reference's shell tests parse ("Resolved
// seeds:", "chain i: kernel-best-val: ... kernel-final-val: ...").  Every
// printed cut is recomputed on the host from the returned spins.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iomanip>
#include <memory>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/rfmf.h"
#include "../host/anneal_gpu.hpp"
#include "../host/tuning.hpp"

namespace {

constexpr uint64_t kSolverSeedOffset = 0xC2B2AE3D27D4EB4FULL;

struct Args {
    std::string filename;
    std::vector<uint64_t> seeds;
    std::vector<double> t0s;
    std::vector<size_t> phases;
    std::vector<uint64_t> iters;
    int device = 0;
};

void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " <filename> --seeds <N,N,...> --t0s <F,F,...>\n"
              << "       --phases <N,N,...> --iters <N,N,...> [--device N]\n"
              << "\nMulti-chain rejection-free multiple-flip Gumbel annealer for Max-Cut on one\n"
              << "MI355X.  One independent chain per (seed, t0) pair; the lists are zipped.\n"
              << "A 0 seed draws from random_device.  --phases gives the group count P per\n"
              << "phase and --iters the sweeps per phase, shared by every chain.  Chain i's\n"
              << "output is identical whether it runs alone or alongside any other chains.\n";
}

template <class T, class F>
std::vector<T> parseList(const std::string& s, F conv) {
    std::vector<T> out;
    std::istringstream is(s);
    std::string tok;
    while (std::getline(is, tok, ',')) out.push_back(conv(tok));
    return out;
}

Args parse(int argc, char** argv) {
    if (argc < 2) { usage(argv[0]); std::exit(1); }
    Args a;
    a.filename = argv[1];
    bool hs = false, hp = false, ht = false, hi = false;
    for (int i = 2; i < argc; i += 2) {
        const std::string flag = argv[i];
        if (i + 1 >= argc) { std::cerr << "rfmf_main: option " << flag << " needs a value\n"; std::exit(1); }
        const std::string val = argv[i + 1];
        try {
            if (flag == "--seed" || flag == "--t0") {
                std::cerr << "rfmf_main: " << flag << " is not accepted: every chain is one (seed, t0) pair, "
                             "given as --seeds a,b,... with --t0s x,y,... (lists of equal length)\n";
                std::exit(1);
            } else if (flag == "--seeds") {
                a.seeds = parseList<uint64_t>(val, [](const std::string& t) { return std::stoull(t); });
                hs = true;
            } else if (flag == "--t0s") {
                a.t0s = parseList<double>(val, [](const std::string& t) { return std::stod(t); });
                ht = true;
            } else if (flag == "--phases") {
                a.phases = parseList<size_t>(val, [](const std::string& t) { return static_cast<size_t>(std::stoull(t)); });
                hp = true;
            } else if (flag == "--iters") {
                a.iters = parseList<uint64_t>(val, [](const std::string& t) { return std::stoull(t); });
                hi = true;
            } else if (flag == "--device") {
                a.device = std::stoi(val);
            } else {
                std::cerr << "rfmf_main: unrecognized option " << flag << "\n";
                usage(argv[0]);
                std::exit(1);
            }
        } catch (const std::exception&) {
            std::cerr << "rfmf_main: cannot parse the value of " << flag << ": '" << val << "'\n";
            std::exit(1);
        }
    }
    // Every failed check is reported, then the usage text once.
    const auto count = [](size_t k) { return std::to_string(k); };
    const std::pair<bool, std::string> checks[] = {
        {!hs, "--seeds missing (required)"},
        {!hp, "--phases missing (required)"},
        {!ht, "--t0s missing (required)"},
        {!hi, "--iters missing (required)"},
        {ht && a.t0s.empty(), "--t0s lists no temperature"},
        {ht && std::any_of(a.t0s.begin(), a.t0s.end(), [](double t) { return !(t > 0.0); }),
         "--t0s entries must all be positive"},
        {hs && ht && a.seeds.size() != a.t0s.size(),
         "--seeds has " + count(a.seeds.size()) + " entries but --t0s has " + count(a.t0s.size()) +
             "; chain i uses seed i with t0 i"},
        {hp && hi && a.iters.size() != a.phases.size(),
         "--iters has " + count(a.iters.size()) + " entries but --phases has " + count(a.phases.size()) +
             "; phase p runs iters[p] sweeps with P = phases[p] groups"},
    };
    bool ok = true;
    for (const auto& c : checks)
        if (c.first) {
            std::cerr << "rfmf_main: " << c.second << "\n";
            ok = false;
        }
    if (!ok) { usage(argv[0]); std::exit(1); }
    return a;
}

std::string baseName(const std::string& path) {
    const auto pos = path.find_last_of('/');
    return pos == std::string::npos ? path : path.substr(pos + 1);
}

}  // namespace

int main(int argc, char** argv) {
    Args a = parse(argc, argv);

    rfmf_graph* gh = nullptr;
    rfmf_read_stats rs{};
    if (rfmf_graph_read_gset(a.filename.c_str(), 0.0, &gh, &rs) != RFMF_OK) {
        std::cerr << "Error: failed to read graph: " << a.filename << "\n";
        return 1;
    }
    // The graph owns the arrays `view` points into: it lives until main returns.
    const std::unique_ptr<rfmf_graph, void (*)(rfmf_graph*)> graphOwner(gh, rfmf_graph_free);
    rfmf_csr view{};
    rfmf_graph_view(gh, &view);
    const int64_t n = view.n;
    anneallib::AnnealCsr csr(std::vector<uint32_t>(view.offsets, view.offsets + n + 1),
                             std::vector<int32_t>(view.edge_to, view.edge_to + view.m2),
                             std::vector<int32_t>(view.edge_weight, view.edge_weight + view.m2),
                             std::vector<int64_t>(rfmf_graph_original_ids(gh), rfmf_graph_original_ids(gh) + n));
    std::cerr << "Loaded: n=" << rs.num_vertices << " m=" << rs.num_edges << " weight-scale=" << rs.scale
              << " max-weighted-degree=" << rs.max_abs_weighted_degree
              << (rs.float_weights ? " (float weights)" : "")
              << (rs.zero_dropped ? " zero-dropped=" + std::to_string(rs.zero_dropped) : std::string()) << "\n";

    for (uint64_t& s : a.seeds)
        if (s == 0) { std::random_device rd; s = rd(); }
    std::cerr << "Resolved seeds:";
    for (size_t i = 0; i < a.seeds.size(); ++i) std::cerr << (i ? "," : " ") << a.seeds[i];
    std::cerr << "\n";
    const size_t C = a.t0s.size();
    for (size_t i = 0; i < C; ++i)
        for (size_t j = i + 1; j < C; ++j)
            if (a.seeds[i] == a.seeds[j] && a.t0s[i] == a.t0s[j])
                std::cerr << "Warning: chains " << i << " and " << j << " have identical (seed, t0) = ("
                          << a.seeds[i] << ", " << a.t0s[i] << ") and will produce identical output.\n";

    std::vector<std::vector<int8_t>> init(C, std::vector<int8_t>(static_cast<size_t>(n)));
    std::vector<uint64_t> solverSeeds(C);
    for (size_t c = 0; c < C; ++c) {
        rfmf_initial_spins(a.seeds[c], n, init[c].data());
        solverSeeds[c] = a.seeds[c] + kSolverSeedOffset;
    }
    const int64_t B = rfmf_max_weighted_degree(&view);
    gpu3::LaunchConfig cfg;
    cfg.device = a.device;
    const double meanDeg = n > 0 ? static_cast<double>(view.m2) / static_cast<double>(n) : 0.0;
    cfg.corr = meanDeg >= gpu3::tuning::corrWarpwalkMinDeg ? gpu3::CorrMode::Warpwalk : gpu3::CorrMode::Walk;
    cfg.sWidth = B <= 2147483647LL ? 32 : 64;
    int64_t total = 0;
    if (rfmf_total_steps(a.iters.data(), static_cast<int64_t>(a.iters.size()), &total) != RFMF_OK) {
        std::cerr << rfmf_last_error() << "\n";
        return 1;
    }
    const double minT0 = *std::min_element(a.t0s.begin(), a.t0s.end());
    std::vector<double> distinct;
    for (double t : a.t0s)
        if (std::find(distinct.begin(), distinct.end(), t) == distinct.end()) distinct.push_back(t);
    std::cerr << "Chains: " << C << " distinct-t0: " << distinct.size()
              << " corr=" << (cfg.corr == gpu3::CorrMode::Warpwalk ? "warpwalk" : "walk") << " s-width=" << cfg.sWidth
              << " B=" << B << " total-sweeps-per-chain=" << total << " score-path="
              << (rfmf_score_safe_int32(total, minT0, B) ? "int32-packed" : "int64-twopass") << "\n";

    std::vector<std::vector<int8_t>> endSpins, bestSpins;
    float ms = 0.0f;
    std::vector<int64_t> kb, kf;
    try {
        gpu3::annealGpu(csr, B, init, a.t0s, a.phases, a.iters, solverSeeds, cfg, endSpins, bestSpins, &ms, &kb, &kf);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
    for (size_t c = 0; c < C; ++c)
        std::cerr << "chain " << c << ": kernel-best-val: " << kb[c] << " kernel-final-val: " << kf[c] << "\n";

    const double sec = static_cast<double>(ms) / 1000.0;
    std::vector<int64_t> bestVals(C);
    const std::string inst = baseName(a.filename);
    for (size_t c = 0; c < C; ++c) {
        const int64_t bv = rfmf_cut_value(&view, bestSpins[c].data());
        const int64_t ev = rfmf_cut_value(&view, endSpins[c].data());
        bestVals[c] = bv;
        std::cout << std::fixed << std::setprecision(4);
        std::cout << "device: GPU instance: " << inst << " --seed " << a.seeds[c] << " --phases ";
        for (size_t i = 0; i < a.phases.size(); ++i) std::cout << (i ? "," : "") << a.phases[i];
        std::cout << " --t0 " << a.t0s[c] << " --iters ";
        for (size_t i = 0; i < a.iters.size(); ++i) std::cout << (i ? "," : "") << a.iters[i];
        std::cout << " best-cut-value: " << bv << " final-cut-value: " << ev << " time-sec: " << sec << "\n";
        if (rs.scale != 1.0)
            std::cout << "weight-scale: " << rs.scale << " descaled-best-cut-value: " << static_cast<double>(bv) / rs.scale
                      << " descaled-final-cut-value: " << static_cast<double>(ev) / rs.scale << "\n";
        std::cout << "Solution-best-partition: ";
        for (int64_t i = 0; i < n; ++i) std::cout << (bestSpins[c][i] < 0 ? '0' : '1');
        std::cout << "\nSolution-final-partition: ";
        for (int64_t i = 0; i < n; ++i) std::cout << (endSpins[c][i] < 0 ? '0' : '1');
        std::cout << "\n";
    }
    size_t bestChain = 0;
    for (size_t i = 1; i < C; ++i)
        if (bestVals[i] > bestVals[bestChain]) bestChain = i;
    size_t atBest = 0;
    for (size_t i = 0; i < C; ++i) atBest += bestVals[i] == bestVals[bestChain];
    std::cout << std::fixed << std::setprecision(4) << "summary: chains: " << C << " best-chain: " << bestChain
              << " best-cut-value: " << bestVals[bestChain] << " time-sec: " << sec << " chains-at-best: " << atBest
              << "\n";
    return 0;
}
