// gpu_search.cpp -- drop-in replacement for the reference's gpu_search_v42
// executable (src/gpu_search_v42.cu:610-632 and Search::run :393-506).
//
//   gpu_search LIMIT FIRST_L LAST_L RESULTS_FILE [GPU [STATE_FILE]]
//
// Same argv, STATE keys (version/limit/begin/end/resume/candidates), results
// lines and stdout "PROGRESS resume R candidates C seconds S" chain of batches
// of exactly min(512, remaining) positions, so scripts/gpu_queue.py and
// scripts/coverage_audit.py accept its output.  Differences: the candidate
// counter and hit list restart per batch on the device (no silent loss past 4096
// hits: an overflow is an error), and device errors exit with status 1.
#include <chrono>
#include <iostream>
#include <vector>

#include "cli_common.h"

int main(int argc, char** argv) {
    if (argc < 5 || argc > 7) {
        std::cerr << "usage: gpu_search LIMIT FIRST_L LAST_L RESULTS_FILE [GPU [STATE_FILE]]\n";
        return 2;
    }
    uint64_t n, begin, end;
    int device;
    try {
        n = std::stoull(argv[1]);
        begin = std::stoull(argv[2]);
        end = std::stoull(argv[3]);
        device = argc >= 6 ? std::stoi(argv[5]) : 0;
    } catch (...) {
        std::cerr << "usage: gpu_search LIMIT FIRST_L LAST_L RESULTS_FILE [GPU [STATE_FILE]]\n";
        return 2;
    }
    const std::string results_file = argv[4];
    const std::string state_file = argc == 7 ? argv[6] : "";
    cli::install_signals();
    auto t0 = std::chrono::steady_clock::now();
    hss_ctx* ctx = nullptr;
    if (hss_ctx_create(n, device, &ctx) != HSS_OK) {
        std::cerr << "gpu_search: " << hss_last_error() << '\n';
        return 1;
    }
    std::cerr << "prepared in " << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()
              << " seconds\n";
    if (!state_file.empty()) {
        auto kv = cli::read_state(state_file);
        uint64_t stored_limit = kv.count("limit") ? kv["limit"] : 0;
        uint64_t stored_end = kv.count("end") ? kv["end"] : 0;
        uint64_t saved = kv.count("resume") ? kv["resume"] : begin;
        if (stored_limit && (stored_limit != n || stored_end != end || saved < begin)) {
            std::cerr << "checkpoint does not match requested range\n";
            hss_ctx_destroy(ctx);
            return 0;
        }
        begin = saved;
    }
    std::ofstream out(results_file, std::ios::app);
    auto seen = cli::read_seen(results_file);
    auto start = std::chrono::steady_clock::now();
    uint64_t next = begin, candidates = 0;
    std::vector<hss_hit> hits(4096);
    while (next <= end && !cli::stop_requested.load()) {
        uint64_t batch = std::min<uint64_t>(512, end - next + 1);
        uint64_t c = 0;
        uint32_t nh = 0;
        int rc = hss_search_interval(ctx, n, next, next + batch - 1, &c, hits.data(), (uint32_t)hits.size(), &nh);
        if (rc != HSS_OK) {
            std::cerr << "gpu_search: " << hss_last_error() << '\n';
            hss_ctx_destroy(ctx);
            return 1;
        }
        candidates += c;
        for (uint32_t i = 0; i < nh; ++i) {
            if (!seen.insert({hits[i].l, hits[i].m, hits[i].n}).second) continue;
            out << hits[i].l << ' ' << hits[i].m << ' ' << hits[i].n << ' ' << cli::hit_area(hits[i]) << '\n';
        }
        out.flush();
        next += batch;
        if (!state_file.empty())
            cli::write_state(state_file, "version 1\nlimit " + std::to_string(n) + "\nbegin " + std::to_string(begin) +
                                             "\nend " + std::to_string(end) + "\nresume " + std::to_string(next) +
                                             "\ncandidates " + std::to_string(candidates) + "\n");
        double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::cout << "PROGRESS resume " << next << " candidates " << candidates << " seconds " << (uint64_t)seconds
                  << std::endl;
    }
    hss_ctx_destroy(ctx);
    return 0;
}
