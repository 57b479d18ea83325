// gpu_search.cpp -- drop-in replacement for the reference's gpu_search_v42
// executable (src/gpu_search_v42.cu:610-632 and Search::run :393-506), and
// with a seventh argument for its heavy-weight variant v36
// (experiments/heavy/gpu_search_v36.cu:305-308, 442-516, 637-652).
//
//   gpu_search LIMIT FIRST_L LAST_L RESULTS_FILE [GPU [STATE_FILE [MIN_WEIGHT]]]
//
// Same argv, STATE keys (version/limit/begin/end/resume/candidates), results
// lines and stdout "PROGRESS resume R candidates C seconds S" chain of batches
// of exactly min(512, remaining) positions, so scripts/gpu_queue.py and
// scripts/coverage_audit.py accept its output.  With MIN_WEIGHT (v36): only
// roots of Gaussian weight >= MIN_WEIGHT are searched, batches are 4096
// positions when MIN_WEIGHT > 0, STATE carries `min_weight` and keeps `begin`
// at FIRST_L, a mismatching checkpoint reports "range or weight", and after
// every batch STATE.done lists "limit", "min_weight" and every searched root
// of the finished prefix (for experiments/heavy/verify_heavy.py and
// scripts/coverage_audit.py:71-99).  Differences: the candidate counter and hit
// list restart per batch on the device (no silent loss past 4096 hits: an
// overflow is an error), and device errors exit with status 1.
#include <chrono>
#include <iostream>
#include <vector>

#include "cli_common.h"

static const char* kUsage = "usage: gpu_search LIMIT FIRST_L LAST_L RESULTS_FILE [GPU [STATE_FILE [MIN_WEIGHT]]]\n";

int main(int argc, char** argv) {
    if (argc < 5 || argc > 8) {
        std::cerr << kUsage;
        return 2;
    }
    uint64_t n, begin, end;
    int device;
    uint32_t min_weight = 0;
    try {
        n = std::stoull(argv[1]);
        begin = std::stoull(argv[2]);
        end = std::stoull(argv[3]);
        device = argc >= 6 ? std::stoi(argv[5]) : 0;
        if (argc == 8) min_weight = (uint32_t)std::stoul(argv[7]);
    } catch (...) {
        std::cerr << kUsage;
        return 2;
    }
    const bool heavy_form = argc == 8;  // the v36 file conventions
    const uint64_t range_begin = begin;
    const std::string results_file = argv[4];
    const std::string state_file = argc >= 7 ? argv[6] : "";
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
        uint64_t stored_weight = kv.count("min_weight") ? kv["min_weight"] : 0;
        if (stored_limit && (stored_limit != n || stored_end != end || saved < begin ||
                             (heavy_form && stored_weight != min_weight))) {
            std::cerr << (heavy_form ? "checkpoint does not match requested range or weight\n"
                                     : "checkpoint does not match requested range\n");
            hss_ctx_destroy(ctx);
            return 0;
        }
        begin = saved;
    }
    // roots of the finished prefix searched under MIN_WEIGHT (for the .done file)
    std::vector<uint64_t> done_roots;
    if (min_weight)
        for (uint64_t l = range_begin; l < begin; ++l)
            if (hss_gaussian_weight(l) >= min_weight) done_roots.push_back(l);
    std::ofstream out(results_file, std::ios::app);
    auto seen = cli::read_seen(results_file);
    auto start = std::chrono::steady_clock::now();
    uint64_t next = begin, candidates = 0;
    std::vector<hss_hit> hits(4096);
    const uint64_t batch_size = min_weight ? 4096 : 512;
    while (next <= end && !cli::stop_requested.load()) {
        uint64_t batch = std::min<uint64_t>(batch_size, end - next + 1);
        uint64_t c = 0;
        uint32_t nh = 0;
        int rc;
        if (min_weight) {
            std::vector<uint64_t> roots;
            for (uint64_t l = next; l < next + batch; ++l)
                if (hss_gaussian_weight(l) >= min_weight) roots.push_back(l);
            rc = hss_search(ctx, n, roots.data(), (uint32_t)roots.size(), nullptr, &c, hits.data(),
                            (uint32_t)hits.size(), &nh);
            done_roots.insert(done_roots.end(), roots.begin(), roots.end());
        } else {
            rc = hss_search_interval(ctx, n, next, next + batch - 1, &c, hits.data(), (uint32_t)hits.size(), &nh);
        }
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
        if (!state_file.empty() && min_weight) {
            std::string body = "limit " + std::to_string(n) + "\nmin_weight " + std::to_string(min_weight) + "\n";
            for (uint64_t l : done_roots) body += std::to_string(l) + "\n";
            cli::write_state(state_file + ".done", body);
        }
        if (!state_file.empty()) {
            std::string body = "version 1\nlimit " + std::to_string(n) + "\nbegin " +
                               std::to_string(heavy_form ? range_begin : begin) + "\nend " + std::to_string(end) +
                               "\nresume " + std::to_string(next) + "\ncandidates " + std::to_string(candidates) +
                               "\n";
            if (heavy_form) body += "min_weight " + std::to_string(min_weight) + "\n";
            cli::write_state(state_file, body);
        }
        double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::cout << "PROGRESS resume " << next << " candidates " << candidates << " seconds " << (uint64_t)seconds
                  << std::endl;
    }
    hss_ctx_destroy(ctx);
    return 0;
}
