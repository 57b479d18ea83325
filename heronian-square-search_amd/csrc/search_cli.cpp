// search_cli.cpp -- the reference's CPU `search` command line
// (src/search_v18.cpp:532-573), served by the GPU core.
//
//   search LIMIT FIRST_L LAST_L THREADS STATE_FILE [RESULTS_FILE]
//
// Keeps the argument validation (exit 2), the checkpoint rules ("checkpoint
// does not match requested range" -> exit 2, "checkpoint already complete" ->
// exit 0), STATE keys version/limit/begin/end/resume/pairs/candidates, the
// "HIT l m n area A" and "PROGRESS resume R completed_l D candidates C seconds S
// pairs/s X" stdout lines, and the coverage skip of HERON_COVERAGE_FILE
// (default coverage-latest.json, src/search_v18.cpp:254-283): positions inside
// the report's complete_intervals count as done without being searched.  As in
// the reference (src/search_v18.cpp:213-222) the report is read again every
// 10 seconds (between batches; covered intervals only accumulate) and a
// PROGRESS line is printed every 10 seconds and at the end; STATE is rewritten
// after every batch.  THREADS is accepted for compatibility; the work runs on
// HIP device $HSS_DEVICE (default 0).  Positions are searched in batches of at
// most 512 and at most BATCH_CELLS planner cells (hss_estimate), so a SIGTERM
// (src/search_v18.cpp:330-335) ends the run within about a second of device
// work.  HSS_SEARCH_REFRESH_S and HSS_SEARCH_PROGRESS_S override the two
// 10-second periods (tests use 0: every batch).
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "cli_common.h"

namespace {

// complete_intervals of a coverage report for this LIMIT (empty on any mismatch)
std::vector<std::pair<uint64_t, uint64_t>> read_coverage(uint64_t n) {
    std::vector<std::pair<uint64_t, uint64_t>> out;
    const char* env = std::getenv("HERON_COVERAGE_FILE");
    std::ifstream in(env ? env : "coverage-latest.json");
    if (!in) return out;
    std::string r((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t lim = r.find("\"limit\""), iv = r.find("\"complete_intervals\""), endp = r.find("\"root_coverage\"");
    if (lim == std::string::npos || iv == std::string::npos || endp == std::string::npos || lim >= iv || iv >= endp)
        return out;
    size_t colon = r.find(':', lim);
    if (colon == std::string::npos || colon >= iv) return out;
    try {
        if (std::stoull(r.substr(colon + 1)) != n) return out;
    } catch (...) {
        return out;
    }
    // every "[ a , b ]" between the two keys
    size_t at = iv;
    std::vector<std::pair<uint64_t, uint64_t>> got;
    while (true) {
        size_t open = r.find('[', at + 1);
        if (open == std::string::npos || open >= endp) break;
        size_t close = r.find(']', open);
        if (close == std::string::npos || close >= endp) break;
        std::string inner = r.substr(open + 1, close - open - 1);
        size_t comma = inner.find(',');
        at = open;
        if (comma == std::string::npos) continue;
        auto digits = [](const std::string& s, uint64_t& v) {
            size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
            if (a == std::string::npos) return false;
            std::string t = s.substr(a, b - a + 1);
            if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos) return false;
            v = std::stoull(t);
            return true;
        };
        uint64_t a, b;
        if (!digits(inner.substr(0, comma), a) || !digits(inner.substr(comma + 1), b)) continue;
        if (!a || b < a || b > n) return out;
        got.emplace_back(a, b);
    }
    return got;
}

// a period in seconds from the environment (default when unset or invalid)
double period_from_env(const char* name, double dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    char* end = nullptr;
    const double x = std::strtod(v, &end);
    return (end && *end == 0 && x >= 0) ? x : dflt;
}

// planner cells per batch: about a second of device work at LIMIT 1e7
constexpr double BATCH_CELLS = 5e11;

}  // namespace

int main(int argc, char** argv) {
    if (argc < 6 || argc > 7) {
        std::cerr << "usage: search LIMIT FIRST_L LAST_L THREADS STATE_FILE [RESULTS_FILE]\n";
        return 2;
    }
    uint64_t n, first, last;
    unsigned long threads;
    try {
        n = std::stoull(argv[1]);
        first = std::stoull(argv[2]);
        last = std::stoull(argv[3]);
        threads = std::stoul(argv[4]);
    } catch (...) {
        return 2;
    }
    if (!n || !first || first > last || last > n || !threads) return 2;
    const std::string state_file = argv[5];
    auto kv = cli::read_state(state_file);
    uint64_t stored_limit = kv.count("limit") ? kv["limit"] : 0;
    if (stored_limit) {
        uint64_t stored_end = kv.count("end") ? kv["end"] : 0;
        uint64_t saved = kv.count("resume") ? kv["resume"] : first;
        if (stored_limit != n || stored_end != last || saved < first) {
            std::cerr << "checkpoint does not match requested range\n";
            return 2;
        }
        first = saved;
        if (first > last) {
            std::cout << "checkpoint already complete\n";
            return 0;
        }
        std::cerr << "resuming at " << first << '\n';
    }
    const std::string results_file = argc == 7 ? argv[6] : "results.txt";
    cli::install_signals();
    const char* dev = std::getenv("HSS_DEVICE");
    auto t0 = std::chrono::steady_clock::now();
    hss_ctx* ctx = nullptr;
    if (hss_ctx_create(n, dev ? std::atoi(dev) : 0, &ctx) != HSS_OK) {
        std::cerr << "search: " << hss_last_error() << '\n';
        return 1;
    }
    std::cerr << "prepared in " << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()
              << " seconds\n";
    { std::ofstream touch(results_file, std::ios::app); }
    auto seen = cli::read_seen(results_file);
    // covered root intervals; a refresh adds the report's intervals, never removes
    std::vector<std::pair<uint64_t, uint64_t>> covered;
    auto refresh = [&]() {
        for (auto& iv : read_coverage(n))
            if (std::find(covered.begin(), covered.end(), iv) == covered.end()) covered.push_back(iv);
    };
    refresh();
    auto is_covered = [&](uint64_t l) {
        for (auto& iv : covered)
            if (iv.first <= l && l <= iv.second) return true;
        return false;
    };
    const double refresh_s = period_from_env("HSS_SEARCH_REFRESH_S", 10.0);
    const double progress_s = period_from_env("HSS_SEARCH_PROGRESS_S", 10.0);
    const uint64_t begin = first;
    uint64_t next = first, pairs = 0, candidates = 0, completed = 0;
    auto start = std::chrono::steady_clock::now();
    auto last_refresh = start, last_progress = start;
    auto since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    };
    auto save = [&]() {
        cli::write_state(state_file, "version 1\nlimit " + std::to_string(n) + "\nbegin " + std::to_string(begin) +
                                         "\nend " + std::to_string(last) + "\nresume " + std::to_string(next) +
                                         "\npairs " + std::to_string(pairs) + "\ncandidates " +
                                         std::to_string(candidates) + "\n");
    };
    auto progress = [&]() {
        double s = since(start);
        std::cout << "PROGRESS resume " << next << " completed_l " << completed << " candidates " << candidates
                  << " seconds " << (uint64_t)s << " pairs/s " << (uint64_t)(pairs / std::max(s, 1.0))
                  << std::endl;
        last_progress = std::chrono::steady_clock::now();
    };
    std::vector<hss_hit> hits(4096);
    std::vector<uint64_t> roots;
    while (next <= last && !cli::stop_requested.load()) {
        if (since(last_refresh) >= refresh_s) {
            refresh();
            last_refresh = std::chrono::steady_clock::now();
        }
        // up to 512 positions and BATCH_CELLS cells of uncovered positions (at least one)
        uint64_t batch = 0;
        double cells = 0;
        roots.clear();
        while (batch < 512 && next + batch <= last) {
            const uint64_t l = next + batch;
            if (!is_covered(l)) {
                double e = 0;
                if (hss_estimate(ctx, n, l, l, &e) != HSS_OK) e = 0;
                if (!roots.empty() && cells + e > BATCH_CELLS) break;
                cells += e;
                roots.push_back(l);
            }
            ++batch;
        }
        uint64_t c = 0;
        uint32_t nh = 0;
        if (!roots.empty()) {
            int rc = hss_search(ctx, n, roots.data(), (uint32_t)roots.size(), nullptr, &c, hits.data(),
                                (uint32_t)hits.size(), &nh);
            if (rc != HSS_OK) {
                std::cerr << "search: " << hss_last_error() << '\n';
                hss_ctx_destroy(ctx);
                return 1;
            }
        }
        {
            std::ofstream out(results_file, std::ios::app);
            for (uint32_t i = 0; i < nh; ++i) {
                if (!seen.insert({hits[i].l, hits[i].m, hits[i].n}).second) continue;
                std::string area = cli::hit_area(hits[i]);
                out << hits[i].l << ' ' << hits[i].m << ' ' << hits[i].n << ' ' << area << '\n';
                std::cout << "HIT " << hits[i].l << ' ' << hits[i].m << ' ' << hits[i].n << " area " << area
                          << std::endl;
            }
        }
        candidates += c;
        for (uint64_t l = next; l < next + batch; ++l) pairs += n - l + 1;
        completed += batch;
        next += batch;
        save();
        if (since(last_progress) >= progress_s) progress();
    }
    save();
    progress();
    hss_ctx_destroy(ctx);
    return 0;
}
