// gpu_serve.cpp -- the served form as a process: interval requests arrive on
// stdin while earlier ones run, and each request gets the files one
// gpu_search_v42 process writes for its interval (src/gpu_search_v42.cu:
// 441-494), so scripts/coverage_audit.py (gpu_complete, :48-68) and the
// supervisors read them unchanged.
//
//   gpu_serve N_MAX [GPU]
//   stdin, one request per line:  LIMIT FIRST_L LAST_L PREFIX
//
// Per request: PREFIX-state.txt (version/limit/begin/end/resume/candidates,
// rewritten atomically whenever the finished prefix passes a step boundary
// begin + 512 j or reaches LAST_L + 1: resumes stay on the +512 chain across
// restarts), PREFIX-results.txt
// (appended "l m n area" lines, de-duplicated against the file, written before
// the STATE resume that covers them) and PREFIX.log (appended lines
// "PROGRESS resume R candidates C seconds S" for every R = begin + 512 j and
// LAST_L + 1 the finished prefix reaches -- the +512 chain the audit checks).
// An existing STATE resumes the interval as gpu_search does; a mismatching one
// skips the request with "checkpoint does not match requested range".  stdout
// gets "DONE PREFIX resume R candidates C" per finished request (R, C as in
// its STATE).  SIGTERM or
// SIGINT cancels every request (the launches in flight finish; STATE keeps the
// finished prefix) and exits 0; EOF on stdin waits for all requests.
#include <chrono>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>
#include <vector>

#include "cli_common.h"

namespace {

struct Job {
    uint64_t id = 0, limit = 0, begin = 0, end = 0;
    std::string prefix;
    uint64_t written = 0;       // the resume last written to STATE and the log
    uint64_t written_cand = 0;  // and its candidates
    std::set<std::array<uint64_t, 3>> seen;
    std::chrono::steady_clock::time_point t0;
};

std::mutex g_lines_mu;
std::vector<std::string> g_lines;
bool g_eof = false;

void read_stdin() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::lock_guard<std::mutex> lk(g_lines_mu);
        g_lines.push_back(line);
    }
    std::lock_guard<std::mutex> lk(g_lines_mu);
    g_eof = true;
}

// Results so far, STATE and PROGRESS lines for the finished prefix [begin,
// resume) cut down to its last step boundary.
bool report(hss_server* srv, Job& j, const hss_request_info& info) {
    const uint64_t r_end = info.resume > j.end ? j.end + 1 : j.begin + 512 * ((info.resume - j.begin) / 512);
    if (r_end <= j.written) return true;
    const uint64_t npos = j.end - j.begin + 1;
    std::vector<hss_root_result> per(npos);
    std::vector<hss_hit> hits(4096);
    uint32_t nh = 0;
    if (hss_server_results(srv, j.id, per.data(), npos, hits.data(), (uint32_t)hits.size(), &nh) != HSS_OK) {
        std::cerr << "gpu_serve: " << hss_last_error() << '\n';
        return false;
    }
    {
        std::ofstream out(j.prefix + "-results.txt", std::ios::app);
        for (uint32_t i = 0; i < nh; ++i) {
            if (!j.seen.insert({hits[i].l, hits[i].m, hits[i].n}).second) continue;
            out << hits[i].l << ' ' << hits[i].m << ' ' << hits[i].n << ' ' << cli::hit_area(hits[i]) << '\n';
        }
    }
    const uint64_t seconds =
        (uint64_t)std::chrono::duration<double>(std::chrono::steady_clock::now() - j.t0).count();
    std::ofstream log(j.prefix + ".log", std::ios::app);
    uint64_t cand = 0, at = j.begin;
    auto cand_upto = [&](uint64_t r) {  // candidates of [begin, r)
        for (; at < r; ++at) cand += per[at - j.begin].candidates;
        return cand;
    };
    for (uint64_t r = std::min(j.begin + 512 * ((j.written - j.begin) / 512 + 1), j.end + 1); r <= r_end;
         r = std::min(r + 512, j.end + 1)) {
        log << "PROGRESS resume " << r << " candidates " << cand_upto(r) << " seconds " << seconds << '\n';
        if (r == j.end + 1) break;
    }
    log.flush();
    j.written_cand = cand_upto(r_end);
    cli::write_state(j.prefix + "-state.txt", "version 1\nlimit " + std::to_string(j.limit) + "\nbegin " +
                                                  std::to_string(j.begin) + "\nend " + std::to_string(j.end) +
                                                  "\nresume " + std::to_string(r_end) + "\ncandidates " +
                                                  std::to_string(j.written_cand) + "\n");
    j.written = r_end;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || argc > 3) {
        std::cerr << "usage: gpu_serve N_MAX [GPU]  (requests on stdin: LIMIT FIRST_L LAST_L PREFIX)\n";
        return 2;
    }
    uint64_t nmax;
    int device;
    try {
        nmax = std::stoull(argv[1]);
        device = argc == 3 ? std::stoi(argv[2]) : 0;
    } catch (...) {
        std::cerr << "usage: gpu_serve N_MAX [GPU]\n";
        return 2;
    }
    cli::install_signals();
    hss_ctx* ctx = nullptr;
    if (hss_ctx_create(nmax, device, &ctx) != HSS_OK) {
        std::cerr << "gpu_serve: " << hss_last_error() << '\n';
        return 1;
    }
    hss_server* srv = nullptr;
    if (hss_server_create(ctx, nullptr, &srv) != HSS_OK) {
        std::cerr << "gpu_serve: " << hss_last_error() << '\n';
        hss_ctx_destroy(ctx);
        return 1;
    }
    std::thread reader(read_stdin);
    reader.detach();  // may stay blocked on stdin after a signal; the process exits anyway
    std::map<uint64_t, Job> jobs;
    bool canceled = false, eof = false;
    int status = 0;
    while (!eof || !jobs.empty()) {
        std::vector<std::string> lines;
        {
            std::lock_guard<std::mutex> lk(g_lines_mu);
            lines.swap(g_lines);
            eof = g_eof;
        }
        for (const std::string& line : lines) {
            std::istringstream in(line);
            Job j;
            if (!(in >> j.limit >> j.begin >> j.end >> j.prefix)) continue;  // blank or malformed line
            if (canceled) continue;
            const std::string state = j.prefix + "-state.txt";
            auto kv = cli::read_state(state);
            const uint64_t stored_limit = kv.count("limit") ? kv["limit"] : 0;
            const uint64_t stored_end = kv.count("end") ? kv["end"] : 0;
            const uint64_t saved = kv.count("resume") ? kv["resume"] : j.begin;
            if (stored_limit && (stored_limit != j.limit || stored_end != j.end || saved < j.begin)) {
                std::cerr << j.prefix << ": checkpoint does not match requested range\n";
                continue;
            }
            if (saved > j.end) {
                std::cout << "DONE " << j.prefix << " resume " << saved << " candidates 0" << std::endl;
                continue;
            }
            j.begin = saved;
            j.written = saved;
            j.seen = cli::read_seen(j.prefix + "-results.txt");
            j.t0 = std::chrono::steady_clock::now();
            if (hss_server_submit(srv, j.limit, j.begin, j.end, &j.id) != HSS_OK) {
                std::cerr << j.prefix << ": " << hss_last_error() << '\n';
                continue;
            }
            jobs.emplace(j.id, std::move(j));
        }
        if (cli::stop_requested.load() && !canceled) {
            canceled = true;
            for (auto& kv : jobs) (void)hss_server_cancel(srv, kv.first);
        }
        for (auto it = jobs.begin(); it != jobs.end();) {
            hss_request_info info;
            const int rc = hss_server_wait(srv, it->first, 0, &info);
            if (rc != HSS_OK && rc != HSS_EAGAIN) {
                std::cerr << it->second.prefix << ": " << hss_last_error() << '\n';
                status = 1;
                it = jobs.erase(it);
                continue;
            }
            if (!report(srv, it->second, info)) status = 1;
            if (rc == HSS_OK) {
                std::cout << "DONE " << it->second.prefix << " resume " << it->second.written << " candidates "
                          << it->second.written_cand << std::endl;
                (void)hss_server_release(srv, it->first);
                it = jobs.erase(it);
            } else {
                ++it;
            }
        }
        if (canceled && jobs.empty()) break;
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
    }
    hss_server_destroy(srv);
    hss_ctx_destroy(ctx);
    return status;
}
