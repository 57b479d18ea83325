// cli_common.h -- file formats shared by the gpu_search and search executables.
//
// STATE files are "key <int>" lines written through a temporary file and an
// atomic rename (src/gpu_search_v42.cu:483-491, src/search_v18.cpp:509-521);
// results files hold "l m n area" lines, appended, with existing lines seeding
// the de-duplication set (src/gpu_search_v42.cu:448-452).
#pragma once

#include <stdint.h>

#include <array>
#include <atomic>
#include <csignal>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <set>
#include <string>

#include "../../include/hss.h"

namespace cli {

inline std::atomic<bool> stop_requested{false};

inline void request_stop(int) { stop_requested.store(true); }

inline void install_signals() {
    std::signal(SIGTERM, request_stop);
    std::signal(SIGINT, request_stop);
}

inline std::string u128_decimal(unsigned __int128 v) {
    if (!v) return "0";
    std::string s;
    while (v) {
        s.push_back(char('0' + (int)(v % 10)));
        v /= 10;
    }
    return std::string(s.rbegin(), s.rend());
}

// area = |4A| / 4 of a hit
inline std::string hit_area(const hss_hit& h) {
    unsigned __int128 a4 = ((unsigned __int128)h.area4_hi << 64) | h.area4_lo;
    return u128_decimal(a4 / 4);
}

// integer-valued keys of a STATE file (missing file: empty map)
inline std::map<std::string, uint64_t> read_state(const std::string& path) {
    std::map<std::string, uint64_t> kv;
    std::ifstream in(path);
    std::string key, value;
    while (in >> key >> value) {
        try {
            kv[key] = std::stoull(value);
        } catch (...) {
        }
    }
    return kv;
}

inline void write_state(const std::string& path, const std::string& body) {
    std::filesystem::path tmp = path + ".tmp";
    {
        std::ofstream out(tmp);
        out << body;
    }
    std::filesystem::rename(tmp, path);
}

// (l, m, n) triples already present in a results file
inline std::set<std::array<uint64_t, 3>> read_seen(const std::string& path) {
    std::set<std::array<uint64_t, 3>> seen;
    std::ifstream in(path);
    uint64_t l, m, n;
    std::string area;
    while (in >> l >> m >> n >> area) seen.insert({l, m, n});
    return seen;
}

}  // namespace cli
