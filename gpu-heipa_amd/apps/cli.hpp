// Command-line handling shared by GPU-HeiPa and GPU-HeiProMap.
//
// Same option table semantics as the reference (src/utility/configuration.h:
// 40-205, src/utility/promap_configuration.h:37-210): long and short keys,
// an option with an empty default is required, --help prints the table.
// Added: --seed (the reference draws std::random_device, configuration.h:119)
// and --arena-gb (device arena size; the reference hard-codes 14 GiB,
// configuration.h:79).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "heipa.h"

namespace cli {

struct Option {
    std::string large_key, small_key, description, default_val, input;
    bool is_set = false;
};

class Options {
public:
    explicit Options(std::vector<Option> o) : opts_(std::move(o)) {}

    void parse(int argc, char **argv) {
        std::vector<std::string> args(argv, argv + argc);
        for (size_t i = 1; i < args.size(); ++i)
            if (args[i] == "--help") {
                print_help();
                std::exit(EXIT_SUCCESS);
            }
        for (size_t i = 1; i < args.size(); ++i) {
            bool known = false;
            for (auto &o: opts_) {
                if (o.large_key == args[i] || (!o.small_key.empty() && o.small_key == args[i])) {
                    if (i + 1 >= args.size()) {
                        std::cerr << "Command Line \"" << args[i] << "\" needs a value!" << std::endl;
                        std::exit(EXIT_FAILURE);
                    }
                    o.input = args[i + 1];
                    o.is_set = true;
                    ++i;
                    known = true;
                    break;
                }
            }
            if (!known) {
                std::cerr << "Unknown option \"" << args[i] << "\"" << std::endl;
                std::exit(EXIT_FAILURE);
            }
        }
    }
    bool is_set(const std::string &k) const { return find(k).is_set; }
    std::string get(const std::string &k) const {
        const Option &o = find(k);
        if (o.input.empty() && o.default_val.empty()) {
            std::cout << "Command Line \"" << k << "\" not set!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        return o.input.empty() ? o.default_val : o.input;
    }
    void print_help() const {
        for (const auto &o: opts_) {
            if (o.small_key.empty()) std::cout << "[ " << o.large_key << "] - " << o.description << std::endl;
            else std::cout << "[ " << o.large_key << ", " << o.small_key << "] - " << o.description << std::endl;
        }
    }

private:
    const Option &find(const std::string &k) const {
        for (const auto &o: opts_)
            if (o.large_key == k || o.small_key == k) return o;
        std::cout << "Command Line \"" << k << "\" is not an allowed name!" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::vector<Option> opts_;
};

template<typename T>
std::vector<T> split_ints(const std::string &s, char sep) {
    std::vector<T> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, sep)) out.push_back((T) std::stoll(item));
    return out;
}

inline double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

inline void die(const char *what, int code, heipa_ctx *ctx = nullptr) {
    std::cerr << "Error: " << what << " failed with status " << code;
    if (ctx) std::cerr << " (" << heipa_ctx_last_error(ctx) << ")";
    std::cerr << std::endl;
    std::exit(EXIT_FAILURE);
}

#define CLI_HIP(x)                                                                  \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::cerr << #x << ": " << hipGetErrorString(e_) << std::endl;          \
            std::exit(EXIT_FAILURE);                                                \
        }                                                                           \
    } while (0)

// device copy of a host graph
struct DeviceGraph {
    heipa_graph g{};
    void *bufs[4] = {nullptr, nullptr, nullptr, nullptr};

    explicit DeviceGraph(const heipa_host_graph *h) {
        uint32_t n, m, flags;
        int64_t gw;
        const uint32_t *xadj, *adj;
        const int32_t *ew, *vw;
        heipa_host_graph_view(h, &n, &m, &gw, &flags, &xadj, &adj, &ew, &vw);
        g.n = n;
        g.m = m;
        g.g_weight = gw;
        g.flags = flags;
        CLI_HIP(hipMalloc(&bufs[0], 4ull * (n + 1)));
        CLI_HIP(hipMalloc(&bufs[1], 4ull * (m ? m : 1)));
        CLI_HIP(hipMemcpy(bufs[0], xadj, 4ull * (n + 1), hipMemcpyHostToDevice));
        if (m) CLI_HIP(hipMemcpy(bufs[1], adj, 4ull * m, hipMemcpyHostToDevice));
        if (ew) {
            CLI_HIP(hipMalloc(&bufs[2], 4ull * (m ? m : 1)));
            if (m) CLI_HIP(hipMemcpy(bufs[2], ew, 4ull * m, hipMemcpyHostToDevice));
        }
        if (vw) {
            CLI_HIP(hipMalloc(&bufs[3], 4ull * (n ? n : 1)));
            if (n) CLI_HIP(hipMemcpy(bufs[3], vw, 4ull * n, hipMemcpyHostToDevice));
        }
        g.xadj = (const uint32_t *) bufs[0];
        g.adjncy = (const uint32_t *) bufs[1];
        g.adjwgt = (const int32_t *) bufs[2];
        g.vwgt = (const int32_t *) bufs[3];
    }
    ~DeviceGraph() {
        for (void *p: bufs)
            if (p) (void) hipFree(p);
    }
};

} // namespace cli
