// Global multisection: the initial mapping of the IM process-mapping path
// (src/initial_partitioning/global_multisection.h:39-287), on the host as in
// the reference (the coarsest graph has at most 8k vertices).
//
// The graph is split into hierarchy[l-1] blocks by the host partitioner,
// each block becomes a subgraph (its vertices in id order with their weights,
// its internal edges in adjacency order) that is split into hierarchy[l-2]
// blocks, and so on; a leaf's block b is written as
// offset(identifier path) + b.  Each split runs under the adaptive imbalance
// that leaves the remaining levels the same slack.  Empty blocks are skipped:
// they have no vertex to place.  The subtrees of the first split run on
// host threads of their own: they share nothing but disjoint entries of the
// output, and every split is deterministic, so the result does not depend on
// the order in which they finish.
#include "host_graph.hpp"

#include <cmath>
#include <exception>
#include <stdexcept>
#include <system_error>
#include <thread>

namespace heipa {

double adaptive_imbalance(double eps, s64 gW, u32 gk, s64 lW, u32 krem, u64 depth) {
    // global_multisection.h:39-48
    const double li = (1.0 + eps) * ((double) ((u64) krem * (u64) gW) / (double) ((u64) gk * (u64) lW));
    return std::pow(li, 1.0 / (double) depth) - 1.0;
}

namespace {

struct Sub {
    u32 n = 0;
    s64 gw = 0;
    std::vector<u32> xadj, adj, n_to_o;
    std::vector<s32> ew, vw;
};

struct Plan {
    std::vector<u32> hier, index_vec, k_rem;
    double eps;
    s64 gW;
    u32 gk;
    u64 seed;
    u32 *out;
};

// recursive_partition (global_multisection.h:139-215) over one subgraph whose
// arrays are given as raw pointers (the top level may have uniform weights:
// null ew / vw)
void split(u32 n, const u32 *xadj, const u32 *adj, const s32 *ew, const s32 *vw, s64 gw, const u32 *n_to_o,
           u32 level, std::vector<u32> &ident, const Plan &p) {
    const u32 l = (u32) p.hier.size();
    const u32 k = p.hier[level];
    const double imb = adaptive_imbalance(p.eps, p.gW, p.gk, gw, p.k_rem[l - 1 - ident.size()], l - ident.size());
    std::vector<u32> local(n, 0);
    initial_partition(n, xadj, adj, ew, vw, k, imb, p.seed, local.data());
    if (ident.size() == l - 1) {
        u32 offset = 0;
        for (u32 i = 0; i + 1 < l; ++i) offset += ident[i] * p.index_vec[p.index_vec.size() - 1 - i];
        for (u32 u = 0; u < n; ++u) p.out[n_to_o[u]] = offset + local[u];
        return;
    }
    // build_subgraphs (global_multisection.h:50-137): one pass sizes every
    // block, a second fills them
    std::vector<Sub> subs(k);
    std::vector<u32> pos(n);
    for (u32 u = 0; u < n; ++u) {
        Sub &s = subs[local[u]];
        pos[u] = s.n++;
        s.gw += vw ? vw[u] : 1;
    }
    for (Sub &s: subs) {
        s.xadj.reserve(s.n + 1);
        s.xadj.push_back(0);
        s.vw.reserve(s.n);
        s.n_to_o.reserve(s.n);
    }
    for (u32 u = 0; u < n; ++u) {
        const u32 b = local[u];
        Sub &s = subs[b];
        for (u32 j = xadj[u]; j < xadj[u + 1]; ++j) {
            const u32 v = adj[j];
            if (local[v] != b) continue;
            s.adj.push_back(pos[v]);
            s.ew.push_back(ew ? ew[j] : 1);
        }
        s.xadj.push_back((u32) s.adj.size());
        s.vw.push_back(vw ? vw[u] : 1);
        s.n_to_o.push_back(n_to_o[u]);
    }
    local.clear();
    local.shrink_to_fit();
    auto child = [&](u32 b, std::vector<u32> &id) {
        Sub &s = subs[b];
        id.push_back(b);
        split(s.n, s.xadj.data(), s.adj.data(), s.ew.data(), s.vw.data(), s.gw, s.n_to_o.data(), level - 1, id, p);
        id.pop_back();
        s = Sub{};
    };
    if (!ident.empty() || k < 2) {
        for (u32 b = 0; b < k; ++b)
            if (subs[b].n) child(b, ident);
        return;
    }
    std::vector<std::thread> workers;
    std::vector<std::exception_ptr> errors(k);
    for (u32 b = 0; b < k; ++b) {
        if (subs[b].n == 0) continue;
        auto run = [&, b]() {
            try {
                std::vector<u32> id;
                id.reserve(p.hier.size());
                child(b, id);
            } catch (...) {
                errors[b] = std::current_exception();
            }
        };
        try {
            workers.emplace_back(run);
        } catch (const std::system_error &) {
            run();  // no thread available: this subtree runs here
        }
    }
    for (auto &t: workers) t.join();
    for (auto &e: errors)
        if (e) std::rethrow_exception(e);
}

} // namespace

void global_multisection(u32 n, const u32 *xadj, const u32 *adjncy, const s32 *adjwgt, const s32 *vwgt,
                         const u32 *hier, u32 l, double imbalance, u64 seed, u32 *part) {
    if (l == 0) throw std::invalid_argument("empty hierarchy");
    Plan p;
    p.hier.assign(hier, hier + l);
    p.eps = imbalance;
    p.seed = seed;
    p.out = part;
    p.gk = 1;
    for (u32 x: p.hier) {
        if (x == 0) throw std::invalid_argument("hierarchy entries must be positive");
        p.gk *= x;
    }
    p.index_vec = {1};  // global_multisection.h:228-230
    for (u32 i = 0; i + 1 < l; ++i) p.index_vec.push_back(p.index_vec[i] * p.hier[i]);
    p.k_rem.resize(l);  // :232-237
    u32 acc = 1;
    for (u32 i = 0; i < l; ++i) {
        p.k_rem[i] = acc * p.hier[i];
        acc *= p.hier[i];
    }
    s64 gw = 0;
    for (u32 u = 0; u < n; ++u) gw += vwgt ? vwgt[u] : 1;
    p.gW = gw;
    std::vector<u32> n_to_o(n);
    for (u32 u = 0; u < n; ++u) n_to_o[u] = u;
    std::vector<u32> ident;
    ident.reserve(l);
    split(n, xadj, adjncy, adjwgt, vwgt, gw, n_to_o.data(), l - 1, ident, p);
}

} // namespace heipa
