// Host initial partitioner for the coarsest graph of the multilevel hierarchy.
//
// The reference runs a modified serial METIS on the host at this point
// (src/initial_partitioning/kway_partitioner/kway_core.h:1527-1588, called from
// src/datastructures/solver.h:477-498).  This is an independent design, not a
// restatement of that code: recursive bisection where every bisection is the
// best of several greedy-graph-growing tries each polished by a 2-way FM pass,
// followed by a greedy k-way cut/balance pass.  It only ever sees graphs of at
// most 8k vertices.  Every choice is driven by a splitmix64 stream seeded with
// `seed`, so the output is a pure function of (graph, k, imbalance, seed); the
// GPU driver and the CPU oracle call this same routine.
#include "host_graph.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <queue>
#include <system_error>
#include <thread>
#include <tuple>

namespace heipa {
namespace {

struct Sub {
    u32 n = 0;
    std::vector<u32> xadj, adj;
    std::vector<s64> ew, vw;
    std::vector<u32> gid;   // vertex id in the caller's graph
    s64 total = 0;
};

struct BisectResult {
    std::vector<u8> side;
    s64 cut = 0;
    s64 over = 0;
};

s64 overload(s64 w0, s64 w1, double max0, double max1) {
    s64 o = 0;
    if ((double) w0 > max0) o += (s64) std::ceil((double) w0 - max0);
    if ((double) w1 > max1) o += (s64) std::ceil((double) w1 - max1);
    return o;
}

bool better(s64 over_a, s64 cut_a, s64 over_b, s64 cut_b) {
    if (over_a != over_b) return over_a < over_b;
    return cut_a < cut_b;
}

s64 compute_cut(const Sub &g, const std::vector<u8> &side) {
    s64 c = 0;
    for (u32 u = 0; u < g.n; ++u)
        for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j)
            if (side[u] != side[g.adj[j]]) c += g.ew[j];
    return c / 2;
}

// Greedy graph growing: side 0 grows from `start` by max-gain frontier
// vertices until it reaches its target weight.
void grow(const Sub &g, u32 start, double tw0, double max0, const std::vector<u64> &prio,
          std::vector<u8> &side) {
    side.assign(g.n, 1);
    std::vector<s64> gain(g.n, 0);
    for (u32 u = 0; u < g.n; ++u) {
        s64 d = 0;
        for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j) d += g.ew[j];
        gain[u] = -d;
    }
    using Item = std::tuple<s64, u64, u32>;
    std::priority_queue<Item> heap;
    s64 w0 = 0;
    u32 scan = 0;
    auto take = [&](u32 v) {
        side[v] = 0;
        w0 += g.vw[v];
        for (u32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
            u32 x = g.adj[j];
            if (side[x] == 1) {
                gain[x] += 2 * g.ew[j];
                heap.emplace(gain[x], prio[x], x);
            }
        }
    };
    take(start);
    std::vector<u8> rejected(g.n, 0);
    while ((double) w0 < tw0) {
        u32 v = g.n;
        while (!heap.empty()) {
            auto [gv, pv, x] = heap.top();
            heap.pop();
            if (side[x] == 1 && !rejected[x] && gv == gain[x]) {
                v = x;
                break;
            }
        }
        if (v == g.n) {
            // frontier exhausted: jump to the next untouched vertex
            while (scan < g.n && (side[scan] == 0 || rejected[scan])) ++scan;
            if (scan == g.n) break;
            v = scan;
        }
        if ((double) (w0 + g.vw[v]) > max0) {
            rejected[v] = 1;
            continue;
        }
        take(v);
    }
}

// 2-way FM with rollback to the best prefix of each pass.
void fm2(const Sub &g, double max0, double max1, std::vector<u8> &side, int max_passes) {
    std::vector<s64> gain(g.n);
    std::vector<u8> moved(g.n);
    for (int pass = 0; pass < max_passes; ++pass) {
        s64 w[2] = {0, 0};
        for (u32 u = 0; u < g.n; ++u) w[side[u]] += g.vw[u];
        s64 cut = 0;
        for (u32 u = 0; u < g.n; ++u) {
            s64 gi = 0;
            for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j) {
                if (side[g.adj[j]] != side[u]) {
                    gi += g.ew[j];
                    cut += g.ew[j];
                } else {
                    gi -= g.ew[j];
                }
            }
            gain[u] = gi;
        }
        cut /= 2;
        using Item = std::pair<s64, u32>;
        auto cmp = [](const Item &a, const Item &b) {
            if (a.first != b.first) return a.first < b.first;
            return a.second > b.second;
        };
        std::priority_queue<Item, std::vector<Item>, decltype(cmp)> heap[2] = {
            std::priority_queue<Item, std::vector<Item>, decltype(cmp)>(cmp),
            std::priority_queue<Item, std::vector<Item>, decltype(cmp)>(cmp)};
        std::fill(moved.begin(), moved.end(), 0);
        for (u32 u = 0; u < g.n; ++u) heap[side[u]].emplace(gain[u], u);
        const double maxs[2] = {max0, max1};
        s64 best_over = overload(w[0], w[1], max0, max1);
        s64 best_cut = cut;
        s64 start_over = best_over, start_cut = best_cut;
        std::vector<u32> seq;
        size_t best_len = 0;
        const size_t limit = std::max<size_t>(25, g.n / 20);
        while (true) {
            // candidate from each side: best non-moved vertex whose move is allowed
            int pick = -1;
            u32 pv = 0;
            s64 pg = 0;
            for (int s = 0; s < 2; ++s) {
                auto &h = heap[s];
                while (!h.empty()) {
                    auto [gv, v] = h.top();
                    if (moved[v] || side[v] != s || gv != gain[v]) {
                        h.pop();
                        continue;
                    }
                    break;
                }
                if (h.empty()) continue;
                auto [gv, v] = h.top();
                const int t = 1 - s;
                bool allowed = (double) (w[t] + g.vw[v]) <= maxs[t] || (double) w[s] > maxs[s];
                if (!allowed) continue;
                if (pick < 0 || gv > pg || (gv == pg && (double) w[s] - maxs[s] > (double) w[1 - pick] - maxs[1 - pick])) {
                    pick = s;
                    pv = v;
                    pg = gv;
                }
            }
            if (pick < 0) break;
            heap[pick].pop();
            const u32 v = pv;
            moved[v] = 1;
            side[v] = (u8) (1 - pick);
            w[pick] -= g.vw[v];
            w[1 - pick] += g.vw[v];
            cut -= gain[v];
            gain[v] = -gain[v];
            for (u32 j = g.xadj[v]; j < g.xadj[v + 1]; ++j) {
                u32 x = g.adj[j];
                if (moved[x]) continue;
                // x's gain changes by +-2w depending on whether x is now on v's side
                if (side[x] == side[v]) gain[x] -= 2 * g.ew[j];
                else gain[x] += 2 * g.ew[j];
                heap[side[x]].emplace(gain[x], x);
            }
            seq.push_back(v);
            s64 ov = overload(w[0], w[1], max0, max1);
            if (better(ov, cut, best_over, best_cut)) {
                best_over = ov;
                best_cut = cut;
                best_len = seq.size();
            } else if (seq.size() - best_len > limit) {
                break;
            }
        }
        for (size_t i = seq.size(); i > best_len; --i) {
            u32 v = seq[i - 1];
            side[v] = (u8) (1 - side[v]);
        }
        if (!better(best_over, best_cut, start_over, start_cut)) break;
    }
}

// Eight tries (graph growing from a random start, then FM), the best kept.
// Every try draws n + 1 numbers from the splitmix64 stream, whose state
// advances by a fixed constant per draw, so try t starts at
// rng + t (n + 1) * constant and the tries of a large subgraph run on host
// threads with the same result as one after the other.
BisectResult bisect(const Sub &g, double tw0, double eps, u64 &rng) {
    constexpr u64 STEP = 0x9e3779b97f4a7c15ull;  // splitmix64's state increment
    const double tw1 = (double) g.total - tw0;
    const double max0 = tw0 * (1.0 + eps), max1 = tw1 * (1.0 + eps);
    const int tries = g.n < 8 ? 1 : 8;
    const u64 draws = (u64) g.n + 1;
    const u64 rng0 = rng;
    rng += (u64) tries * draws * STEP;
    std::vector<BisectResult> res(tries);
    auto one = [&](int t) {
        u64 st = rng0 + (u64) t * draws * STEP;
        std::vector<u64> prio(g.n);
        for (u32 u = 0; u < g.n; ++u) prio[u] = splitmix64(st);
        const u32 start = (u32) (splitmix64(st) % g.n);
        std::vector<u8> side;
        grow(g, start, tw0, max0, prio, side);
        fm2(g, max0, max1, side, 8);
        s64 w0 = 0;
        for (u32 u = 0; u < g.n; ++u) if (side[u] == 0) w0 += g.vw[u];
        res[t].cut = compute_cut(g, side);
        res[t].over = overload(w0, g.total - w0, max0, max1);
        res[t].side = std::move(side);
    };
    std::vector<std::exception_ptr> errors(tries);
    auto guarded = [&](int t) {
        try {
            one(t);
        } catch (...) {
            errors[t] = std::current_exception();
        }
    };
    std::vector<std::thread> workers;
    for (int t = 1; t < tries; ++t) {
        if (g.n >= 512) {
            try {
                workers.emplace_back(guarded, t);
                continue;
            } catch (const std::system_error &) {
            }
        }
        guarded(t);  // small subgraph, or no thread available: this try runs here
    }
    guarded(0);
    for (auto &w: workers) w.join();
    for (auto &e: errors)
        if (e) std::rethrow_exception(e);
    int best = 0;
    for (int t = 1; t < tries; ++t)
        if (better(res[t].over, res[t].cut, res[best].over, res[best].cut)) best = t;
    return std::move(res[best]);
}

Sub extract(const Sub &g, const std::vector<u8> &side, u8 s) {
    Sub c;
    std::vector<u32> loc(g.n, UINT32_MAX);
    for (u32 u = 0; u < g.n; ++u) {
        if (side[u] == s) {
            loc[u] = c.n++;
            c.gid.push_back(g.gid[u]);
            c.vw.push_back(g.vw[u]);
            c.total += g.vw[u];
        }
    }
    c.xadj.assign(c.n + 1, 0);
    for (u32 u = 0; u < g.n; ++u) {
        if (side[u] != s) continue;
        for (u32 j = g.xadj[u]; j < g.xadj[u + 1]; ++j) {
            u32 v = g.adj[j];
            if (side[v] == s) {
                c.adj.push_back(loc[v]);
                c.ew.push_back(g.ew[j]);
            }
        }
        c.xadj[loc[u] + 1] = (u32) c.adj.size();
    }
    return c;
}

void recurse(const Sub &g, u32 k, u32 offset, double eps, u64 &rng, u32 *part) {
    if (g.n == 0) return;
    if (k == 1 || g.n == 1) {
        for (u32 u = 0; u < g.n; ++u) part[g.gid[u]] = offset;
        return;
    }
    const u32 k0 = k / 2, k1 = k - k0;
    const double tw0 = (double) g.total * (double) k0 / (double) k;
    BisectResult r = bisect(g, tw0, eps, rng);
    Sub a = extract(g, r.side, 0);
    Sub b = extract(g, r.side, 1);
    recurse(a, k0, offset, eps, rng, part);
    recurse(b, k1, offset + k0, eps, rng, part);
}

// Greedy k-way pass: positive-gain moves that keep the target block within
// lmax, and forced moves out of overloaded blocks.
void kway_greedy(u32 n, const u32 *xadj, const u32 *adjncy, const s32 *adjwgt, const s32 *vwgt,
                 u32 k, s64 lmax, u32 *part) {
    std::vector<s64> bw(k, 0), conn(k, 0);
    std::vector<u32> touched;
    for (u32 u = 0; u < n; ++u) bw[part[u]] += vwgt ? vwgt[u] : 1;
    for (int pass = 0; pass < 8; ++pass) {
        u32 moved = 0;
        for (u32 u = 0; u < n; ++u) {
            const s64 wu = vwgt ? vwgt[u] : 1;
            const u32 own = part[u];
            touched.clear();
            for (u32 j = xadj[u]; j < xadj[u + 1]; ++j) {
                u32 b = part[adjncy[j]];
                if (conn[b] == 0) touched.push_back(b);
                conn[b] += adjwgt ? adjwgt[j] : 1;
            }
            const s64 own_c = conn[own];
            const bool over = bw[own] > lmax;
            s64 best_g = 0;
            u32 best = UINT32_MAX;
            for (u32 b: touched) {
                if (b == own || bw[b] + wu > lmax) continue;
                s64 g = conn[b] - own_c;
                bool ok = over ? true : g > 0;
                if (!ok) continue;
                if (best == UINT32_MAX || g > best_g || (g == best_g && b < best)) {
                    best = b;
                    best_g = g;
                }
            }
            if (over && best == UINT32_MAX) {
                // no adjacent block has room: lightest block with room
                for (u32 b = 0; b < k; ++b) {
                    if (b == own || bw[b] + wu > lmax) continue;
                    if (best == UINT32_MAX || bw[b] < bw[best]) best = b;
                }
            }
            for (u32 b: touched) conn[b] = 0;
            if (best != UINT32_MAX) {
                bw[own] -= wu;
                bw[best] += wu;
                part[u] = best;
                ++moved;
            }
        }
        if (moved == 0) break;
    }
}

} // namespace

void initial_partition(u32 n, const u32 *xadj, const u32 *adjncy, const s32 *adjwgt,
                       const s32 *vwgt, u32 k, double imbalance, u64 seed, u32 *part) {
    for (u32 u = 0; u < n; ++u) part[u] = 0;
    if (k <= 1 || n == 0) return;
    Sub g;
    g.n = n;
    g.xadj.assign(xadj, xadj + n + 1);
    g.adj.assign(adjncy, adjncy + xadj[n]);
    g.ew.resize(xadj[n]);
    for (u32 j = 0; j < xadj[n]; ++j) g.ew[j] = adjwgt ? adjwgt[j] : 1;
    g.vw.resize(n);
    g.gid.resize(n);
    for (u32 u = 0; u < n; ++u) {
        g.vw[u] = vwgt ? vwgt[u] : 1;
        g.gid[u] = u;
        g.total += g.vw[u];
    }
    // per-bisection tolerance such that the product over log2(k) levels stays
    // within (1 + imbalance)
    const double depth = std::ceil(std::log2((double) k));
    const double eps_b = std::pow(1.0 + imbalance, 1.0 / std::max(1.0, depth)) - 1.0;
    u64 rng = seed ^ 0x6a09e667f3bcc909ull;
    recurse(g, k, 0, eps_b, rng, part);
    const s64 lmax = (s64) std::ceil((1.0 + imbalance) * ((double) g.total / (double) k));
    kway_greedy(n, xadj, adjncy, adjwgt, vwgt, k, lmax, part);
}

} // namespace heipa
