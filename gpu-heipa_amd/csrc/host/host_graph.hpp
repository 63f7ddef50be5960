// Host-side graph container and helpers shared by the product library, the CLI
// apps and the oracle build.  Plain C++20, no HIP.
//
// Layout follows the reference's HostGraph (src/datastructures/host_graph.h:38-52):
// CSR with both directions stored, 0-based u32 ids, s32 weights, and uniform
// weights not stored (empty vectors + flags).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace heipa {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using s32 = int32_t;
using s64 = int64_t;

struct HostGraph {
    u32 n = 0;
    u32 m = 0;             // directed CSR entries (2 x undirected edges)
    s64 g_weight = 0;      // total vertex weight
    bool uniform_edge_weights = true;
    bool uniform_vertex_weights = true;
    std::vector<u32> xadj;     // n + 1
    std::vector<u32> adjncy;   // m
    std::vector<s32> adjwgt;   // m, empty if uniform
    std::vector<s32> vwgt;     // n, empty if uniform

    s32 ew(u32 j) const { return uniform_edge_weights ? 1 : adjwgt[j]; }
    s32 vw(u32 u) const { return uniform_vertex_weights ? 1 : vwgt[u]; }
};

// splitmix64 step; used as the seed expander of every generator and of the
// initial partitioner's random choices.
inline u64 splitmix64(u64 &state) {
    u64 z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Stateless 64-bit mixer (the splitmix64 finalizer).
inline u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded by splitmix64; uniform doubles are (x >> 11) * 2^-53.
struct Xoshiro256 {
    u64 s[4];
    explicit Xoshiro256(u64 seed) {
        u64 st = seed;
        for (auto &x: s) x = splitmix64(st);
    }
    static u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }
    u64 next() {
        const u64 result = rotl(s[1] * 5, 7) * 9;
        const u64 t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double) (next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, bound) by multiply-shift (bound > 0)
    u64 below(u64 bound) { return (u64) (((__uint128_t) next() * bound) >> 64); }
};

// ---- METIS I/O (reference: src/datastructures/host_graph.h:154-247, src/utility/util.h:288-329)
// Returns an empty string on success, else an error message.
std::string read_metis(const std::string &path, HostGraph &g);
std::string write_metis(const std::string &path, const HostGraph &g);
std::string write_partition(const std::string &path, const u32 *part, u32 n);
std::string read_partition(const std::string &path, std::vector<u32> &part);

// ---- Synthetic generators (seeded, deterministic)
// 2D 5-point grid rows x cols, unit weights, adjacency sorted ascending.
HostGraph gen_grid(u32 rows, u32 cols);
// 2D random geometric graph: n points in [0,1)^2, edge iff squared distance <= r^2
// with r = sqrt(avg_deg / (pi n)); edge weight = 1 + hash(seed,min,max) mod wmod
// (wmod == 0 -> uniform weights); adjacency sorted.
HostGraph gen_rgg(u32 n, double avg_deg, u64 seed, u32 wmod);
// R-MAT of the given scale with edge_factor * 2^scale samples, quadrant probs
// (a,b,c,d); self loops dropped, symmetrised, deduplicated; ids >= n rejected
// when n < 2^scale (n == 0 means 2^scale); weight = 1 + hash mod wmod;
// explicit_vw stores unit vertex weights explicitly.
HostGraph gen_rmat(u32 scale, u32 n, double edge_factor, double a, double b, double c,
                   u64 seed, u32 wmod, bool explicit_vw);

// edge weight hash used by the generators
inline s32 edge_weight_hash(u64 seed, u32 u, u32 v, u32 wmod) {
    u32 lo = u < v ? u : v, hi = u < v ? v : u;
    u64 h = mix64(mix64(seed ^ 0x5bd1e995ull) ^ (((u64) lo << 32) | hi));
    return (s32) (1 + h % wmod);
}

// ---- Initial partitioner (host, serial).  Replaces the reference's
// modified-METIS kway_partition (src/initial_partitioning/kway_partitioner/
// kway_core.h:1527-1588) with an own recursive bisection (greedy graph
// growing + FM) followed by a greedy k-way pass.  Deterministic for a given
// (graph, k, imbalance, seed).  adjwgt / vwgt may be null (uniform).
void initial_partition(u32 n, const u32 *xadj, const u32 *adjncy, const s32 *adjwgt,
                       const s32 *vwgt, u32 k, double imbalance, u64 seed, u32 *part);

// ---- Global multisection (host, serial): the IM path's initial mapping
// (src/initial_partitioning/global_multisection.h:139-287) over
// initial_partition; part receives ids < prod(hier).  Throws
// std::invalid_argument on an empty hierarchy or a zero entry.
void global_multisection(u32 n, const u32 *xadj, const u32 *adjncy, const s32 *adjwgt, const s32 *vwgt,
                         const u32 *hier, u32 l, double imbalance, u64 seed, u32 *part);
// determine_adaptive_imbalance (global_multisection.h:39-48)
double adaptive_imbalance(double eps, s64 gW, u32 gk, s64 lW, u32 krem, u64 depth);

} // namespace heipa
