// Per-graph launch metadata: the lane-group width used by the group-per-vertex
// kernels (G lanes cooperate on one vertex's adjacency, G = next power of two
// of the average degree, 4..64) and the list of hub vertices whose degree is
// too large for a G-lane group; hubs get a whole workgroup each.
#pragma once

#include "prim.hpp"

namespace heipa {

struct GraphMeta {
    u32 G = 16;          // lanes per vertex
    u32 hub_deg = 1024;  // degree above which a vertex is a hub
    u32 max_deg = 0;
    u32 n_hubs = 0;
    u32 *hubs = nullptr; // device list (arena, owner's lifetime)
};

// G halved down to at most cap (never below 4)
inline u32 cap_group(u32 G, u32 cap) {
    while (G > 4 && G > cap) G >>= 1;
    return G;
}

inline u32 group_for_avg(double avg) {
    u32 g = 4;
    while (g < 64 && (double) g < avg) g <<= 1;
    return g;
}

__global__ void max_degree_kernel(const u32 *xadj, u32 n, u32 *out);
__global__ void hub_count_kernel(const u32 *xadj, u32 n, u32 hub_deg, u32 *count, u32 *list);

// Computes metadata; the hub list is pushed on the arena's BACK stack when
// `back` is true (temporary) or the FRONT stack otherwise.
GraphMeta make_meta(Ctx &c, const DGraph &g, bool back);

// Dispatch a templated kernel launcher on the group width.
#define HEIPA_DISPATCH_G(Gval, FN, ...)            \
    switch (Gval) {                                \
        case 4: FN<4>(__VA_ARGS__); break;         \
        case 8: FN<8>(__VA_ARGS__); break;         \
        case 16: FN<16>(__VA_ARGS__); break;       \
        case 32: FN<32>(__VA_ARGS__); break;       \
        default: FN<64>(__VA_ARGS__); break;       \
    }

} // namespace heipa
