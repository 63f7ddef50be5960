// Internal C++ interfaces of the device hot path (one header for all .hip files).
#pragma once

#include "graph_meta.hpp"

#include <vector>

namespace heipa {

// ---- input validation (eval.hip); throw Error(HEIPA_ERR_INVALID)
void validate_graph(Ctx &c, const DGraph &g);
void validate_part(Ctx &c, const u32 *part, u32 n, u32 k);

// ---- evaluators (eval.hip)
s64 edge_cut(Ctx &c, const DGraph &g, const GraphMeta &meta, const u32 *part);
s64 comm_cost(Ctx &c, const DGraph &g, const GraphMeta &meta, const u32 *part, const s32 *D, u32 k);
void block_weights_dev(Ctx &c, const DGraph &g, const u32 *part, u32 k, s64 *bw_dev);

// ---- matching (match.hip): mapping is a device array of n entries; returns nc
u32 match(Ctx &c, const DGraph &g, const GraphMeta &meta, u32 seed, u32 *mapping, heipa_match_stats *stats);

// ---- contraction (contract.hip): the coarse arrays are pushed on the arena's
// front stack in the order vwgt, xadj, adjncy, adjwgt.
struct CoarseGraph {
    DGraph g;
    size_t front_depth_before = 0;  // release_front(front_depth_before) frees it
};
CoarseGraph contract(Ctx &c, const DGraph &fine, const GraphMeta &meta, const u32 *mapping, u32 nc);
void project(Ctx &c, const u32 *mapping, u32 fine_n, const u32 *coarse_part, u32 *fine_part);

// ---- Jet refinement (jet.hip)
// Largest k one Jet call takes: the sweep keeps one LDS counter per block for
// each lane group (12k bytes against a 48 KiB budget).
constexpr u32 JET_MAX_K = 4093;
struct JetResult {
    s64 cut = 0, maxw = 0;
    u32 iterations = 0;
};
// D (device, k x k, row-major) selects the mapping objective: `cut` is then
// the comm cost J and the loop is ProMap_jet_refine's
// (src/refinement/promap_jet_label_propagation.h:1219-1311).
JetResult jet_refine(Ctx &c, const DGraph &g, const GraphMeta &meta, u32 *part, u32 k, s64 lmax, bool ultra,
                     u32 level, s64 cut, s64 maxw, std::vector<heipa_jet_trace> *trace, const s32 *D = nullptr);

// ---- drivers (solver.hip)
struct SolveResult {
    s64 cut = 0, maxw = 0, lmax = 0, initial_cut = 0;
    u32 levels = 0, jet_iterations = 0;
    double ms[5] = {0, 0, 0, 0, 0};
};
// The IM process-mapping objective: hierarchy and the device distance matrix
// (k x k with k = prod(hier)).
struct MapSpec {
    std::vector<u32> hier;
    const s32 *D = nullptr;
};
// map != null: ProMapSolver::solve (IM); cut / initial_cut / level cuts hold J
SolveResult solve(Ctx &c, const DGraph &g, u32 k, double imbalance, u64 seed, bool ultra, u32 *part,
                  std::vector<heipa_level_trace> *levels, const MapSpec *map = nullptr);
void hm_partition(Ctx &c, const DGraph &g, const std::vector<u32> &hier, double imbalance, u64 seed, bool ultra,
                  u32 *part);

} // namespace heipa
