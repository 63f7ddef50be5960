/* CPU oracle for the GPU-HeiPa multilevel hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A serial C++ restatement (compiled with g++) of the reference's two-hop
 * matching, contraction, projection, Jet refinement with weak/strong
 * rebalancing, the edge-cut / J(C,D,Pi) evaluators, the distance matrix, the
 * Solver level loop and the HM recursion, in the deterministic tie-break mode
 * ("det") that the GPU library must reproduce bit for bit.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker.  Every function cites the reference file:line it follows.
 *
 * The initial partitioner is NOT restated here: the reference's modified METIS
 * cannot be built offline, so the oracle links the product's own host routine
 * (gpu-heipa_amd/csrc/host/initial_partition.cpp) to keep end-to-end runs
 * comparable.  That routine's parity is therefore "shared code", checked only
 * by invariants.
 */
#ifndef HEIPA_ORACLE_H
#define HEIPA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;
    uint32_t m;               /* directed CSR entries */
    const uint32_t *xadj;     /* n + 1 */
    const uint32_t *adjncy;   /* m */
    const int32_t *adjwgt;    /* m, or NULL for uniform edge weights */
    const int32_t *vwgt;      /* n, or NULL for uniform vertex weights */
    int64_t g_weight;         /* total vertex weight */
} orc_graph;

typedef struct {
    uint32_t ratio_idx;   /* index into the filter-ratio list */
    uint32_t iteration;   /* value of the iteration counter after ++ */
    uint32_t mode;        /* 0 = LP, 1 = weak rebalance, 2 = strong rebalance */
    uint32_t n_cand;      /* LP: candidates after the first filter; rebalance: candidates scored */
    uint32_t n_moves;
    uint32_t pad;
    int64_t cut;          /* working cut after the moves */
    int64_t maxw;         /* working max block weight after the moves */
    uint64_t moves_hash;  /* sum over moves of mix64((u << 32) | dest) */
    uint64_t gain_hash;   /* LP: sum over candidates of mix64(mix64(u) + gain1[u]); else 0 */
} orc_jet_trace;

typedef struct {
    uint32_t n, m;        /* level graph size */
    uint32_t nc;          /* coarse size produced from this level (0 for the coarsest) */
    uint32_t pad;
    uint64_t mapping_hash;/* sum over u of mix64((u << 32) | M(u)) */
    int64_t cut;          /* cut after refining this level (initial cut for the coarsest) */
    int64_t maxw;
} orc_level_trace;

typedef struct {
    uint32_t rounds;          /* heavy-edge matching rounds */
    uint32_t after_hem;       /* unmapped after heavy-edge matching */
    uint32_t after_leaf;      /* unmapped after the leaf pass (== after_hem if skipped) */
    uint32_t after_twin;
    uint32_t after_relative;
    uint32_t ran_leaf, ran_twin, ran_relative;
} orc_match_stats;

/* hashing (two_hop_matching.h:42, :93-100; Kokkos pod_hash = MurmurHash3_x86_32) */
/* Block connectivity rows (block_conn.h).  mode 1 (default): the reference's
 * rows, whose entries can keep weight 0 (they count in strong rebalance and are
 * mapping-LP candidates); mode 0: positive connections only.  With the check
 * on, every row is compared with the adjacency after each update; a mismatch
 * or a capacity overrun sets orc_row_error(). */
void orc_set_row_mode(int mode);
void orc_set_row_check(int on);
int orc_row_error(void);
uint32_t orc_murmur3_x86_32(const void *key, int32_t len, uint32_t seed);
uint32_t orc_pod_hash_u32(uint32_t x);
uint32_t orc_xorshift(uint32_t x);
uint64_t orc_mix64(uint64_t x);

/* jet_label_propagation.h:147-180 (det form: exact comparison with powers of 1.5) */
uint32_t orc_gain_bucket(int64_t gx, int64_t vwgt);

/* edge_cut.h:71-90 / comm_cost.h:36-56 with 64-bit accumulation */
int64_t orc_edge_cut(const orc_graph *g, const uint32_t *part);
int64_t orc_comm_cost(const orc_graph *g, const uint32_t *part, const int32_t *D, uint32_t k);
/* distance_oracle_matrix.h:41-90; D_out is k*k with k = prod(hier) */
int orc_distance_matrix(const uint32_t *hier, const int32_t *dist, uint32_t l, int32_t *D_out);

/* two_hop_matching.h:573-650 */
int orc_match(const orc_graph *g, uint32_t seed, uint32_t *mapping, uint32_t *nc, orc_match_stats *stats);
/* graph.h:127-390 with canonical (sorted) coarse rows; cadjncy/cadjwgt need room for g->m entries */
int orc_contract(const orc_graph *g, const uint32_t *mapping, uint32_t nc, uint32_t *cxadj,
                 uint32_t *cadjncy, int32_t *cadjwgt, int32_t *cvwgt, uint32_t *cm);
/* partition.h:92-105 */
void orc_project(const uint32_t *mapping, uint32_t fine_n, const uint32_t *coarse_part, uint32_t *fine_part);
/* jet_label_propagation.h:926-1033 */
int orc_jet_refine(const orc_graph *g, uint32_t *part, uint32_t k, int64_t lmax, int ultra, uint32_t level,
                   int64_t *cut, int64_t *maxw, orc_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len);
/* solver.h:340-556 (plain Solver); level trace optional */
int orc_partition(const orc_graph *g, uint32_t k, double imbalance, uint64_t seed, int ultra, uint32_t *part,
                  int64_t *cut, int64_t *maxw, orc_level_trace *levels, uint32_t level_cap, uint32_t *n_levels);
/* hierarchical_multisection.h:237-307 */
int orc_hm_partition(const orc_graph *g, const uint32_t *hier, uint32_t l, double imbalance, uint64_t seed,
                     int ultra, uint32_t *part);
/* ProMap_jet_refine (promap_jet_label_propagation.h:1219-1311); *J is the comm cost in/out,
 * trace records carry J in their cut field */
int orc_map_refine(const orc_graph *g, uint32_t *part, const int32_t *D, uint32_t k, int64_t lmax, int64_t *J,
                   int64_t *maxw, orc_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len);
/* global_multisection_host (global_multisection.h:217-259) over the shared host partitioner */
int orc_global_multisection(const orc_graph *g, const uint32_t *hier, uint32_t l, double imbalance, uint64_t seed,
                            uint32_t *part);
/* ProMapSolver::solve, --config IM (promap_solver.h:146-246); level trace cut fields hold J */
int orc_im_partition(const orc_graph *g, const uint32_t *hier, const int32_t *dist, uint32_t l, double imbalance,
                     uint64_t seed, uint32_t *part, int64_t *J, int64_t *maxw, orc_level_trace *levels,
                     uint32_t level_cap, uint32_t *n_levels);
/* global_multisection.h:39-48 */
double orc_adaptive_imbalance(double global_imbalance, int64_t global_g_weight, uint32_t global_k,
                              int64_t local_g_weight, uint32_t local_k_rem, uint64_t depth);

#ifdef __cplusplus
}
#endif

#endif
