/* heipa.h — C-ABI of the MI355X-native GPU-HeiPa / GPU-HeiProMap hot path.
 *
 * Plain pointers and sizes only.  Graph and partition arrays passed to the
 * device entry points are DEVICE pointers (hipMalloc'd or from any allocator
 * on the context's device); results come back as host scalars or in device
 * arrays the caller provides.  Every call is synchronous at return (the
 * reference fences at the end of every phase, solver.h:440/467/519/548) and
 * returns a status code instead of calling exit()/abort().
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef HEIPA_H
#define HEIPA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference aborts or exits instead:
 * kokkos_memory_stack.h:83-94, host_graph.h:148-151, configuration.h:141-164) */
#define HEIPA_OK 0
#define HEIPA_ERR_OOM 1       /* context arena exhausted */
#define HEIPA_ERR_INVALID 2   /* bad argument or malformed graph/file */
#define HEIPA_ERR_OVERFLOW 3  /* a coarse weight does not fit the s32 graph format */
#define HEIPA_ERR_DEVICE 4    /* HIP runtime error */
#define HEIPA_ERR_IO 5        /* file error */

/* heipa_graph.flags (the reference's uniform_vertex_weights / uniform_edge_weights,
 * src/datastructures/graph.h:51-52: uniform weights are not stored) */
#define HEIPA_UNIFORM_VW 1u
#define HEIPA_UNIFORM_EW 2u

typedef struct heipa_ctx heipa_ctx;

/* Device CSR graph; replaces `Graph` (src/datastructures/graph.h:47-61).
 * Both directions stored, m = directed entries, u32 ids and offsets, s32
 * weights.  The reference's redundant COO array edges_u is not kept. */
typedef struct {
    uint32_t n;
    uint32_t m;
    int64_t g_weight;          /* total vertex weight */
    uint32_t flags;            /* HEIPA_UNIFORM_VW | HEIPA_UNIFORM_EW */
    uint32_t reserved;
    const uint32_t *xadj;      /* device, n + 1 */
    const uint32_t *adjncy;    /* device, m */
    const int32_t *adjwgt;     /* device, m, NULL iff HEIPA_UNIFORM_EW */
    const int32_t *vwgt;       /* device, n, NULL iff HEIPA_UNIFORM_VW */
} heipa_graph;

typedef struct {
    uint32_t rounds;
    uint32_t after_hem, after_leaf, after_twin, after_relative;
    uint32_t ran_leaf, ran_twin, ran_relative;
} heipa_match_stats;

/* one record per performed Jet iteration (see heipa_jet_refine_traced) */
typedef struct {
    uint32_t ratio_idx, iteration, mode, n_cand, n_moves, pad;
    int64_t cut, maxw;
    uint64_t moves_hash, gain_hash;
} heipa_jet_trace;

/* one record per level of a Solver run, finest first */
typedef struct {
    uint32_t n, m, nc, pad;
    uint64_t mapping_hash;
    int64_t cut, maxw;
} heipa_level_trace;

typedef struct {
    int64_t cut;               /* final edge cut (64-bit) */
    int64_t maxw;              /* final max block weight */
    int64_t lmax;              /* ceil((1+eps) W / k) */
    int64_t initial_cut;       /* cut of the initial partition on the coarsest graph */
    uint32_t levels;           /* number of graphs in the hierarchy */
    uint32_t jet_iterations;   /* Jet iterations over all levels */
    double ms_coarsening, ms_contraction, ms_initial, ms_uncontraction, ms_refinement;
} heipa_solve_stats;

/* ---- context: device, stream, LIFO arena (replaces KokkosMemoryStack,
 * src/datastructures/kokkos_memory_stack.h:36-187, and the default execution
 * space).  One context per host thread. */
int heipa_ctx_create(int device, uint64_t arena_bytes, heipa_ctx **out);
int heipa_ctx_destroy(heipa_ctx *ctx);
const char *heipa_ctx_last_error(const heipa_ctx *ctx);
void *heipa_ctx_stream(heipa_ctx *ctx);          /* the hipStream_t all work runs on */
uint64_t heipa_arena_mark(const heipa_ctx *ctx); /* front-stack depth, for heipa_arena_release */
int heipa_arena_release(heipa_ctx *ctx, uint64_t mark);
uint64_t heipa_arena_peak(const heipa_ctx *ctx); /* peak bytes in use since creation or the last reset */
int heipa_arena_reset_peak(heipa_ctx *ctx);      /* the peak restarts from the bytes in use now */
/* copies between device and host memory on the context stream (synchronous);
 * lets callers without their own HIP runtime handle read results */
int heipa_download(heipa_ctx *ctx, void *host_dst, const void *dev_src, uint64_t bytes);
int heipa_upload(heipa_ctx *ctx, void *dev_dst, const void *host_src, uint64_t bytes);

/* ---- evaluators */
/* edge_cut<uew>(const Graph&, const Partition&, exec) — src/utility/edge_cut.h:71-90 (64-bit sum, halved) */
int heipa_edge_cut(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, int64_t *cut);
/* comm_cost<uew,O>(const Graph&, const Partition&, O&, exec) — src/utility/comm_cost.h:36-56
 * (64-bit, not halved); D is a device k*k s32 matrix */
int heipa_comm_cost(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, const int32_t *D, uint32_t k,
                    int64_t *J);
/* No counterpart in the reference: heipa_edge_cut for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It takes
 * no context and allocates nothing: its partial sums live in a static device array of 32 KiB
 * (2048 workgroup slots) that each device holds until the process ends; every input size fits it
 * (the launch uses at most 2048 workgroups).  Asynchronous on `stream` (a hipStream_t of the
 * current device, NULL for its default stream): *cut_dev (device int64) receives the cut, or -1
 * when the graph is malformed (where heipa_edge_cut returns HEIPA_ERR_INVALID).  Returns
 * HEIPA_ERR_INVALID for null arguments, HEIPA_ERR_DEVICE if a launch fails.  Calls that overlap on
 * one device share the array and race: order them on one stream. */
int heipa_edge_cut_noalloc(const heipa_graph *g, const uint32_t *part, int64_t *cut_dev, void *stream);
/* initialize_distance_oracle_matrix(k, hierarchy, distance, ...) —
 * src/distance_oracles/distance_oracle_matrix.h:41-90; host output, k = prod(hier) */
int heipa_distance_matrix(const uint32_t *hier, const int32_t *dist, uint32_t l, int32_t *D_host);
/* recalculate_weights + max_weight — src/datastructures/partition.h:107-134; bw_host has k entries (any k:
 * up to 8192 blocks are summed in LDS, more with global atomics) */
int heipa_block_weights(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, uint32_t k, int64_t *bw_host,
                        int64_t *maxw);

/* ---- coarsening */
/* two_hop_matcher_get_mapping<uvw,uew>(const Graph&, const Partition&, lmax, stack, exec) —
 * src/coarsening/two_hop_matching.h:573-650.  mapping: device, n entries; seed 12345 in the Solver. */
int heipa_match(heipa_ctx *ctx, const heipa_graph *g, uint32_t seed, uint32_t *mapping, uint32_t *nc,
                heipa_match_stats *stats);
/* from_Graph_Mapping<uvw,uew>(const Graph&, const Mapping&, stack, exec) — src/datastructures/graph.h:127-390.
 * The coarse arrays are allocated on the context arena's front stack (LIFO, like the reference); free
 * them with heipa_arena_release(ctx, mark taken before the call).  Rows are sorted by neighbour id. */
int heipa_contract(heipa_ctx *ctx, const heipa_graph *fine, const uint32_t *mapping, uint32_t nc,
                   heipa_graph *coarse);
/* uncontract(Partition&, const Mapping&, exec) — src/datastructures/partition.h:92-105 */
int heipa_project(heipa_ctx *ctx, const uint32_t *mapping, uint32_t fine_n, const uint32_t *coarse_part,
                  uint32_t *fine_part);

/* ---- refinement */
/* jet_refine<uvw,uew>(Graph&, Partition&, k, lmax, use_ultra, level, cut, maxw, stack, exec) —
 * src/refinement/jet_label_propagation.h:926-1033.  part is in/out (device); cut and maxw are in/out
 * (the current values on entry, the best partition's on return).  bweights (device, k entries, may be
 * NULL) receives the block weights of the returned partition, as the reference's partition.bweights
 * (copy_into, src/datastructures/partition.h:192-200).  k <= 4093 (the block-connectivity counters live
 * in LDS; the reference has no such limit): a larger k returns HEIPA_ERR_INVALID.  The same limit holds
 * for heipa_map_refine, for heipa_partition and heipa_im_partition on graphs with more than 8k vertices,
 * and for each split (one hierarchy entry) of heipa_hm_partition. */
int heipa_jet_refine(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, uint32_t k,
                     int64_t lmax, int ultra, uint32_t level, int64_t *cut, int64_t *maxw);
/* same, recording one heipa_jet_trace per iteration (test hook; adds a readback per iteration) */
int heipa_jet_refine_traced(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, uint32_t k,
                            int64_t lmax, int ultra, uint32_t level, int64_t *cut, int64_t *maxw,
                            heipa_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len);

/* ---- IM process mapping (mapping objective J = sum over directed edges w(u,v) * D(part(u), part(v))) */
/* ProMap_jet_refine<uvw,uew> — src/refinement/promap_jet_label_propagation.h:1219-1311.  D: device, k x k
 * row-major distances (heipa_distance_matrix, uploaded).  part in/out (device); J and maxw in/out (current
 * values on entry, the best partition's on return); bweights as for heipa_jet_refine.  level is accepted
 * for parity with the reference's signature; the mapping LP reads no filter ratio. */
int heipa_map_refine(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, const int32_t *D,
                     uint32_t k, int64_t lmax, uint32_t level, int64_t *J, int64_t *maxw);
/* same, recording one heipa_jet_trace per iteration that moved vertices (cut fields hold J) */
int heipa_map_refine_traced(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights,
                            const int32_t *D, uint32_t k, int64_t lmax, uint32_t level, int64_t *J, int64_t *maxw,
                            heipa_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len);
/* ProMapSolver<DistanceOracleMatrix>(config).solve(host_g) with --config IM — src/datastructures/
 * promap_solver.h:146-347: coarsening, global multisection, J-driven Jet per level.  part: device.
 * stats->cut / initial_cut and the level trace cut fields hold J.  prod(hier) <= 65536, and <= 4093 when
 * n > 8 prod(hier) (the mapping refinement runs then); checked before the k x k matrix is built. */
int heipa_im_partition(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *hier, const int32_t *dist, uint32_t l,
                       double imbalance, uint64_t seed, uint32_t *part, heipa_solve_stats *stats,
                       heipa_level_trace *levels, uint32_t level_cap);

/* ---- drivers */
/* Solver(config).solve(host_g) / the device-graph Solver ctor — src/datastructures/solver.h:144-246, 248,
 * 340-556.  part: device, n entries.  stats and levels may be NULL. */
int heipa_partition(heipa_ctx *ctx, const heipa_graph *g, uint32_t k, double imbalance, uint64_t seed, int ultra,
                    uint32_t *part, heipa_solve_stats *stats, heipa_level_trace *levels, uint32_t level_cap);
/* ProMapSolver<DistanceOracleMatrix>(config).solve_multisection(host_g) —
 * src/datastructures/promap_solver.h:238-241 -> src/initial_partitioning/hierarchical_multisection.h:237-307.
 * imbalance >= 0; prod(hier) <= 65536.  The subtrees below the first split run on worker contexts
 * (HEIPA_OPT_HM_WORKERS); a worker whose arena runs out hands its unfinished subtrees back to ctx. */
int heipa_hm_partition(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *hier, uint32_t l, double imbalance,
                       uint64_t seed, int ultra, uint32_t *part);
/* determine_adaptive_imbalance — src/initial_partitioning/global_multisection.h:39-48 */
double heipa_adaptive_imbalance(double global_imbalance, int64_t global_g_weight, uint32_t global_k,
                                int64_t local_g_weight, uint32_t local_k_rem, uint64_t depth);

/* ---- batched and served forms.  No counterpart in the reference: its apps solve one graph per process.
 * A pool of `slots` workers on one device; each worker owns a context (HIP stream, arena of
 * arena_bytes, pinned staging) and a host thread.  Jobs wait in one FIFO queue and start when a worker
 * is free and the device-memory estimates of the running jobs plus the next one stay within
 * memory_budget.  A job runs exactly heipa_partition or heipa_hm_partition on its worker's context, so
 * its result does not depend on the other jobs; jobs on different workers overlap on the device (their
 * streams map to separate hardware queues up to the runtime's GPU_MAX_HW_QUEUES). */
typedef struct heipa_pool heipa_pool;
#define HEIPA_JOB_PARTITION 0 /* Solver(config).solve: k blocks */
#define HEIPA_JOB_HM 1        /* ProMapSolver::solve_multisection: hierarchy hier[0..l) */
typedef struct {
    /* inputs */
    heipa_graph graph;        /* device arrays; valid until the job is done */
    uint32_t kind;            /* HEIPA_JOB_PARTITION or HEIPA_JOB_HM */
    uint32_t k;               /* partition jobs */
    uint32_t hier[8];         /* HM jobs: a_1..a_l */
    uint32_t l;
    int32_t ultra;
    double imbalance;
    uint64_t seed;
    uint32_t *part;           /* device, graph.n entries */
    /* outputs (valid once done) */
    int32_t status;           /* HEIPA_OK or the error code of the run */
    int32_t slot;             /* worker that ran the job, -1 if it never ran */
    int64_t cut, maxw;        /* edge cut and largest block weight of the result */
    uint64_t estimate;        /* device bytes the job was admitted with */
    double t_submit, t_start, t_done; /* ms on the pool's clock (heipa_pool_clock) */
    uint32_t state;           /* 0 new, 1 queued, 2 running, 3 done (read through heipa_pool_wait) */
    uint32_t reserved;
    uint64_t peak;            /* device bytes the run used at most (its worker's arena peak) */
    uint64_t graph_key;       /* input: the graph's key for resident hierarchies (HEIPA_OPT_GRAPH_KEY), 0 none */
} heipa_job;
int heipa_pool_create(int device, uint32_t slots, uint64_t arena_bytes, uint64_t memory_budget, heipa_pool **out);
int heipa_pool_destroy(heipa_pool *pool);  /* waits for queued and running jobs */
/* device bytes a job needs in a worker's arena (an upper estimate from n, m and the number of blocks) */
uint64_t heipa_job_estimate(const heipa_job *job);
/* batched form: runs all jobs and returns when every one is done (HEIPA_OK even if some jobs failed;
 * see each job's status) */
int heipa_pool_run(heipa_pool *pool, heipa_job *jobs, uint32_t count);
/* served form: queues a job and returns at once; a job whose estimate exceeds the arena or the budget
 * is done at once with HEIPA_ERR_OOM */
int heipa_pool_submit(heipa_pool *pool, heipa_job *job);
/* waits until the job is done or timeout_ms passes (negative: no limit); 1 done, 0 not yet */
int heipa_pool_wait(heipa_pool *pool, heipa_job *job, double timeout_ms);
double heipa_pool_clock(const heipa_pool *pool);
/* heipa_ctx_set_option on every worker's context (call while no job runs); HEIPA_OPT_HIER_CACHE makes
 * one cache that all workers share */
int heipa_pool_set_option(heipa_pool *pool, int option, uint64_t value);
/* Autotuner of the batched form (no counterpart in the reference: its launch shapes are fixed).  Runs
 * the sample jobs through the pool once per candidate setting of three options that change launch
 * shapes but never results -- HEIPA_OPT_GROUP_CAP in {64, 16, 8}, HEIPA_OPT_BLOCK_GROUP_CAP in
 * {16, 8, 4}, HEIPA_OPT_BLOCK_JET in {1024, 2048, 4096} -- one option at a time starting from the
 * current setting (each value other than the option's starting one runs once: 2 + 6 runs from the
 * defaults), takes a candidate when the sample runs at least 2% faster with it than the best so far,
 * and leaves the chosen setting on every worker (call while no job runs).  The sample's outputs are those of its
 * last run; a failed sample job ends the tuning with its status. */
typedef struct {
    uint32_t group_cap, block_group_cap, block_jet;  /* the chosen setting */
    uint32_t runs;                                   /* sample runs, the untimed first one included */
    double ms_start, ms_best;                        /* sample wall time at the starting and the chosen setting */
} heipa_tune;
int heipa_pool_autotune(heipa_pool *pool, heipa_job *sample, uint32_t count, heipa_tune *out);
/* heipa_hier_cache_stats of the pool's shared cache */
int heipa_pool_hier_cache_stats(const heipa_pool *pool, uint64_t *entries, uint64_t *bytes, uint64_t *levels_reused,
                                uint64_t *levels_built);

/* ---- context options.  No counterpart in the reference (its sizes are compile-time).
 * HEIPA_OPT_RESOLVE_CAP: most rebalance resolve entries one workgroup sorts in LDS before an
 * iteration takes the device-wide radix-sort path (default and maximum: what the device's LDS
 * allows; 0 forces the device-wide path).  Results do not depend on it; tests use it to cover
 * both paths. */
#define HEIPA_OPT_RESOLVE_CAP 1
/* HEIPA_OPT_BLOCK_JET: the largest level (vertices) whose Jet runs in one workgroup when it has no
 * hubs and at most 32 edges per vertex on average (default 2048; 0 forces the multi-launch form
 * everywhere).  Results do not depend on it. */
#define HEIPA_OPT_BLOCK_JET 2
/* HEIPA_OPT_GRAPHS: 1 (default) replays each level's batch of Jet iterations as a captured HIP graph;
 * 0 launches every kernel.  Results do not depend on it. */
#define HEIPA_OPT_GRAPHS 3
/* HEIPA_OPT_PROF_SAMPLE: while profiling, bracket about one launch in `value` of a kernel with HIP
 * events, chosen by a hash of the launch count (default 1 = every launch); heipa_prof_get extrapolates
 * the timed launches' mean to all launches. */
#define HEIPA_OPT_PROF_SAMPLE 4
/* HEIPA_OPT_BUCKET_COPIES: copies of the rebalance minibucket totals that the multi-launch Jet
 * spreads its atomic additions over (0 = automatic, default: about one per 64 sweep workgroups, at
 * most 32; otherwise rounded down to a power of two, at most 32).  Results do not depend on it. */
#define HEIPA_OPT_BUCKET_COPIES 5
/* HEIPA_OPT_BLOCK_LDS: 1 (default) keeps the per-vertex Jet state of a single-workgroup level in LDS
 * for the whole launch when it fits (copied in at the start, back at the end); 0 keeps it in device
 * memory.  Results do not depend on it. */
#define HEIPA_OPT_BLOCK_LDS 6
/* HEIPA_OPT_HM_WORKERS: worker contexts (each its own stream, arena and host thread, kept by the
 * context between calls) on which heipa_hm_partition solves the subtrees below its first split at
 * the same time (default 8, at most 64; 0 or 1: one after another on the context's stream).  Setting
 * it closes the cached workers beyond the new count, which frees their arenas.  Results do not
 * depend on it. */
#define HEIPA_OPT_HM_WORKERS 7
/* HEIPA_OPT_HM_WORKER_BYTES: arena bytes of each HM worker context (default 0: an estimate from the
 * largest child's n and m).  A worker whose arena runs out hands its unfinished subtrees back to the
 * calling context, so a small value still gives the same result.  Setting it closes the cached
 * workers. */
#define HEIPA_OPT_HM_WORKER_BYTES 8
/* HEIPA_OPT_RACE: 1 resolves the reference's races by atomic order, as its GPU build does (a
 * matching receiver goes to the first proposer, two_hop_matching.h:321-329; a rebalancing
 * candidate's offset in its minibucket and a strong overflow position are atomic fetch-add results,
 * jet_label_propagation.h:439, :684, :580), instead of the deterministic rules (default 0).  Results
 * then vary from run to run; every partition still meets lmax wherever the deterministic run does. */
#define HEIPA_OPT_RACE 9
/* HEIPA_OPT_HIER_CACHE: keep coarse hierarchies in device memory between calls, up to `value` bytes
 * (least recently used graphs go first; 0 drops the cache).  The hierarchy of a graph does not depend
 * on k (two_hop_matching.h:590 fixes the matching seed; solver.h:343-347 only stops at n <= 8k), so a
 * later heipa_partition / heipa_im_partition / heipa_hm_partition (its first split) of the same graph
 * reuses the levels it needs, builds and keeps the ones beyond them, and returns the partition a fresh
 * solve returns.  A graph is named by HEIPA_OPT_GRAPH_KEY. */
#define HEIPA_OPT_HIER_CACHE 10
/* HEIPA_OPT_GRAPH_KEY: the caller's key of the input graph of the following calls on this context
 * (0, the default: no reuse).  A key that comes back with another n, m or total weight replaces its
 * entry; the caller keeps one key per graph. */
#define HEIPA_OPT_GRAPH_KEY 11
/* HEIPA_OPT_GROUP_CAP: the widest lane group (lanes per vertex) of the multi-launch Jet's afterburner,
 * apply and row launches (default 64, i.e. the power of two at or above the level's average degree,
 * at most 64; a smaller value rounds down to a power of two, at least 4).  Results do not depend on it.
 * HEIPA_OPT_BLOCK_GROUP_CAP: the same for the single-workgroup Jet (default 8: more vertices in flight
 * in the one workgroup).  heipa_pool_autotune picks both for a pool. */
#define HEIPA_OPT_GROUP_CAP 12
#define HEIPA_OPT_BLOCK_GROUP_CAP 13
/* HEIPA_OPT_REGION_EDGES: the multi-launch Jet gives each sweep workgroup a region of the level's
 * vertices (its candidate and move lists); on a dense level the regions shrink by halves from the
 * workgroup's thread count (to at least 8 vertices) while a region's vertices times the level's
 * average degree exceed `value`, so that more workgroups share the level (default 1024; 0 keeps one
 * vertex per thread).  The row updates of wide vertices and the matching's repick passes size their
 * pieces the same way (the repick only while that adds workgroups).  Results do not depend on it. */
#define HEIPA_OPT_REGION_EDGES 14
int heipa_ctx_set_option(heipa_ctx *ctx, int option, uint64_t value);
/* dst uses src's hierarchy cache (HEIPA_OPT_HIER_CACHE) from now on: one cache for several contexts */
int heipa_ctx_share_hier_cache(heipa_ctx *dst, const heipa_ctx *src);
/* graphs held, device bytes, levels reused and levels built since the cache was created */
int heipa_hier_cache_stats(const heipa_ctx *ctx, uint64_t *entries, uint64_t *bytes, uint64_t *levels_reused,
                           uint64_t *levels_built);

/* ---- per-kernel timing (HIP events on the context stream) */
int heipa_prof_enable(heipa_ctx *ctx, int on);
/* time only kernels whose name starts with `prefix` (NULL or "" = all) */
int heipa_prof_filter(heipa_ctx *ctx, const char *prefix);
int heipa_prof_reset(heipa_ctx *ctx);
uint32_t heipa_prof_count(heipa_ctx *ctx);
/* algorithmic bytes = the compulsory HBM bytes of each launch, summed; total_ms = the mean of the
 * event-timed launches times the number of launches (see HEIPA_OPT_PROF_SAMPLE) */
int heipa_prof_get(heipa_ctx *ctx, uint32_t idx, char *name, uint32_t name_cap, uint64_t *launches,
                   double *total_ms, double *alg_bytes);

/* ---- host-side graph tools (no device use) */
typedef struct heipa_host_graph heipa_host_graph;
/* from_file — src/datastructures/host_graph.h:154-247 */
int heipa_read_metis(const char *path, heipa_host_graph **out);
int heipa_write_metis(const char *path, const heipa_host_graph *g);
/* write_partition — src/utility/util.h:288-329 */
int heipa_write_partition(const char *path, const uint32_t *part_host, uint32_t n);
int heipa_gen_grid(uint32_t rows, uint32_t cols, heipa_host_graph **out);
int heipa_gen_rgg(uint32_t n, double avg_deg, uint64_t seed, uint32_t wmod, heipa_host_graph **out);
int heipa_gen_rmat(uint32_t scale, uint32_t n, double edge_factor, double a, double b, double c, uint64_t seed,
                   uint32_t wmod, int explicit_vw, heipa_host_graph **out);
int heipa_host_graph_from_arrays(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                                 const int32_t *vwgt, heipa_host_graph **out);
/* host views into the graph (valid until heipa_host_graph_free); adjwgt / vwgt NULL when uniform */
int heipa_host_graph_view(const heipa_host_graph *g, uint32_t *n, uint32_t *m, int64_t *g_weight, uint32_t *flags,
                          const uint32_t **xadj, const uint32_t **adjncy, const int32_t **adjwgt,
                          const int32_t **vwgt);
void heipa_host_graph_free(heipa_host_graph *g);
/* the shared host initial partitioner (replaces kway_partition, kway_core.h:1548-1588); host arrays */
int heipa_initial_partition(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                            const int32_t *vwgt, uint32_t k, double imbalance, uint64_t seed, uint32_t *part);
/* global_multisection_host (src/initial_partitioning/global_multisection.h:217-259) over the shared host
 * partitioner: the IM path's initial mapping; host arrays, part receives ids < prod(hier) */
int heipa_global_multisection(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                              const int32_t *vwgt, const uint32_t *hier, uint32_t l, double imbalance, uint64_t seed,
                              uint32_t *part);

#ifdef __cplusplus
}
#endif

#endif /* HEIPA_H */
