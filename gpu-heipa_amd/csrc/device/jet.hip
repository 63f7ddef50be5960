// Jet refinement on the GPU, deterministic, with the refinement loop's control
// flow kept on the device.
//
// Follows jet_refine (src/refinement/jet_label_propagation.h:926-1033) with
// jet_lp (:182-354), rebalance_weak (:592-751), rebalance_strong (:356-590) and
// perform_moves (:753-924).  Differences in mechanism, not in result:
//   * no BlockConn table (block_conn.h): a vertex's block connectivity is
//     rebuilt on demand in LDS — G lanes per vertex add their edges into a
//     dense per-group k-counter array with 64-bit LDS atomics; the first adder
//     of a block appends it to a per-group touched list.  The reference's
//     dest_cache/gain_cache (jet:194-237) decide which vertices are rebuilt.
//     Ties between blocks go to the smallest block id.  What a BlockConn
//     row holds beyond the positive connections is tracked for the vertices
//     whose row capacity is k (degree >= k, "wide"): the row's size and the
//     set of its valid entries, which keeps a block whose connection fell to
//     0 while the size is k (block_conn.h:379) until update_large rebuilds
//     the row.  Strong rebalance counts such entries in its mean (jet:418-428)
//     and the mapping LP takes them as candidates (promap:237-246).
//   * the cut change of a move set is computed exactly in 64 bits from the old
//     and new blocks of each moved vertex and its neighbours; it equals the
//     reference's (c1 + c2) / 2 (jet:763-923).
//   * minibucket offsets: a vertex's offset inside its minibucket is the weight
//     of smaller ids in it; strong-rebalance overflow (jet:576-583) is resolved
//     serially in id order.  Only candidates whose offset can decide their
//     selection are sorted: one workgroup sorts them by (minibucket, id) in
//     LDS; longer lists go through a device-wide radix sort.
//   * gain_bucket (jet:147-180): float quotient as in the reference, then an
//     exact comparison against float thresholds of powers of 1.5.
//   * control: the loop of jet_refine (ratio index, iteration counters,
//     weak/strong switching, best-partition tracking) runs in the last
//     workgroup of each iteration's apply launch.  Every kernel reads the
//     control block first and returns at once when the loop has ended, so the
//     host enqueues iterations in batches and reads the control block once per
//     batch instead of once per iteration.
// Launches per iteration: sweep (LP best block | rebalance score), buckets
// (rebalance only: per-block prefix of the minibucket totals), post
// (afterburner | rebalance classify), resolve (rebalance only), apply, and a
// one-workgroup control launch (block weights, counters, loop state).  No
// kernel synchronises across workgroups: every hand-off is a launch boundary.
#include "eval.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace heipa {
namespace {

constexpr u32 NO_MOVE = 0xFFFFFFFFu - 3;
constexpr u32 NULL_PART = 0xFFFFFFFFu - 1;
constexpr s64 GAIN_MIN = INT64_MIN;
constexpr u32 N_MAX_ITERATIONS = 12;
constexpr u32 N_MAX_WEAK_ITERATIONS = 2;
constexpr double PHI = 0.999;
constexpr u32 MAX_SECTIONS = 128;
constexpr u32 MAX_BUCKETS = 50;
constexpr int MID_BUCKETS = 25;
constexpr int JB = 256;                  // threads of the group-per-vertex kernels
constexpr int RB = 1024;                 // threads of the resolve workgroup
constexpr u32 RESOLVE_CAP = 8192;        // resolve entries sorted in LDS (power of two) ...
constexpr u32 RESOLVE_CAP_SMALL = 2048;  // ... or this many where a workgroup gets at most 64 KiB of LDS
constexpr u32 OVF_CAP = 512;             // strong-rebalance overflow entries kept in LDS
constexpr u32 ES_CAP = 1025;             // eviction chunk bounds kept in LDS (k + 1)
constexpr u32 TRACE_RING = 256;          // trace records kept on the device between host reads
constexpr u32 MAX_RATIOS = 17;
constexpr size_t SWEEP_LDS = 48 * 1024;  // block-connectivity counters per sweep workgroup
constexpr u32 SWEEP_PIECE = 1024;        // vertices listed in LDS per sweep phase
constexpr u32 BATCH_MAX = 8;             // iterations enqueued per host read of the control block

// float thresholds: LO[k] = least float >= 1.5^-k, HI[k] = greatest float <= 1.5^k
__constant__ float kLo[24] = {
    0x1.000000p+0f, 0x1.555556p-1f, 0x1.c71c72p-2f, 0x1.2f684cp-2f, 0x1.948b10p-3f, 0x1.0db20cp-3f,
    0x1.679810p-4f, 0x1.df756ap-5f, 0x1.3fa39cp-5f, 0x1.aa2f7ap-6f, 0x1.1c1fa6p-6f, 0x1.7ad4dep-7f,
    0x1.f91bd2p-8f, 0x1.50bd38p-8f, 0x1.c0fc4ap-9f, 0x1.2b52dcp-9f, 0x1.8f1926p-10f, 0x1.0a10c4p-10f,
    0x1.62c104p-11f, 0x1.d9015cp-12f, 0x1.3b563ep-12f, 0x1.a472fcp-13f, 0x1.184ca8p-13f, 0x1.75bb8ap-14f};
__constant__ float kHi[25] = {
    0x1.000000p+0f, 0x1.800000p+0f, 0x1.200000p+1f, 0x1.b00000p+1f, 0x1.440000p+2f, 0x1.e60000p+2f,
    0x1.6c8000p+3f, 0x1.116000p+4f, 0x1.9a1000p+4f, 0x1.338c00p+5f, 0x1.cd5200p+5f, 0x1.59fd80p+6f,
    0x1.037e20p+7f, 0x1.853d30p+7f, 0x1.23ede4p+8f, 0x1.b5e4d6p+8f, 0x1.486ba0p+9f, 0x1.eca170p+9f,
    0x1.717914p+10f, 0x1.151acep+11f, 0x1.9fa836p+11f, 0x1.37be28p+12f, 0x1.d39d3ep+12f, 0x1.5eb5eep+13f,
    0x1.070872p+14f};

__device__ u32 gain_bucket(s64 gx, s64 vwgt) {
    if (gx > 0) return 0;
    if (gx == 0) return 1;
    const float g = (float) (-gx) / (float) vwgt;
    if (g < 1.0f) {
        for (int k = 1; k <= 23; ++k)
            if (g >= kLo[k]) return (u32) (MID_BUCKETS - k);
        return 2;
    }
    for (int k = 0; k <= 24; ++k)
        if (g <= kHi[k]) return (u32) (MID_BUCKETS + k);
    return MAX_BUCKETS - 1;
}

__device__ __forceinline__ u64 mix64d(u64 z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Counters of the iteration in flight.  Lists have one region per sweep
// workgroup, filled without global atomics (an LDS counter per workgroup),
// plus an extra list for the few producers outside the sweep grid (hub
// workgroups, resolve, global path), appended with global atomics; st[]
// holds the extra lists' lengths and the hub workgroups' statistics, and the
// per-workgroup statistics go to a partial array.  The control launch sums
// both into the iteration's totals and clears st[].
enum Part : u32 { PT_MISS = 0, PT_MISS_DEG, PT_CAND_DEG, PT_GHASH, PT_DELTA, PT_MOVE_DEG, PT_MHASH, PT_COUNT = 8 };
enum Stat : u32 {
    ST_MISS = 0,  // vertices rebuilt (LP) / scored (rebalance)
    ST_CAND,      // LP candidates / rebalance candidates (extra list: length)
    ST_MOVES,     // moves (extra list: length)
    ST_RLIST,     // rebalance resolve list length
    ST_DELTA,     // s64 raw cut delta (twice the cut change)
    ST_MHASH,
    ST_GHASH,
    ST_MISS_DEG,  // sum of degrees of the rebuilt / scored vertices
    ST_CAND_DEG,  // sum of degrees of the afterburner candidates
    ST_MOVE_DEG,  // sum of degrees of the moved vertices
    ST_COUNT
};

// The state of jet_refine's loop (jet:960-1030), in device memory.
// The head of the control block: every launch reads it at its start, in
// one load issued before the live check (loads behind the check's early
// return would be a second dependent memory round trip).
struct CtlHead {
    double ratio;  // ratio of the current ratio index
    u32 active;    // 0 once the loop has ended
    u32 stalled;   // the resolve list did not fit: the host runs the global path
    u32 mode;
    u32 save;      // the partition after the last iteration is the best so far
    u32 stamp;
    u32 zero_lp;   // mapping objective: LP iterations in a row without moves
};
struct JetCtl {
    double ratio;
    u32 active, stalled, mode, save, stamp, zero_lp;
    unsigned long long st[ST_COUNT];
    s64 cur_cut, cur_maxw, best_cut, best_maxw;
    double bytes[5];  // algorithmic bytes of the sweep, buckets, post, resolve and apply launches
    u32 ri, it, bal;
    u32 iters;        // iterations that changed the partition
    u32 n_trace;
    u32 last_moves;   // length of the last iteration's move list (update_large when > n / 10)
};
static_assert(sizeof(CtlHead) == 32 && offsetof(JetCtl, st) == sizeof(CtlHead), "control block head layout");

struct JetDev {
    u32 n, k;
    s64 lmax, opt;  // opt = ceil(W / k)
    const u32 *xadj, *adj;
    const s32 *ew, *vw;
    u32 *work, *dest, *dcache, *mark, *oldp, *part;
    s64 *gcache, *gain1;
    s64 *icache;  // connection to the own block of a vertex whose neighbours all share its block; -1: unknown
    u8 *lock;
    u32 *cand, *moves;      // regions of cap / cap_mv entries, one per sweep workgroup
    u32 *cand_n, *moves_n;  // region lengths
    u32 *cand_x, *moves_x;  // extra lists (lengths in st[ST_CAND], st[ST_MOVES])
    unsigned long long *pstat;  // PT_COUNT partial statistics per sweep workgroup
    u32 nreg, cap, cap_mv;
    u32 count_bytes;  // hub kernels add their degree statistics (the byte accounting of the profiler)
    s64 *bw;
    JetCtl *ctl;
    heipa_jet_trace *trace;  // ring of TRACE_RING records; null: no trace, no hashes
    const double *ratios;  // MAX_RATIOS filter ratios
    u32 n_ratios, hub_deg;
    u32 rcap;  // resolve entries sorted in LDS (<= the kernel's capacity)
    // rebalancing
    u32 sections, width, t;
    u32 nshard;         // workgroup w adds into copy w % nshard (spreads same-address atomics)
    u32 *gid;           // minibucket per vertex
    s64 *btot;          // nshard copies of the t + 1 minibucket totals, zero between iterations
    s64 *Q, *R;         // t: exclusive / inclusive prefix of btot inside each block's range
    s64 *Tseg;          // k: total of each block's range
    u64 *rlist;         // resolve list: minibucket << 32 | vertex
    s64 *evict_adjust;  // k
    s64 *evict_start;   // k + 1
    s64 *counter;       // strong-rebalance overflow position
    const s32 *D;       // k x k distances: mapping objective J (ProMap); null: edge cut
    // BlockConn rows of the wide vertices (degree >= k); rflag null when the level has none
    u8 *rflag;          // n: bit 0 wide, bit 1 a neighbour is in this iteration's move list,
                        // bit 2 its neighbour counts are kept (wcnt)
    u32 *wid;           // n: rank among the wide vertices
    u32 *wsz;           // per wide vertex: the row's size (BlockConn sizes)
    u32 *wV, *wP;       // per wide vertex, W words each: blocks of its valid entries / of positive ones
    u32 *wcnt;          // per wide vertex, k counts of its neighbours per block, kept by the moves for
                        // the counted ones: k <= 64 and ROW_THREAD_DEG < degree <= hub_deg
    u32 W;              // words per block mask, ceil(k / 32)
    u32 race;           // atomic order: minibucket offsets and strong overflow positions from fetch-adds
};

// Per-vertex staleness of the caches, the n bytes after the locks: bit 0
// the decision cache (dcache, gcache), bit 1 the interior cache.  Moving a
// vertex marks it and its neighbours stale with one byte each instead of
// rewriting both caches (12 bytes): the bytes stay in L2 where the caches
// (12n) do not, and most moves of a rebalancing iteration touch vertices
// many times.
__device__ __forceinline__ u8 *stale_of(const JetDev &d) { return d.lock + d.n; }
__device__ __forceinline__ s64 vwt(const JetDev &d, u32 u) { return d.vw ? (s64) d.vw[u] : 1; }

// a candidate's weight into minibucket g: rebalancing puts most candidates of
// an overloaded block into a few dozen minibuckets, so each workgroup adds into
// its own copy of the totals (the bucket kernel sums the copies)
// Racing mode keeps one copy and records the add's previous total as the
// vertex's offset in its minibucket (the reference's temp_gain =
// atomic_fetch_add, jet:439, :684) in gain1, which no rebalancing iteration
// reads otherwise.
__device__ __forceinline__ void btot_add(const JetDev &d, u32 u, u32 g, s64 w) {
    const u64 copy = d.nshard > 1 ? blockIdx.x % d.nshard : 0;
    const unsigned long long old =
        atomicAdd((unsigned long long *) &d.btot[copy * ((u64) d.t + 1) + g], (unsigned long long) w);
    if (d.race) d.gain1[u] = (s64) old;
}
__device__ __forceinline__ s64 ewt(const JetDev &d, u32 j) { return d.ew ? (s64) d.ew[j] : 1; }
__device__ __forceinline__ u32 degree(const JetDev &d, u32 u) { return d.xadj[u + 1] - d.xadj[u]; }
__device__ __forceinline__ bool live(const JetCtl *ctl) { return ctl->active && !ctl->stalled; }
__device__ __forceinline__ CtlHead ctl_head(const JetCtl *ctl) {
    CtlHead h;
    __builtin_memcpy(&h, ctl, sizeof(h));
    return h;
}
__device__ __forceinline__ bool live(const CtlHead &h) { return h.active && !h.stalled; }

// eviction threshold of rebalance_strong (jet:363) / rebalance_weak (jet:600-604)
__device__ __forceinline__ s64 max_block_weight(const JetDev &d, bool strong) {
    const s64 a = (s64) ((double) d.lmax * 0.99);
    if (strong) return d.opt + 1 > a ? d.opt + 1 : a;
    return a < d.lmax - 100 ? d.lmax - 100 : a;
}

constexpr u32 ROW_THREAD_DEG = 32;  // wide vertices up to this degree walk their adjacency in one thread

// a neighbour v of a vertex moving from block ou to nu: a wide v's row is
// updated after the iteration (benign race on the flag: every writer stores
// the same value).  A counted v's neighbour counts follow the move, as the
// reference's update_small follows each moved vertex's edges
// (block_conn.h:208-345): its row update then reads k counts instead of its
// adjacency.  Vertices of low degree are cheaper to walk, and a hub's counts
// would serialise the moves' atomics on k addresses, so neither is counted.
// f = rflag[v] and wid = wid[v] (read only if v is counted), loaded by the
// caller together with the neighbour's other state: a load after the stores
// of the previous neighbour waited for them
__device__ __forceinline__ void row_touch(const JetDev &d, u32 v, u8 f, u32 wid, u32 ou, u32 nu) {
    if (f & 1u) {
        if (!(f & 2u)) d.rflag[v] = f | 2u;  // most neighbours of a move list are touched many times
        if ((f & 4u) && ou != nu) {
            const u64 base = (u64) wid * d.k;
            atomicSub(&d.wcnt[base + ou], 1u);
            atomicAdd(&d.wcnt[base + nu], 1u);
        }
    }
}

// the block mask of a counted vertex's positive entries from its counts
__device__ __forceinline__ u64 row_count_mask(const JetDev &d, u32 v) {
    const u64 base = (u64) d.wid[v] * d.k;
    u64 mask = 0;
    for (u32 b = 0; b < d.k; ++b)
        if (__hip_atomic_load(&d.wcnt[base + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) mask |= 1ull << b;
    return mask;
}

// the wide vertex u's row holds k entries: its valid entries include the
// blocks whose connection fell to 0 since the row filled up
__device__ __forceinline__ bool row_full(const JetDev &d, u32 u, u32 &r) {
    if (!d.rflag || !(d.rflag[u] & 1u)) return false;
    r = d.wid[u];
    return d.wsz[r] == d.k;
}

// orders the LDS traffic of one wave (lane groups never span waves)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// wave-aggregated appends to device lists
__device__ __forceinline__ void append(bool pred, u32 val, u32 *list, unsigned long long *counter) {
    const u64 bal = __ballot(pred);
    if (!bal) return;
    const u32 first = (u32) __builtin_ctzll(bal);
    u32 base = 0;
    if (lane_id() == first) base = (u32) atomicAdd(counter, (unsigned long long) __popcll(bal));
    base = __shfl(base, (int) first, WAVE);
    if (pred) list[base + (u32) __popcll(bal & lanemask_lt())] = val;
}
__device__ __forceinline__ void append64(bool pred, u64 val, u64 *list, unsigned long long *counter) {
    const u64 bal = __ballot(pred);
    if (!bal) return;
    const u32 first = (u32) __builtin_ctzll(bal);
    u32 base = 0;
    if (lane_id() == first) base = (u32) atomicAdd(counter, (unsigned long long) __popcll(bal));
    base = __shfl(base, (int) first, WAVE);
    if (pred) list[base + (u32) __popcll(bal & lanemask_lt())] = val;
}

// wave-aggregated append to this workgroup's region (LDS counter)
__device__ __forceinline__ void append_local(bool pred, u32 val, u32 *region, u32 *s_cnt) {
    const u64 bal = __ballot(pred);
    if (!bal) return;
    const u32 first = (u32) __builtin_ctzll(bal);
    u32 base = 0;
    if (lane_id() == first) base = atomicAdd(s_cnt, (u32) __popcll(bal));
    base = __shfl(base, (int) first, WAVE);
    if (pred) region[base + (u32) __popcll(bal & lanemask_lt())] = val;
}

// append_local of a vertex with two more words at the same index of two
// parallel arrays
__device__ __forceinline__ void append_local3(bool pred, u32 val, u32 a, u32 b, u32 *region, u32 *ra, u32 *rb,
                                              u32 *s_cnt) {
    const u64 bal = __ballot(pred);
    if (!bal) return;
    const u32 first = (u32) __builtin_ctzll(bal);
    u32 base = 0;
    if (lane_id() == first) base = atomicAdd(s_cnt, (u32) __popcll(bal));
    base = __shfl(base, (int) first, WAVE);
    if (pred) {
        const u32 i = base + (u32) __popcll(bal & lanemask_lt());
        region[i] = val;
        ra[i] = a;
        rb[i] = b;
    }
}

// sum over a workgroup of any multiple-of-64 size; the result is valid in
// thread 0; `lds` holds blockDim.x / 64 entries
__device__ __forceinline__ u64 block_total(u64 v, u64 *lds) {
    v = wave_sum(v);
    __syncthreads();
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = v;
    __syncthreads();
    u64 t = 0;
    if (threadIdx.x == 0)
        for (u32 w = 0; w < blockDim.x / WAVE; ++w) t += lds[w];
    return t;
}

__device__ __forceinline__ void wave_add(unsigned long long *ctr, u64 v) {
    v = wave_sum(v);
    if (lane_id() == 0 && v) atomicAdd(ctr, (unsigned long long) v);
}

// ---------------------------------------------------------------- LDS block connectivity
// per group: cnt[k] u64, touched[k] u32, tcount u32; every group region is
// padded to 16 bytes so the next group's u64 counters stay 8-byte aligned for
// the 64-bit LDS atomics (12k + 16 is only a multiple of 8 for even k)
__host__ __device__ inline size_t conn_group_bytes(u32 k) { return (12ull * k + 16 + 15) & ~(size_t) 15; }

struct GroupConn {
    unsigned long long *cnt;
    u32 *touched;
    u32 *tcount;
};

__device__ __forceinline__ GroupConn group_conn(char *smem, u32 k, u32 gidx) {
    char *p = smem + conn_group_bytes(k) * gidx;
    GroupConn g;
    g.cnt = (unsigned long long *) p;
    g.touched = (u32 *) (p + 8ull * k);
    g.tcount = (u32 *) (p + 12ull * k);
    return g;
}

// Adjacency loops take ADJ_UNROLL edges per lane per round and issue all
// loads of a round before using any, so a vertex of degree D costs about
// D / (stride * ADJ_UNROLL) rounds of dependent memory round trips instead of
// D / stride (a hub's workgroup walks thousands of entries per thread group).
constexpr u32 ADJ_UNROLL = 4;
// unroll of a lane group of G lanes: narrow groups walk short adjacencies,
// where the unrolled form only adds instructions and registers
template<int G>
constexpr u32 group_unroll() { return G >= 32 ? ADJ_UNROLL : 1; }

template<u32 ADJ_UNROLL>
__device__ __forceinline__ void conn_accumulate_range(const JetDev &d, const GroupConn &gc, u64 b0, u64 e, u32 lane,
                                                      u32 G) {
    for (u64 j0 = b0 + lane; j0 < e; j0 += (u64) G * ADJ_UNROLL) {
        u32 v[ADJ_UNROLL], bl[ADJ_UNROLL];
        s64 w[ADJ_UNROLL];
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            const u64 j = j0 + (u64) q * G;
            v[q] = j < e ? d.adj[j] : 0u;
            w[q] = j < e ? ewt(d, (u32) j) : 0;
        }
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) bl[q] = j0 + (u64) q * G < e ? d.work[v[q]] : 0u;
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            if (j0 + (u64) q * G < e) {
                const unsigned long long old = atomicAdd(&gc.cnt[bl[q]], (unsigned long long) w[q]);
                if (old == 0) gc.touched[atomicAdd(gc.tcount, 1u)] = bl[q];
            }
        }
    }
}

template<u32 ADJ_UNROLL>
__device__ __forceinline__ void conn_accumulate(const JetDev &d, const GroupConn &gc, u32 u, u32 lane, u32 G) {
    conn_accumulate_range<ADJ_UNROLL>(d, gc, d.xadj[u], d.xadj[u + 1], lane, G);
}

__device__ __forceinline__ void conn_reset(const GroupConn &gc, u32 T, u32 lane, u32 G) {
    for (u32 i = lane; i < T; i += G) gc.cnt[gc.touched[i]] = 0;
}

// lexicographic best (conn desc, block asc)
struct Best {
    s64 c;
    u32 b;
};
__device__ __forceinline__ bool best_better(const Best &x, const Best &y) {
    if (x.b == NO_MOVE) return false;
    if (y.b == NO_MOVE) return true;
    if (x.c != y.c) return x.c > y.c;
    return x.b < y.b;
}

// reduce over the G lanes of a group (G <= 64) or, for G > 64, over the block
template<int G>
__device__ __forceinline__ Best best_reduce(Best x, Best *red) {
    if constexpr (G <= WAVE) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            Best y{__shfl_xor(x.c, o, G), __shfl_xor(x.b, o, G)};
            if (best_better(y, x)) x = y;
        }
        return x;
    } else {
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            Best y{__shfl_xor(x.c, o, WAVE), __shfl_xor(x.b, o, WAVE)};
            if (best_better(y, x)) x = y;
        }
        if (lane_id() == 0) red[threadIdx.x / WAVE] = x;
        __syncthreads();
        Best r = red[0];
        for (u32 w = 1; w < G / WAVE; ++w)
            if (best_better(red[w], r)) r = red[w];
        __syncthreads();
        return r;
    }
}

template<int G, typename T>
__device__ __forceinline__ T group_sum(T x, T *red) {
    if constexpr (G <= WAVE) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, G);
        return x;
    } else {
        x = wave_sum(x);
        if (lane_id() == 0) red[threadIdx.x / WAVE] = x;
        __syncthreads();
        T r = 0;
        for (u32 w = 0; w < G / WAVE; ++w) r += red[w];
        __syncthreads();
        return r;
    }
}

// ---------------------------------------------------------------- LP evaluation
// best_block (:200-236) + filter (:247-258) of one rebuilt vertex; lane 0 of
// the group writes the results; returns whether u is a candidate
template<int G>
__device__ __forceinline__ bool lp_eval(const JetDev &d, const GroupConn &gc, u32 u, u32 own_b, bool locked, u32 lane,
                                        double ratio, Best *red, s64 *red2, bool active) {
    const u32 T = active ? *gc.tcount : 0;
    Best best{0, NO_MOVE};
    if (d.D) {
        // ProMap_jet_lp best_block (promap_jet_label_propagation.h:230-258):
        // gain_delta (:180-204) of a move to b is J(own) - J(b) with
        // J(x) = sum over the touched blocks c of conn(c) * D(x, c); the lanes
        // split the candidate blocks, each summing over all touched blocks
        const s32 *Do = d.D + (u64) own_b * d.k;
        s64 jo = 0;
        for (u32 i = lane; i < T; i += G) {
            const u32 c = gc.touched[i];
            jo += (s64) gc.cnt[c] * (s64) Do[c];
        }
        jo = group_sum<G>(jo, red2);
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (b == own_b) continue;
            const s32 *Db = d.D + (u64) b * d.k;
            s64 jb = 0;
            for (u32 j = 0; j < T; ++j) {
                const u32 c = gc.touched[j];
                jb += (s64) gc.cnt[c] * (s64) Db[c];
            }
            Best x{jo - jb, b};
            if (best_better(x, best)) best = x;
        }
        // blocks of weight-0 entries of a full row are candidates as well
        // (promap:237-246 walks every valid entry)
        u32 r = 0;
        if (active && row_full(d, u, r)) {
            for (u32 i = lane; i < d.W; i += G) {
                u32 z = d.wV[(u64) r * d.W + i] & ~d.wP[(u64) r * d.W + i];
                if (own_b >> 5 == i) z &= ~(1u << (own_b & 31));
                while (z) {
                    const u32 b = i * 32 + (u32) __builtin_ctz(z);
                    z &= z - 1;
                    const s32 *Db = d.D + (u64) b * d.k;
                    s64 jb = 0;
                    for (u32 j = 0; j < T; ++j) {
                        const u32 c = gc.touched[j];
                        jb += (s64) gc.cnt[c] * (s64) Db[c];
                    }
                    Best x{jo - jb, b};
                    if (best_better(x, best)) best = x;
                }
            }
        }
    } else {
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (b == own_b) continue;
            Best x{(s64) gc.cnt[b], b};
            if (best_better(x, best)) best = x;
        }
    }
    best = best_reduce<G>(best, red);
    bool cand = false;
    if (active && lane == 0) {
        const s64 own = (s64) gc.cnt[own_b];
        d.icache[u] = (T == 0 || (T == 1 && gc.touched[0] == own_b)) ? own : -1;
        u32 best_id = best.b;
        s64 gain = 0;
        if (best_id != NO_MOVE) {
            const s64 bc = best.c;
            if (d.D) {
                if (bc >= 0) gain = bc;  // promap:232: best_delta starts at -1
                else best_id = NO_MOVE;
            } else if (bc >= own || ((double) own - (double) bc) < floor(ratio * (double) own)) {
                gain = bc - own;
            } else {
                best_id = NO_MOVE;
            }
        }
        d.gcache[u] = gain;
        d.dcache[u] = best_id;
        stale_of(d)[u] = 0;  // both caches rewritten
        d.dest[u] = best_id;
        if (best_id != NO_MOVE && !locked) {
            d.gain1[u] = gain;
            cand = true;
        } else {
            d.gain1[u] = GAIN_MIN;
            d.lock[u] = 0;
        }
    }
    return cand;
}

// ---------------------------------------------------------------- rebalance evaluation
// per-workgroup view of the block weights: the underloaded blocks in ascending
// order (weak, jet:614-632) or the largest remaining capacity below the
// threshold (strong, jet:374-388; -1 if none)
struct RebShared {
    s64 max_vwgt;
    u32 n_under, pad;
    u32 under[1];  // k entries
};

// computed by the first wave, 64 blocks per step
__device__ __forceinline__ void reb_prologue(const JetDev &d, bool strong, s64 mbw, RebShared *rs) {
    if (threadIdx.x < WAVE) {
        const u32 lane = threadIdx.x;
        if (!strong) {
            u32 c = 0;
            for (u32 b0 = 0; b0 < d.k; b0 += WAVE) {
                const u32 b = b0 + lane;
                const bool f = b < d.k && d.bw[b] < mbw;
                const u64 bal = __ballot(f);
                if (f) rs->under[c + (u32) __popcll(bal & lanemask_lt())] = b;
                c += (u32) __popcll(bal);
            }
            if (lane == 0) rs->n_under = c;
        } else {
            s64 mx = -1;
            for (u32 b = lane; b < d.k; b += WAVE)
                if (d.bw[b] < mbw && mbw - d.bw[b] > mx) mx = mbw - d.bw[b];
            mx = wave_max(mx);
            if (lane == 0) rs->max_vwgt = mx;
        }
    }
    __syncthreads();
}

// vertices considered for eviction (weak jet:641-646, strong jet:394-398)
__device__ __forceinline__ bool reb_eligible(const JetDev &d, bool strong, const RebShared *rs, u32 u, u32 ub) {
    const s64 bwu = d.bw[ub];
    const s64 wu = vwt(d, u);
    if (!strong) return bwu > d.lmax && (double) wu < 1.5 * (double) (bwu - d.opt);
    const s64 mv = rs->max_vwgt;
    return mv >= 0 && bwu > d.lmax && wu <= 2 * mv && wu < 2 * (bwu - d.opt);
}

// weak best block (jet:658-694) / strong score (jet:399-435) of one vertex;
// lane 0 records dest (weak), the minibucket and its weight; returns whether
// u is a candidate
template<int G>
__device__ __forceinline__ bool reb_eval(const JetDev &d, bool strong, s64 mbw, const RebShared *rs,
                                         const GroupConn &gc, u32 u, u32 own_b, u32 lane, Best *red, s64 *red2,
                                         bool active) {
    const u32 T = active ? *gc.tcount : 0;
    bool listed = false;
    if (!strong) {
        Best best{0, NO_MOVE};
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (d.bw[b] >= mbw) continue;
            Best x{(s64) gc.cnt[b], b};
            if (best_better(x, best)) best = x;
        }
        best = best_reduce<G>(best, red);
        if (active && lane == 0) {
            const s64 own = (s64) gc.cnt[own_b];
            d.icache[u] = (T == 0 || (T == 1 && gc.touched[0] == own_b)) ? own : -1;
            stale_of(d)[u] = stale_of(d)[u] & 1u;  // the interior cache rewritten
            u32 dst;
            s64 gain;
            if (best.b != NO_MOVE && best.c > 0) {
                dst = best.b;
                gain = best.c - own;
            } else {
                const u32 nu = rs->n_under;
                dst = nu ? rs->under[u % nu] : own_b;
                gain = -own;
            }
            d.dest[u] = dst;
            if (dst != own_b) {
                const u32 bt = gain_bucket(gain, vwt(d, u));
                const u32 g = (MAX_BUCKETS * own_b + bt) * d.sections + (u % d.sections);
                d.gid[u] = g;
                btot_add(d, u, g, vwt(d, u));
                listed = true;
            }
        }
    } else {
        s64 sum = 0, count = 0;
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (b == own_b) continue;
            if (d.bw[b] < mbw) {
                sum += (s64) gc.cnt[b];
                count += 1;
            }
        }
        // entries of weight 0 in a full row count too (jet:418-428)
        u32 r = 0;
        if (active && row_full(d, u, r)) {
            for (u32 i = lane; i < d.W; i += G) {
                u32 z = d.wV[(u64) r * d.W + i] & ~d.wP[(u64) r * d.W + i];
                if (own_b >> 5 == i) z &= ~(1u << (own_b & 31));
                while (z) {
                    const u32 b = i * 32 + (u32) __builtin_ctz(z);
                    z &= z - 1;
                    count += d.bw[b] < mbw;
                }
            }
        }
        sum = group_sum<G>(sum, red2);
        count = group_sum<G>(count, red2);
        if (active && lane == 0) {
            const s64 own = (s64) gc.cnt[own_b];
            d.icache[u] = (T == 0 || (T == 1 && gc.touched[0] == own_b)) ? own : -1;
            stale_of(d)[u] = stale_of(d)[u] & 1u;  // the interior cache rewritten
            if (count == 0) count = 1;
            const s64 gain = sum / count - own;
            const s64 wu = vwt(d, u);
            const s64 lim = d.bw[own_b] - d.lmax;
            const u32 bt = gain_bucket(gain, wu < lim ? wu : lim);
            const u32 g = (MAX_BUCKETS * own_b + bt) * d.sections + (u % d.sections) + 1;
            d.gid[u] = g;
            btot_add(d, u, g, wu);
            listed = true;
        }
    }
    return listed;
}

// reb_eval of an eligible vertex whose neighbours all lie in its own block
// (own connection `own`), without its adjacency: no other block is touched,
// so weak sends it to underloaded[u % count] with gain -own and strong
// scores it 0 / 1 - own.  Returns whether u is a candidate.
__device__ __forceinline__ bool reb_interior(const JetDev &d, bool strong, const RebShared *rs, u32 u, s64 own) {
    const u32 own_b = d.work[u];
    const s64 wu = vwt(d, u);
    if (!strong) {
        const u32 nu = rs->n_under;
        const u32 dst = nu ? rs->under[u % nu] : own_b;
        d.dest[u] = dst;
        if (dst == own_b) return false;
        const u32 g = (MAX_BUCKETS * own_b + gain_bucket(-own, wu)) * d.sections + (u % d.sections);
        d.gid[u] = g;
        btot_add(d, u, g, wu);
        return true;
    }
    const s64 lim = d.bw[own_b] - d.lmax;
    const u32 g = (MAX_BUCKETS * own_b + gain_bucket(-own, wu < lim ? wu : lim)) * d.sections + (u % d.sections) + 1;
    d.gid[u] = g;
    btot_add(d, u, g, wu);
    return true;
}

// score offset of minibucket g for a vertex of block ub: the weight of the
// candidates in minibuckets [ub * width, g).  A strong candidate in the last
// minibucket of its block lands at index (ub + 1) * width (the reference's +1
// minibucket offset, jet:437), i.e. after its block's whole range.
__device__ __forceinline__ s64 bucket_prefix(const JetDev &d, u32 g, u32 ub) {
    return g >= (ub + 1) * d.width ? d.Tseg[ub] : d.Q[g];
}

// ---------------------------------------------------------------- sweep
// Workgroup r owns vertices [r * cap, (r + 1) * cap), taken in pieces of up to
// SWEEP_PIECE.  Phase A streams the piece, one vertex per thread: LP takes
// the cached decision of cached vertices, rebalance tests eligibility; the
// vertices to rebuild or score go to an LDS list.  Phase B hands that list to
// the workgroup's lane groups, G lanes per vertex.  Hubs are left to
// hub_sweep_kernel.  A pending best-partition copy is done in phase A.
template<int G>
__global__ __launch_bounds__(JB) void sweep_kernel(JetDev d) {
    extern __shared__ __align__(16) char smem[];
    __shared__ u32 s_cnt, s_nlist;
    __shared__ u64 s_red[JB / WAVE];
    const CtlHead hd = ctl_head(d.ctl);
    if (!live(hd)) return;
    u32 *region = d.cand + (u64) blockIdx.x * d.cap;
    const u32 mode = hd.mode;
    const bool save = hd.save != 0;
    const double ratio = hd.ratio;
    const u32 groups = blockDim.x / G;
    const u32 lane = threadIdx.x % G, gi = threadIdx.x / G;
    const size_t cgb = conn_group_bytes(d.k);
    GroupConn gc = group_conn(smem, d.k, gi);
    // listed vertices to rebuild, each with its adjacency start and a word
    // own block << 16 | locked << 15 | degree (degree <= hub_deg <= 2048, k < 4096)
    u32 *list = (u32 *) (smem + cgb * groups);
    u32 *lbeg = list + SWEEP_PIECE, *linfo = lbeg + SWEEP_PIECE;
    RebShared *rs = (RebShared *) (smem + cgb * groups + 12ull * SWEEP_PIECE);
    for (u32 i = lane; i < d.k; i += G) gc.cnt[i] = 0;
    if (lane == 0) *gc.tcount = 0;
    if (threadIdx.x == 0) s_cnt = 0;
    const bool strong = mode == 2;
    const s64 mbw = mode ? max_block_weight(d, strong) : 0;
    if (mode) reb_prologue(d, strong, mbw, rs);
    __syncthreads();
    u64 nev = 0, degs = 0;
    const u64 vbegin = (u64) blockIdx.x * d.cap;
    const u64 vend = vbegin + d.cap < d.n ? vbegin + d.cap : d.n;
    for (u64 p0 = vbegin; p0 < vend; p0 += SWEEP_PIECE) {
        const u64 p1 = p0 + SWEEP_PIECE < vend ? p0 + SWEEP_PIECE : vend;
        if (threadIdx.x == 0) s_nlist = 0;
        __syncthreads();
        for (u64 v0 = p0; v0 < p1; v0 += blockDim.x) {
            const u64 vv = v0 + threadIdx.x;
            const u32 v = (u32) vv;
            bool heavy = false, cand = false;
            u32 hb = 0, hinfo = 0;
            if (vv < p1) {
                if (save) d.part[v] = d.work[v];
                if (mode == 0) {
                    const u32 dc0 = d.dcache[v];
                    const u8 lk = d.lock[v], sv = stale_of(d)[v];  // loaded with the cache entry: one round trip
                    const u32 dc = (sv & 1u) ? NULL_PART : dc0;
                    if (dc == NULL_PART) {
                        // offsets and own block in one round trip, kept for phase B
                        const u32 xb = d.xadj[v], xe = d.xadj[v + 1], own = d.work[v];
                        heavy = xe - xb <= d.hub_deg;
                        hb = xb;
                        hinfo = own << 16 | (lk ? 1u << 15 : 0u) | (xe - xb);
                    } else {
                        d.dest[v] = dc;
                        if (dc != NO_MOVE && lk == 0) {
                            cand = true;
                            d.gain1[v] = d.gcache[v];
                        } else {
                            d.gain1[v] = GAIN_MIN;
                            d.lock[v] = 0;
                        }
                    }
                } else {
                    const u32 wv = d.work[v], xb = d.xadj[v], xe = d.xadj[v + 1];
                    if (xe - xb <= d.hub_deg && reb_eligible(d, strong, rs, v, wv)) {
                        const s64 ic0 = d.icache[v];
                        const s64 ic = (stale_of(d)[v] & 2u) ? -1 : ic0;
                        if (ic >= 0) {
                            cand = reb_interior(d, strong, rs, v, ic);
                        } else {
                            heavy = true;
                            hb = xb;
                            hinfo = wv << 16 | (xe - xb);
                        }
                    }
                }
            }
            append_local(cand, v, region, &s_cnt);
            append_local3(heavy, v, hb, hinfo, list, lbeg, linfo, &s_nlist);
        }
        __syncthreads();
        const u32 nl = s_nlist;
        for (u32 r0 = 0; r0 < nl; r0 += groups) {
            const u32 idx = r0 + gi;
            const bool active = idx < nl;
            const u32 u = active ? list[idx] : 0;
            const u32 ub = active ? lbeg[idx] : 0, info = active ? linfo[idx] : 0;
            const u32 own = info >> 16, deg = info & 0x7FFFu;
            if (active) {
                if (lane == 0) {
                    nev++;
                    degs += deg;
                }
                conn_accumulate_range<group_unroll<G>()>(d, gc, ub, (u64) ub + deg, lane, G);
            }
            wave_sync();
            const bool listed = mode == 0
                                    ? lp_eval<G>(d, gc, u, own, (info >> 15) & 1u, lane, ratio, nullptr, nullptr, active)
                                    : reb_eval<G>(d, strong, mbw, rs, gc, u, own, lane, nullptr, nullptr, active);
            wave_sync();
            if (active) conn_reset(gc, *gc.tcount, lane, G);
            wave_sync();
            if (active && lane == 0) *gc.tcount = 0;
            wave_sync();
            append_local(listed, u, region, &s_cnt);
        }
        __syncthreads();
    }
    nev = block_total(nev, s_red);
    degs = block_total(degs, s_red);
    if (threadIdx.x == 0) {
        d.cand_n[blockIdx.x] = s_cnt;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MISS] = nev;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MISS_DEG] = degs;
    }
}

// Per-block prefix of the minibucket totals (the reference's scan over all
// minibuckets, jet:440-447 / :696-703, restricted to what selection reads:
// differences inside one block's range).  Workgroup p scans block p's range
// [p * width, (p + 1) * width) and clears the totals for the next iteration.
constexpr int BKT_ITEMS = 8;
constexpr u32 BKT_CHUNK = JB * BKT_ITEMS;
constexpr u32 BKT_UNROLL = 8;  // copies loaded per round (loads in flight per thread)
__global__ __launch_bounds__(JB) void bucket_kernel(JetDev d) {
    __shared__ s64 red[JB / WAVE];
    __shared__ unsigned long long acc[BKT_CHUNK];
    const CtlHead hd = ctl_head(d.ctl);
    if (!live(hd) || hd.mode == 0) return;
    const u32 p = blockIdx.x;
    // racing mode: the classify launch records strong evictions' ends here
    if (d.race && hd.mode == 2 && threadIdx.x == 0) d.evict_adjust[p] = 0;
    const u64 base = (u64) p * d.width;
    const u64 stride = (u64) d.t + 1;
    s64 carry = 0;
    for (u32 c0 = 0; c0 < d.width; c0 += BKT_CHUNK) {
        const u32 len = d.width - c0 < BKT_CHUNK ? d.width - c0 : BKT_CHUNK;
        // sum the copies of this chunk into LDS and clear them; the loads of
        // BKT_UNROLL copies are issued before any is used
        for (u32 i = threadIdx.x; i < len; i += JB) acc[i] = 0;
        __syncthreads();
        for (u32 cb = 0; cb < d.nshard; cb += BKT_UNROLL) {
            for (u32 i = threadIdx.x; i < len; i += JB) {
                s64 x[BKT_UNROLL];
#pragma unroll
                for (u32 q = 0; q < BKT_UNROLL; ++q)
                    x[q] = cb + q < d.nshard ? d.btot[(cb + q) * stride + base + c0 + i] : 0;
#pragma unroll
                for (u32 q = 0; q < BKT_UNROLL; ++q) {
                    if (x[q]) {
                        atomicAdd(&acc[i], (unsigned long long) x[q]);
                        d.btot[(cb + q) * stride + base + c0 + i] = 0;
                    }
                }
            }
        }
        __syncthreads();
        const u32 first = threadIdx.x * BKT_ITEMS;
        s64 v[BKT_ITEMS];
        s64 sum = 0;
#pragma unroll
        for (int j = 0; j < BKT_ITEMS; ++j) {
            v[j] = first + j < len ? (s64) acc[first + j] : 0;
            sum += v[j];
        }
        s64 tot;
        s64 run = carry + block_exclusive_scan<JB>(sum, red, tot);
#pragma unroll
        for (int j = 0; j < BKT_ITEMS; ++j) {
            if (first + j < len) {
                d.Q[base + c0 + first + j] = run;
                run += v[j];
                d.R[base + c0 + first + j] = run;
            }
        }
        carry += tot;
        __syncthreads();  // acc is refilled by the next chunk
    }
    if (threadIdx.x == 0) {
        d.Tseg[p] = carry;
        if (p == d.k - 1)  // strong +1 entry of the last block
            for (u32 c = 0; c < d.nshard; ++c) d.btot[c * stride + (u64) d.k * d.width] = 0;
    }
}

// hubs: one workgroup each.  Phase 0 runs before the main sweep and scores
// rebalance candidates (their weight must be in the minibucket totals before
// the main sweep's last workgroup scans them); phase 1 runs after it and
// rebuilds uncached LP vertices (cached ones were taken by the main sweep).
__global__ __launch_bounds__(JB) void hub_sweep_kernel(JetDev d, const u32 *hubs, int phase) {
    extern __shared__ __align__(16) char smem[];
    __shared__ Best red[JB / WAVE];
    __shared__ s64 red2[JB / WAVE];
    JetCtl *ctl = d.ctl;
    const CtlHead hd = ctl_head(ctl);
    if (!live(hd)) return;
    const u32 mode = hd.mode;
    if ((mode == 0) != (phase == 1)) return;
    const u32 u = hubs[blockIdx.x];
    GroupConn gc = group_conn(smem, d.k, 0);
    RebShared *rs = (RebShared *) (smem + conn_group_bytes(d.k));
    const bool strong = mode == 2;
    s64 mbw = 0;
    if (mode == 0) {
        const u32 dc0 = d.dcache[u];
        if (!(stale_of(d)[u] & 1u) && dc0 != NULL_PART) return;
    } else {
        mbw = max_block_weight(d, strong);
        reb_prologue(d, strong, mbw, rs);
        if (!reb_eligible(d, strong, rs, u, d.work[u])) return;
    }
    for (u32 i = threadIdx.x; i < d.k; i += JB) gc.cnt[i] = 0;
    if (threadIdx.x == 0) *gc.tcount = 0;
    __syncthreads();
    const u32 own = d.work[u];
    const bool locked = d.lock[u] != 0;
    conn_accumulate<ADJ_UNROLL>(d, gc, u, threadIdx.x, JB);
    __syncthreads();
    const bool listed = mode == 0 ? lp_eval<JB>(d, gc, u, own, locked, threadIdx.x, hd.ratio, red, red2, true)
                                  : reb_eval<JB>(d, strong, mbw, rs, gc, u, own, threadIdx.x, red, red2, true);
    append(listed, u, d.cand_x, &ctl->st[ST_CAND]);
    if (threadIdx.x == 0) {
        if (d.count_bytes) {  // one update per hub workgroup of two counters: only when profiling
            atomicAdd(&ctl->st[ST_MISS], 1ull);
            atomicAdd(&ctl->st[ST_MISS_DEG], (unsigned long long) degree(d, u));
        }
    }
}

// ---------------------------------------------------------------- post: afterburner / classify
// afterburner (jet:269-328): selected candidates are locked, stamped and
// appended to the move list
template<u32 ADJ_UNROLL>
__device__ __forceinline__ s64 afterburner_change_range(const JetDev &d, u32 u, u64 b, u64 e, u32 lane, u32 stride) {
    const s64 ug = d.gain1[u];
    const u32 o = d.work[u], nn = d.dest[u];
    const s32 *Do = d.D ? d.D + (u64) o * d.k : nullptr, *Dn = d.D ? d.D + (u64) nn * d.k : nullptr;
    s64 change = 0;
    for (u64 j0 = b + lane; j0 < e; j0 += (u64) stride * ADJ_UNROLL) {
        u32 v[ADJ_UNROLL], vn[ADJ_UNROLL], vo[ADJ_UNROLL];
        s64 vg[ADJ_UNROLL], w[ADJ_UNROLL];
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            const u64 j = j0 + (u64) q * stride;
            v[q] = j < e ? d.adj[j] : 0u;
            w[q] = j < e ? ewt(d, (u32) j) : 0;
        }
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            const bool ok = j0 + (u64) q * stride < e;
            vg[q] = ok ? d.gain1[v[q]] : GAIN_MIN;
            vn[q] = ok ? d.dest[v[q]] : 0u;
            vo[q] = ok ? d.work[v[q]] : 0u;
        }
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            if (j0 + (u64) q * stride >= e) continue;
            const bool first = vg[q] > ug || (vg[q] == ug && v[q] < u);
            if (!first) continue;
            if (Do)  // promap:297-314: the gain of u once v has moved minus its gain before
                change += w[q] * (((s64) Do[vn[q]] - (s64) Dn[vn[q]]) - ((s64) Do[vo[q]] - (s64) Dn[vo[q]]));
            else
                change += w[q] * ((s64) (vn[q] == nn) - (s64) (vn[q] == o) + (s64) (vo[q] == o) - (s64) (vo[q] == nn));
        }
    }
    return change;
}
template<u32 ADJ_UNROLL>
__device__ __forceinline__ s64 afterburner_change(const JetDev &d, u32 u, u32 lane, u32 stride) {
    return afterburner_change_range<ADJ_UNROLL>(d, u, d.xadj[u], d.xadj[u + 1], lane, stride);
}

// afterburner over `len` candidates of `region`, G lanes each; selected
// vertices go to `mregion` (LDS counter); per-thread statistics in degs / gh.
// Each round loads the group's next candidate and its adjacency offsets, so
// that they arrive with this round's edges.
template<int G>
__device__ __forceinline__ void afterburner_list(const JetDev &d, u32 stamp, const u32 *region, u32 len, u32 *mregion,
                                                 u32 *s_cnt, u64 &degs, u64 &gh) {
    const bool do_hash = d.trace != nullptr;
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    u32 nxu = 0, nxb = 0, nxe = 0;
    if (gi < len) {
        nxu = region[gi];
        nxb = d.xadj[nxu];
        nxe = d.xadj[nxu + 1];
    }
    for (u32 i0 = 0; i0 < len; i0 += groups) {
        const u32 i = i0 + gi;
        const bool active = i < len;
        const u32 u = active ? nxu : 0, ub = nxb, ue = nxe;
        if (i + groups < len) {
            nxu = region[i + groups];
            nxb = d.xadj[nxu];
            nxe = d.xadj[nxu + 1];
        }
        const bool hub = active && ue - ub > d.hub_deg;
        s64 ch = (active && !hub) ? afterburner_change_range<group_unroll<G>()>(d, u, ub, ue, lane, G) : 0;
        ch = group_sum<G>(ch, (s64 *) nullptr);
        bool sel = false;
        if (active && !hub && lane == 0) {
            degs += ue - ub;
            const s64 ug = d.gain1[u];
            if (do_hash) gh += mix64d(mix64d(u) + (u64) ug);
            if (ug + ch >= 0) {
                d.lock[u] = 1;
                d.mark[u] = stamp;
                d.oldp[u] = d.work[u];
                sel = true;
            }
        }
        append_local(sel, u, mregion, s_cnt);
    }
}

// Rebalance selection (weak jet:718-744, strong jet:472-500) splits the
// candidates by score(u) = P[gid] - P[block begin] + off(u), off(u) in
// [0, btot[gid] - w(u)]: not taken if the minibucket starts at or beyond the
// limit; taken (weak) if the whole minibucket ends below the limit; the rest
// (and, for strong, every possibly taken candidate, whose exact position feeds
// adjust_scores) goes to the resolve list.  Both tests depend on the
// minibucket only, so a minibucket is either wholly in the resolve list or not
// at all, and offsets computed over the list are offsets in the minibucket.
__device__ __forceinline__ void classify_one(const JetDev &d, bool strong, u32 stamp, bool valid, u32 u,
                                             u32 *mregion, u32 *s_cnt) {
    JetCtl *ctl = d.ctl;
    {
        bool sure = false, boundary = false;
        u64 key = 0;
        if (valid && d.race) {
            // the offset came with the minibucket add: the score is known here;
            // strong candidates still need the cookie cutter (resolve launch)
            const u32 g = d.gid[u], ub = d.work[u];
            const s64 limit = d.bw[ub] - d.lmax;
            const s64 sc = bucket_prefix(d, g, ub) + d.gain1[u];
            if (sc < limit) {
                if (!strong) {
                    sure = true;
                    d.mark[u] = stamp;
                    d.oldp[u] = ub;
                } else {
                    boundary = true;
                    key = ((u64) g << 32) | u;
                    if (sc + vwt(d, u) >= limit) d.evict_adjust[ub] = sc + vwt(d, u);
                }
            }
        } else if (valid) {
            const u32 g = d.gid[u], ub = d.work[u];
            const s64 limit = d.bw[ub] - d.lmax;
            if (bucket_prefix(d, g, ub) < limit) {
                if (!strong && d.R[g] < limit) {
                    sure = true;
                    d.mark[u] = stamp;
                    d.oldp[u] = ub;
                } else {
                    boundary = true;
                    key = ((u64) g << 32) | u;
                }
            }
        }
        append_local(sure, u, mregion, s_cnt);
        append64(boundary, key, d.rlist, &ctl->st[ST_RLIST]);
    }
}

// this workgroup's candidate region, then its share of the extra list
__device__ __forceinline__ void classify_body(const JetDev &d, bool strong, u32 stamp, u32 len, u32 *mregion,
                                              u32 *s_cnt) {
    const u32 *region = d.cand + (u64) blockIdx.x * d.cap;
    for (u32 i0 = 0; i0 < len; i0 += blockDim.x) {
        const u32 i = i0 + threadIdx.x;
        classify_one(d, strong, stamp, i < len, i < len ? region[i] : 0u, mregion, s_cnt);
    }
    // the extra list in contiguous chunks of ceil(length / workgroups): the
    // move region has room for exactly that many extra entries
    const u32 lx = (u32) d.ctl->st[ST_CAND];
    const u32 chunk = (lx + gridDim.x - 1) / gridDim.x;
    const u32 x0 = blockIdx.x * chunk;
    const u32 x1 = x0 + chunk < lx ? x0 + chunk : lx;
    for (u32 i0 = x0; i0 < x1; i0 += blockDim.x) {
        const u32 i = i0 + threadIdx.x;
        classify_one(d, strong, stamp, i < x1, i < x1 ? d.cand_x[i] : 0u, mregion, s_cnt);
    }
}

// ---------------------------------------------------------------- edge-balanced lists
// A workgroup stages up to FLAT_CHUNK listed vertices in LDS with the
// exclusive prefix of their degrees (hubs count 0: their own kernels handle
// them) and walks the concatenated adjacencies with all its threads,
// ADJ_UNROLL edges per thread per round; the vertex of edge x comes from a
// binary search over the prefix.  Every thread gets the same number of edges
// whatever the degree mix: with one vertex per lane group a workgroup waited
// on its highest-degree vertex, and held few vertices in flight when the
// lists are long (rebalancing on skewed graphs moves a large share of all
// vertices per iteration).
constexpr u32 FLAT_CHUNK = JB;
struct FlatLds {
    u32 u[FLAT_CHUNK];
    u32 start[FLAT_CHUNK];  // xadj[u]
    u32 off[FLAT_CHUNK + 1];
    u32 a[FLAT_CHUNK], b[FLAT_CHUNK];
    s64 g[FLAT_CHUNK];
    unsigned long long acc[FLAT_CHUNK];
    u32 red[JB / WAVE];
};

// last i in [0, cnt) with off[i] <= x (off non-decreasing, off[0] = 0)
__device__ __forceinline__ u32 flat_find(const u32 *off, u32 cnt, u32 x) {
    u32 lo = 0, hi = cnt;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// degrees of the staged vertices -> L.off; returns the chunk's edge count
__device__ __forceinline__ u32 flat_prefix(FlatLds &L, u32 cnt, u32 deg) {
    u32 tot;
    const u32 ex = block_exclusive_scan<JB>(threadIdx.x < cnt ? deg : 0u, L.red, tot);
    if (threadIdx.x < cnt) L.off[threadIdx.x] = ex;
    if (threadIdx.x == 0) L.off[cnt] = tot;
    __syncthreads();
    return tot;
}

// afterburner (promap:284-322 / jet:269-328) over `len` candidates of
// `region`, edge-balanced; selected vertices go to `mregion`
__device__ void afterburner_flat(const JetDev &d, u32 stamp, const u32 *region, u32 len, u32 *mregion, u32 *s_cnt,
                                 FlatLds &L, u64 &degs, u64 &gh) {
    const bool do_hash = d.trace != nullptr;
    for (u32 c0 = 0; c0 < len; c0 += FLAT_CHUNK) {
        const u32 cnt = len - c0 < FLAT_CHUNK ? len - c0 : FLAT_CHUNK;
        const u32 t = threadIdx.x;
        u32 u = 0, deg = 0;
        bool hub = true;
        if (t < cnt) {
            u = region[c0 + t];
            const u32 b = d.xadj[u], e = d.xadj[u + 1];
            hub = e - b > d.hub_deg;
            deg = hub ? 0u : e - b;
            L.u[t] = u;
            L.start[t] = b;
            L.g[t] = d.gain1[u];
            L.a[t] = d.work[u];
            L.b[t] = d.dest[u];
            L.acc[t] = 0;
        }
        const u32 E = flat_prefix(L, cnt, deg);
        for (u32 x0 = t; x0 < E; x0 += JB * ADJ_UNROLL) {
            u32 iv[ADJ_UNROLL], v[ADJ_UNROLL], vn[ADJ_UNROLL], vo[ADJ_UNROLL];
            s64 vg[ADJ_UNROLL], w[ADJ_UNROLL];
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                const u32 x = x0 + q * JB;
                const bool ok = x < E;
                iv[q] = ok ? flat_find(L.off, cnt, x) : 0u;
                const u32 j = L.start[iv[q]] + (x - L.off[iv[q]]);
                v[q] = ok ? d.adj[j] : 0u;
                w[q] = ok ? ewt(d, j) : 0;
            }
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                const bool ok = x0 + q * JB < E;
                vg[q] = ok ? d.gain1[v[q]] : GAIN_MIN;
                vn[q] = ok ? d.dest[v[q]] : 0u;
                vo[q] = ok ? d.work[v[q]] : 0u;
            }
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                if (x0 + q * JB >= E) continue;
                const u32 i = iv[q], cu = L.u[i];
                const s64 ug = L.g[i];
                if (!(vg[q] > ug || (vg[q] == ug && v[q] < cu))) continue;
                const u32 o = L.a[i], nn = L.b[i];
                s64 c;
                if (d.D)  // promap:297-314
                    c = w[q] * (((s64) d.D[(u64) o * d.k + vn[q]] - (s64) d.D[(u64) nn * d.k + vn[q]]) -
                                ((s64) d.D[(u64) o * d.k + vo[q]] - (s64) d.D[(u64) nn * d.k + vo[q]]));
                else
                    c = w[q] * ((s64) (vn[q] == nn) - (s64) (vn[q] == o) + (s64) (vo[q] == o) - (s64) (vo[q] == nn));
                if (c) atomicAdd(&L.acc[i], (unsigned long long) c);
            }
        }
        __syncthreads();
        bool sel = false;
        if (t < cnt && !hub) {
            degs += deg;
            const s64 ug = L.g[t];
            if (do_hash) gh += mix64d(mix64d(u) + (u64) ug);
            if (ug + (s64) L.acc[t] >= 0) {
                d.lock[u] = 1;
                d.mark[u] = stamp;
                d.oldp[u] = L.a[t];
                sel = true;
            }
        }
        append_local(sel, u, mregion, s_cnt);
        __syncthreads();  // the next chunk restages the LDS
    }
}

template<int G, bool FLAT>
__global__ __launch_bounds__(JB) void post_kernel(JetDev d) {
    __shared__ u32 s_cnt;
    __shared__ u64 s_red[JB / WAVE];
    JetCtl *ctl = d.ctl;
    const u32 len = d.cand_n[blockIdx.x];  // issued with the control block: one round trip
    const CtlHead hd = ctl_head(ctl);
    if (!live(hd)) return;
    const u32 mode = hd.mode, stamp = hd.stamp;
    u32 *mregion = d.moves + (u64) blockIdx.x * d.cap_mv;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    if (mode == 0) {
        u64 gh = 0, degs = 0;
        if constexpr (FLAT) {
            __shared__ FlatLds L;
            afterburner_flat(d, stamp, d.cand + (u64) blockIdx.x * d.cap, len, mregion, &s_cnt, L, degs, gh);
        } else {
            afterburner_list<G>(d, stamp, d.cand + (u64) blockIdx.x * d.cap, len, mregion, &s_cnt, degs, gh);
        }
        degs = block_total(degs, s_red);
        gh = block_total(gh, s_red);
        if (threadIdx.x == 0) {
            d.pstat[(u64) blockIdx.x * PT_COUNT + PT_CAND_DEG] = degs;
            d.pstat[(u64) blockIdx.x * PT_COUNT + PT_GHASH] = gh;
        }
    } else {
        classify_body(d, mode == 2, stamp, len, mregion, &s_cnt);
        if (threadIdx.x == 0) {
            d.pstat[(u64) blockIdx.x * PT_COUNT + PT_CAND_DEG] = 0;
            d.pstat[(u64) blockIdx.x * PT_COUNT + PT_GHASH] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) d.moves_n[blockIdx.x] = s_cnt;
}

__global__ __launch_bounds__(JB) void hub_post_kernel(JetDev d, const u32 *hubs) {
    __shared__ s64 red[JB / WAVE];
    JetCtl *ctl = d.ctl;
    const CtlHead hd = ctl_head(ctl);
    const u32 u = hubs[blockIdx.x];
    if (!live(hd) || hd.mode != 0) return;
    const u32 stamp = hd.stamp;
    const s64 ug = d.gain1[u];
    if (ug == GAIN_MIN) return;  // not a candidate
    s64 ch = afterburner_change<ADJ_UNROLL>(d, u, threadIdx.x, JB);
    ch = group_sum<JB>(ch, red);
    bool sel = false;
    if (threadIdx.x == 0) {
        if (d.count_bytes) atomicAdd(&ctl->st[ST_CAND_DEG], (unsigned long long) degree(d, u));
        if (d.trace) atomicAdd(&ctl->st[ST_GHASH], (unsigned long long) mix64d(mix64d(u) + (u64) ug));
        if (ug + ch >= 0) {
            d.lock[u] = 1;
            d.mark[u] = stamp;
            d.oldp[u] = d.work[u];
            sel = true;
        }
    }
    append(sel, u, d.moves_x, &ctl->st[ST_MOVES]);
}

// ---------------------------------------------------------------- resolve
// the chunk containing position pos: the last id in [0, k] with es[id] <= pos
// (es = evict_start, non-decreasing from es[0] = 0; the reference walks it
// linearly, jet:564-566)
template<typename P>
__device__ __forceinline__ s64 chunk_index(P es, u32 k, s64 pos) {
    u32 lo = 0, hi = k + 1;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (es[mid] <= pos) lo = mid + 1;
        else hi = mid;
    }
    return (s64) lo - 1;
}

// destination chunk of an evicted vertex at position pos (adjust_scores,
// jet:563-575, first pass): id < k with room -> that block; id >= k -> stays;
// otherwise the vertex must take an overflow position (returns NO_MOVE)
__device__ __forceinline__ u32 chunk_of(const s64 *es, u32 k, s64 pos, s64 wu, u32 ub) {
    const s64 id = chunk_index(es, k, pos);
    if (id >= 0 && id < (s64) k && wu / 2 <= es[id + 1] - pos) return (u32) id;
    if (id < (s64) k) return NO_MOVE;
    return ub;
}

// the smallest multiple of wu at or above dist (both positive); 32-bit
// division where both fit 31 bits (the 64-bit one is a long instruction
// sequence, and the overflow step runs it serially)
__device__ __forceinline__ s64 round_up_to(s64 dist, s64 wu) {
    if (((u64) dist | (u64) wu) < (1ull << 31)) {
        const u32 q = ((u32) dist + (u32) wu - 1u) / (u32) wu;
        return (s64) q * wu;
    }
    return (dist + wu - 1) / wu * wu;
}

// the serial adjust_scores loop of one overflowing vertex (jet:563-585): its
// chunk, or k when the vertex stays in its block.  After its own position the
// vertex takes the counter's positions, counter, counter + wu, ..., until one
// lands in a chunk with room for wu / 2 (or past the last chunk).  Positions
// later in a chunk that has no room for the first have less, so the loop
// jumps to the first position at or past the next chunk's start: the same
// position and counter as one take at a time, in at most one step per chunk.
template<typename P>
__device__ __forceinline__ u32 overflow_chunk(P es, u32 k, s64 pos, s64 wu, s64 &counter) {
    s64 id = chunk_index(es, k, pos);
    if (id >= 0 && id < (s64) k && wu / 2 <= es[id + 1] - pos) return (u32) id;
    if (id >= (s64) k) return k;
    pos = counter;
    while (true) {
        id = chunk_index(es, k, pos);
        if (id >= (s64) k || (id >= 0 && wu / 2 <= es[id + 1] - pos)) break;
        if (wu <= 0) {  // a weightless vertex never leaves a position before the first chunk
            id = k;
            break;
        }
        const s64 next = es[id + 1];  // > pos; es[0] when id = -1
        pos += round_up_to(next - pos, wu);
    }
    counter = pos + wu;
    return id >= (s64) k ? k : (u32) id;
}

// The chunk bounds in LDS through an LDS-typed pointer: a pointer that may
// point to either memory compiles to flat accesses, and the overflow step is
// a chain of dependent reads.
typedef __attribute__((address_space(3))) const s64 lds_s64;
__device__ __forceinline__ lds_s64 *les_of(const s64 *p) { return (lds_s64 *) p; }

// cookie cutter (jet:508-552), by the first wave of the workgroup; on entry
// evict_adjust holds the filter's values, on exit the eviction position of
// each block's first candidate, so that position = score + evict_adjust[block].
// With les, the chunk bounds also go to that LDS array; with lcnt, the
// overflow counter's start goes there instead of to device memory.
__device__ void cookie_cutter(const JetDev &d, s64 mbw, s64 *les = nullptr, s64 *lcnt = nullptr) {
    const u32 lane = threadIdx.x;
    s64 carry = 0, carry2 = 0;
    for (u32 b0 = 0; b0 < d.k; b0 += WAVE) {
        const u32 p = b0 + lane;
        s64 add = 0, cap = 0;
        if (p < d.k) {
            add = d.evict_adjust[p];
            if (add == 0) add = d.Tseg[p];
            cap = mbw > d.bw[p] ? mbw - d.bw[p] : 0;
        }
        const s64 inc = wave_inclusive_scan(add), inc2 = wave_inclusive_scan(cap);
        if (p < d.k) {
            d.evict_adjust[p] = carry + inc - add;
            d.evict_start[p + 1] = carry2 + inc2;
            if (les) les[p + 1] = carry2 + inc2;
        }
        carry += __shfl(inc, WAVE - 1, WAVE);
        carry2 += __shfl(inc2, WAVE - 1, WAVE);
    }
    if (lane == 0) {
        if (lcnt) *lcnt = carry;
        else *d.counter = carry;
        d.evict_start[0] = 0;
        if (les) les[0] = 0;
    }
}

// One workgroup sorts the resolve list by (minibucket, id) in LDS, derives each
// entry's offset inside its minibucket and selects; strong rebalancing then
// runs the cookie cutter and adjust_scores.  Lists longer than CAP, or more
// than OVF_CAP overflowing vertices, stall the loop for the host's global path
// (nothing has been appended to the move list at that point).
struct ResolveLds {
    u64 *skey;
    s64 *soff;
    u8 *ssel;
    s64 *red;    // RB / WAVE
    u32 *ocount; // [0] overflow entries, [1] the first that takes the counter twice
    s64 *cnt;    // the overflow counter during one resolve step
    u32 *olist;  // OVF_CAP
    u32 *ovid;   // OVF_CAP
    u32 *oent;   // OVF_CAP
    s64 *s_es;   // ES_CAP
};

// LDS bytes of ResolveLds for CAP entries (16-byte aligned parts)
__host__ __device__ constexpr size_t resolve_lds_bytes(u32 cap) {
    return 16ull * cap + ((cap + 15) & ~15u) + 8ull * (RB / WAVE) + 16 + 12ull * OVF_CAP + 8ull * ES_CAP;
}
__device__ inline ResolveLds resolve_lds(char *base, u32 cap) {
    ResolveLds r;
    r.skey = (u64 *) base;
    r.soff = (s64 *) (base + 8ull * cap);
    r.ssel = (u8 *) (base + 16ull * cap);
    char *p = base + 16ull * cap + ((cap + 15) & ~15u);
    r.red = (s64 *) p;
    p += 8ull * (RB / WAVE);
    r.ocount = (u32 *) p;
    r.cnt = (s64 *) (p + 8);
    p += 16;
    r.olist = (u32 *) p;
    p += 4ull * OVF_CAP;
    r.ovid = (u32 *) p;
    p += 4ull * OVF_CAP;
    r.oent = (u32 *) p;
    p += 4ull * OVF_CAP;
    r.s_es = (s64 *) p;
    return r;
}

// The strong overflow step (jet:563-585) of oc <= OVF_CAP entries, by the
// whole workgroup (NT threads, NT >= OVF_CAP).  olist[0, oc) index the resolve
// arrays; skey holds weight << 32 | vertex and soff the position that did not
// fit.  Serially, in vertex id order, each entry takes the counter's position
// and adds its weight, and takes again while that position lands in a chunk
// without room.  If every entry takes once, the entry of rank r gets the
// counter plus the weights of the entries ranked before it: ranks come from
// counting smaller ids, the positions from a scan, and all are checked at
// once; from the first entry that must take again on, one thread finishes
// serially.  On exit soff holds each entry's chunk (k: it stays).
template<int NT, typename P>
__device__ __forceinline__ void overflow_block(const JetDev &d, P es, const ResolveLds &lds, u32 oc) {
    static_assert(NT >= (int) OVF_CAP, "one thread per overflow entry");
    u64 *skey = lds.skey;
    s64 *soff = lds.soff;
    const u32 t = threadIdx.x, k = d.k;
    if (t < oc) lds.ovid[t] = (u32) skey[lds.olist[t]];
    if (t == 0) lds.ocount[1] = oc;
    __syncthreads();
    if (t < oc) {
        const u32 my = lds.ovid[t];
        u32 r = 0;
        for (u32 q = 0; q < oc; ++q) r += lds.ovid[q] < my ? 1u : 0u;  // ids are distinct
        lds.oent[r] = lds.olist[t];
    }
    __syncthreads();
    const bool act = t < oc;
    const u32 e = act ? lds.oent[t] : 0;
    const s64 w = act ? (s64) (skey[e] >> 32) : 0;
    if (act) lds.ovid[t] = (u32) w;  // the weights in rank order, for the serial part
    s64 tot;
    const s64 ex = block_exclusive_scan<NT>(w, lds.red, tot);
    const s64 c0 = *lds.cnt;
    const s64 pos = c0 + ex;
    u32 id = k;
    if (act) {
        const s64 cid = chunk_index(es, k, pos);
        if (cid >= 0 && cid < (s64) k && w / 2 <= es[cid + 1] - pos) id = (u32) cid;
        else if (cid < (s64) k) atomicMin(&lds.ocount[1], t);
    }
    __syncthreads();
    const u32 first = lds.ocount[1];
    if (act && t < first) soff[e] = id;
    if (t == first) lds.red[0] = pos;  // the counter when that entry comes
    __syncthreads();
    if (t == 0) {
        // the rest one at a time: each takes the counter (its own position is
        // known to fail), and since the counter only grows, the chunk search
        // walks forward from the previous position's chunk
        s64 counter = first < oc ? lds.red[0] : c0 + tot;
        s64 cid = first < oc ? chunk_index(es, k, counter) : 0;
        for (u32 r = first; r < oc; ++r) {
            const u32 i = lds.oent[r];
            const s64 wu = (s64) lds.ovid[r];
            s64 pos = counter;
            u32 got;
            while (true) {
                while (cid < (s64) k && es[cid + 1] <= pos) ++cid;
                if (cid >= (s64) k) {
                    got = k;
                    break;
                }
                if (cid >= 0 && wu / 2 <= es[cid + 1] - pos) {
                    got = (u32) cid;
                    break;
                }
                if (wu <= 0) {  // as overflow_chunk
                    got = k;
                    break;
                }
                pos += round_up_to(es[cid + 1] - pos, wu);  // first take at or past the next chunk
            }
            soff[i] = got;
            counter = pos + wu;
        }
    }
}

// Bitonic sort of up to RB keys ascending into skey[0, P2) (P2 a power of
// two, entries from L on padded with ~0), key i in thread i's register:
// stages whose partner lies in the same wave exchange by lane shuffles
// without a barrier, only stages across waves go through LDS.  For longer
// lists the LDS form is faster: a shuffle costs about as much as an LDS
// access, and the LDS form skips the idle threads' work.
__device__ __forceinline__ void sort_keys_one_per_thread(const u64 *src, u32 L, u32 P2, u64 *skey) {
    const u32 i = threadIdx.x;
    u64 r = i < L ? src[i] : ~0ull;
    for (u32 kk = 2; kk <= P2; kk <<= 1) {
        for (u32 j = kk >> 1; j > 0; j >>= 1) {
            u64 b;
            if (j >= WAVE) {
                __syncthreads();  // the previous LDS stage has read its partners
                if (i < P2) skey[i] = r;
                __syncthreads();
                b = i < P2 ? skey[i ^ j] : r;
            } else {
                b = __shfl_xor(r, (int) j);
            }
            const bool take_min = ((i & j) == 0) == ((i & kk) == 0);
            r = take_min ? (b < r ? b : r) : (b > r ? b : r);
        }
    }
    __syncthreads();
    if (i < P2) skey[i] = r;
    __syncthreads();
}

// Resolves a list of L entries with NT threads; false: the list (or the
// strong overflow) does not fit and nothing was appended.  Moves go to the
// extra move list, counted in LDS (`local_cnt`) or in st[ST_MOVES].
template<u32 CAP, int NT = RB>
__device__ __forceinline__ bool resolve_body(const JetDev &d, bool strong, u32 stamp, u32 L, const ResolveLds &lds, u32 *local_cnt) {
    u64 *skey = lds.skey;
    s64 *soff = lds.soff;
    u8 *ssel = lds.ssel;
    s64 *red = lds.red;
    u32 &ocount = *lds.ocount;
    u32 *olist = lds.olist;
    s64 *s_es = lds.s_es;
    if (L == 0) return true;
    if (L > CAP || L > d.rcap) return false;
    u32 P2 = 1;
    while (P2 < L) P2 <<= 1;
    if (threadIdx.x == 0) ocount = 0;
    if (strong)
        for (u32 b = threadIdx.x; b < d.k; b += NT) d.evict_adjust[b] = 0;
    if (P2 <= NT) {
        sort_keys_one_per_thread(d.rlist, L, P2, skey);
    } else {
    for (u32 i = threadIdx.x; i < P2; i += NT) skey[i] = i < L ? d.rlist[i] : ~0ull;
    __syncthreads();
    // bitonic sort, ascending
    for (u32 kk = 2; kk <= P2; kk <<= 1) {
        for (u32 j = kk >> 1; j > 0; j >>= 1) {
            for (u32 i = threadIdx.x; i < P2; i += NT) {
                const u32 p = i ^ j;
                if (p > i) {
                    const u64 a = skey[i], b = skey[p];
                    if (((i & kk) == 0) ? (a > b) : (a < b)) {
                        skey[i] = b;
                        skey[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    }
    // exclusive prefix of the weights in sorted order
    s64 carry = 0;
    for (u32 i0 = 0; i0 < L; i0 += NT) {
        const u32 i = i0 + threadIdx.x;
        const s64 w = i < L ? vwt(d, (u32) skey[i]) : 0;
        s64 tot;
        const s64 e = block_exclusive_scan<NT>(w, red, tot);
        if (i < L) soff[i] = carry + e;
        carry += tot;
    }
    __syncthreads();
    // offset inside the minibucket = prefix - prefix at the minibucket's first entry
    s64 offv[CAP / NT];
#pragma unroll
    for (u32 q = 0; q < CAP / NT; ++q) {
        const u32 i = q * NT + threadIdx.x;
        offv[q] = 0;
        if (i < L) {
            const u64 head = skey[i] & 0xFFFFFFFF00000000ull;
            u32 lo = 0, hi = i;
            while (lo < hi) {
                const u32 mid = (lo + hi) >> 1;
                if (skey[mid] < head) lo = mid + 1;
                else hi = mid;
            }
            offv[q] = soff[i] - soff[lo];
        }
    }
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < CAP / NT; ++q) {
        const u32 i = q * NT + threadIdx.x;
        if (i < L) soff[i] = offv[q];
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < L; i += NT) {
        const u32 u = (u32) skey[i];
        const u32 g = (u32) (skey[i] >> 32);
        const u32 ub = d.work[u];
        const s64 score = bucket_prefix(d, g, ub) + soff[i];
        const s64 limit = d.bw[ub] - d.lmax;
        const bool take = score < limit;
        ssel[i] = take;
        if (take && strong && score + vwt(d, u) >= limit) d.evict_adjust[ub] = score + vwt(d, u);
        if (take && !strong) {
            d.mark[u] = stamp;
            d.oldp[u] = ub;
        }
    }
    __syncthreads();
    if (strong) {
        const bool es_lds = d.k + 1 <= ES_CAP;
        if (threadIdx.x < WAVE) cookie_cutter(d, max_block_weight(d, true), es_lds ? s_es : nullptr, lds.cnt);
        __syncthreads();
        const s64 *es = es_lds ? s_es : d.evict_start;
        // adjust_scores, first pass; overflowing vertices are resolved below
        for (u32 i = threadIdx.x; i < L; i += NT) {
            if (!ssel[i]) continue;
            const u32 u = (u32) skey[i];
            const u32 g = (u32) (skey[i] >> 32);
            const u32 ub = d.work[u];
            const s64 pos = bucket_prefix(d, g, ub) + soff[i] + d.evict_adjust[ub];
            const u32 dst = chunk_of(es, d.k, pos, vwt(d, u), ub);
            d.mark[u] = stamp;
            d.oldp[u] = ub;
            if (dst == NO_MOVE) {
                // the serial step below reads its position and weight from LDS
                // (vertex weights fit 31 bits: coarse weights are s32)
                const u32 q = atomicAdd(&ocount, 1u);
                if (q < OVF_CAP) olist[q] = i;
                ssel[i] = 2;
                soff[i] = pos;
                skey[i] = ((u64) vwt(d, u) << 32) | u;
            } else {
                d.dest[u] = dst;
            }
        }
        __syncthreads();
        const u32 oc = ocount;
        if (oc > OVF_CAP) return false;
        if (oc) {
            if (es_lds) overflow_block<NT>(d, les_of(s_es), lds, oc);
            else overflow_block<NT>(d, (const s64 *) d.evict_start, lds, oc);
        }
        __syncthreads();
        for (u32 q = threadIdx.x; q < oc; q += NT) {
            const u32 i = olist[q], u = (u32) skey[i], id = (u32) soff[i];
            d.dest[u] = id < d.k ? id : d.work[u];
        }
    }
    for (u32 i0 = 0; i0 < L; i0 += NT) {
        const u32 i = i0 + threadIdx.x;
        const bool take = i < L && ssel[i] != 0;
        if (local_cnt) append_local(take, take ? (u32) skey[i] : 0u, d.moves_x, local_cnt);
        else append(take, take ? (u32) skey[i] : 0u, d.moves_x, &d.ctl->st[ST_MOVES]);
    }
    return true;
}

// Racing mode, strong rebalancing: the list holds every taken candidate with
// its score known; the cookie cutter, then each candidate's chunk, and an
// overflowing candidate's next position from a fetch-add on the overflow
// counter (jet:576-583) in whatever order the lanes get there.
__device__ void resolve_race(const JetDev &d, u32 stamp, u32 L, s64 *s_es) {
    if (threadIdx.x < WAVE) cookie_cutter(d, max_block_weight(d, true));
    __syncthreads();
    const bool es_lds = d.k + 1 <= ES_CAP;
    if (es_lds)
        for (u32 i = threadIdx.x; i <= d.k; i += RB) s_es[i] = d.evict_start[i];
    __syncthreads();
    const s64 *es = es_lds ? s_es : d.evict_start;
    for (u32 i0 = 0; i0 < L; i0 += RB) {
        const u32 i = i0 + threadIdx.x;
        u32 u = 0;
        if (i < L) {
            const u64 key = d.rlist[i];
            u = (u32) key;
            const u32 g = (u32) (key >> 32), ub = d.work[u];
            const s64 wu = vwt(d, u);
            s64 pos = bucket_prefix(d, g, ub) + d.gain1[u] + d.evict_adjust[ub];
            u32 dst = ub;
            while (true) {
                const s64 id = chunk_index(es, d.k, pos);
                if (id >= 0 && id < (s64) d.k && wu / 2 <= es[id + 1] - pos) {
                    dst = (u32) id;
                    break;
                }
                if (id >= (s64) d.k) break;
                pos = (s64) atomicAdd((unsigned long long *) d.counter, (unsigned long long) wu);
            }
            d.dest[u] = dst;
            d.mark[u] = stamp;
            d.oldp[u] = ub;
        }
        append(i < L, u, d.moves_x, &d.ctl->st[ST_MOVES]);
    }
}

template<u32 CAP>
__global__ __launch_bounds__(RB) void resolve_kernel(JetDev d) {
    __shared__ __align__(16) char smem[resolve_lds_bytes(CAP)];
    JetCtl *ctl = d.ctl;
    const CtlHead hd = ctl_head(ctl);
    const u32 L = (u32) ctl->st[ST_RLIST];  // the same cache line as the head
    if (!live(hd) || hd.mode == 0) return;
    if (d.race) {
        if (hd.mode == 2 && L) resolve_race(d, hd.stamp, L, (s64 *) smem);
        return;
    }
    const ResolveLds lds = resolve_lds(smem, CAP);
    const bool ok = resolve_body<CAP>(d, hd.mode == 2, hd.stamp, L, lds, nullptr);
    if (!ok && threadIdx.x == 0) ctl->stalled = 1;
}

// ---------------------------------------------------------------- apply moves and control
// exact cut change, cache invalidation (u and N(u)), partition and block
// weight update.  Non-moved neighbours are never written and moved
// neighbours' old block comes from oldp, so concurrent writes to work[] are safe.
// the change of the objective over edges [b, e) of moved vertex u (lanes
// `lane`, `lane + stride`, ...), invalidating the caches of the neighbours
template<u32 ADJ_UNROLL>
__device__ __forceinline__ s64 move_delta_range(const JetDev &d, u32 u, u64 b, u64 e, u32 lane, u32 stride,
                                                u32 stamp) {
    const u32 ou = d.oldp[u], nu = d.dest[u];
    const s32 *Dn = d.D ? d.D + (u64) nu * d.k : nullptr, *Do = d.D ? d.D + (u64) ou * d.k : nullptr;
    s64 delta = 0;
    for (u64 j0 = b + lane; j0 < e; j0 += (u64) stride * ADJ_UNROLL) {
        u32 v[ADJ_UNROLL], mk[ADJ_UNROLL], op[ADJ_UNROLL], ds[ADJ_UNROLL], wk[ADJ_UNROLL], wi[ADJ_UNROLL];
        u8 rf[ADJ_UNROLL];
        s64 w[ADJ_UNROLL];
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            const u64 j = j0 + (u64) q * stride;
            v[q] = j < e ? d.adj[j] : 0u;
            w[q] = j < e ? ewt(d, (u32) j) : 0;
        }
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            const bool ok = j0 + (u64) q * stride < e;
            mk[q] = ok ? d.mark[v[q]] : 0u;
            op[q] = ok ? d.oldp[v[q]] : 0u;
            ds[q] = ok ? d.dest[v[q]] : 0u;
            wk[q] = ok ? d.work[v[q]] : 0u;
            rf[q] = ok && d.rflag ? d.rflag[v[q]] : (u8) 0;
            wi[q] = ok && d.wcnt ? d.wid[v[q]] : 0u;
        }
#pragma unroll
        for (u32 q = 0; q < ADJ_UNROLL; ++q) {
            if (j0 + (u64) q * stride >= e) continue;
            const bool mv = mk[q] == stamp;
            const u32 ov = mv ? op[q] : wk[q];
            const u32 nv = mv ? ds[q] : ov;
            if (Dn) {
                // comm cost change (promap:1174-1207): moves that keep their
                // block contribute nothing; an edge to a vertex that did not
                // change block counts twice (its other direction is not visited)
                if (ou != nu) {
                    const s64 x = w[q] * ((s64) Dn[nv] - (s64) Do[ov]);
                    delta += nv != ov ? x : 2 * x;
                }
            } else {
                const s64 x = w[q] * ((s64) (nu != nv) - (s64) (ou != ov));
                delta += mv ? x : 2 * x;
            }
            stale_of(d)[v[q]] = 3;
            row_touch(d, v[q], rf[q], wi[q], ou, nu);
        }
    }
    return delta;
}

template<u32 U>
__device__ __forceinline__ s64 move_delta(const JetDev &d, u32 u, u32 lane, u32 stride, u32 stamp) {
    return move_delta_range<U>(d, u, d.xadj[u], d.xadj[u + 1], lane, stride, stamp);
}

// commits a move; the block weight changes go to `bw` (the global array, or a
// workgroup's LDS copy that is added to it once per block at the end)
__device__ __forceinline__ void move_commit(const JetDev &d, u32 u, s64 *bw) {
    const u32 ou = d.oldp[u], nu = d.dest[u];
    const s64 w = vwt(d, u);
    stale_of(d)[u] = 3;
    d.work[u] = nu;
    atomicAdd((unsigned long long *) &bw[ou], (unsigned long long) (-w));
    atomicAdd((unsigned long long *) &bw[nu], (unsigned long long) w);
}
__device__ __forceinline__ void move_commit(const JetDev &d, u32 u) { move_commit(d, u, d.bw); }

// algorithmic bytes of one iteration's launches (w = 2 when edge weights are
// stored, else 1; E(X) = summed degree of the vertex set X):
//   LP sweep: 25n (cache, lock, dest, gain) + 12 per rebuilt vertex + 4w E +
//     4 per distinct neighbour block read (capped at 4n);
//   rebalance sweep: (8 + 4 if vertex weights) n + 4w E(scored) + min(4E, 4n)
//     + 8 per candidate;
//   buckets: 32 t (read and clear the totals, write two prefixes);
//   afterburner: 16|C| + 4w E(C) + min(16 E(C), 16n);
//   classify: (20 + 4 if vertex weights) per candidate;
//   resolve: 32 per resolve-list entry;
//   apply: 16|M| + 4w E(M) + min(16 E(M), 16n);
//   a pending best-partition copy adds 8n to the sweep.
__device__ void account(const JetDev &d, JetCtl *ctl, const unsigned long long *rep, u32 mode, bool saved) {
    const double w = d.ew ? 2.0 : 1.0, n = (double) d.n, wv = d.vw ? 4.0 : 0.0;
    const double nm = (double) rep[ST_MISS], em = (double) rep[ST_MISS_DEG];
    const double nc = (double) rep[ST_CAND], ec = (double) rep[ST_CAND_DEG];
    const double nmv = (double) rep[ST_MOVES], emv = (double) rep[ST_MOVE_DEG];
    double sweep = saved ? 8.0 * n : 0.0, buckets = 0.0, post, resolve = 0.0;
    if (mode == 0) {
        sweep += 25.0 * n + 12.0 * nm + 4.0 * w * em + fmin(4.0 * em, 4.0 * n);
        post = 16.0 * nc + 4.0 * w * ec + fmin(16.0 * ec, 16.0 * n);
    } else {
        sweep += (8.0 + wv) * n + 4.0 * w * em + fmin(4.0 * em, 4.0 * n) + 8.0 * nc;
        buckets = 32.0 * (double) d.t;
        post = (20.0 + wv) * nc;
        resolve = 32.0 * (double) rep[ST_RLIST];
    }
    ctl->bytes[0] += sweep;
    ctl->bytes[1] += buckets;
    ctl->bytes[2] += post;
    ctl->bytes[3] += resolve;
    ctl->bytes[4] += 16.0 * nmv + 4.0 * w * emv + fmin(16.0 * emv, 16.0 * n);
}

__device__ void trace_record(const JetDev &d, JetCtl *ctl, const unsigned long long *rep, u32 mode, s64 cur,
                             s64 cur_maxw) {
    heipa_jet_trace &t = d.trace[ctl->n_trace % TRACE_RING];
    t.ratio_idx = ctl->ri;
    t.iteration = ctl->it;
    t.mode = mode;
    t.n_cand = (u32) rep[ST_CAND];
    t.n_moves = (u32) rep[ST_MOVES];
    t.pad = 0;
    t.cut = cur;
    t.maxw = cur_maxw;
    t.moves_hash = rep[ST_MHASH];
    t.gain_hash = mode == 0 ? rep[ST_GHASH] : 0;
    ctl->n_trace++;
}

// One step of ProMap_jet_refine's loop (promap_jet_label_propagation.h:
// 1254-1305) for the mapping objective: an iteration without moves is
// counted and skipped; the balance counter is raised before every rebalance
// iteration (one weak, then strong); the second best-partition test is
// strict on the max block weight.  The loop ends early where it can only
// repeat itself: after a strong iteration without moves (nothing it reads
// changed) or after two LP iterations in a row without moves (the first
// clears every lock, so the second and all later ones see the same state);
// the remaining iterations of the reference change nothing.
__device__ void map_control(const JetDev &d, JetCtl *ctl, const unsigned long long *rep, s64 maxw) {
    const u32 mode = ctl->mode;
    const u32 n_moves = (u32) rep[ST_MOVES];
    u32 bal = ctl->bal, it = ctl->it;
    u32 save = 0;
    bool done = false;
    if (n_moves) {
        ctl->iters++;
        const s64 cur = ctl->cur_cut + (s64) rep[ST_DELTA];
        ctl->cur_cut = cur;
        ctl->cur_maxw = maxw;
        if (d.trace) trace_record(d, ctl, rep, mode, cur, maxw);
        const s64 best = ctl->best_cut, best_maxw = ctl->best_maxw;
        if (best_maxw > d.lmax && maxw < best_maxw) {
            save = 1;
            ctl->best_cut = cur;
            ctl->best_maxw = maxw;
            it = 0;
        } else if (cur < best && (maxw <= d.lmax || maxw < best_maxw)) {
            if ((double) cur < PHI * (double) best) it = 0;
            save = 1;
            ctl->best_cut = cur;
            ctl->best_maxw = maxw;
        }
        ctl->zero_lp = 0;
    } else if (mode == 0) {
        done = ++ctl->zero_lp >= 2;
    } else {
        done = mode == 2;
    }
    ctl->save = save;
    if (done || it >= N_MAX_ITERATIONS) {
        ctl->active = 0;
    } else {
        it += 1;
        if (ctl->cur_maxw <= d.lmax) {
            bal = 0;
            ctl->mode = 0;
        } else {
            bal++;
            ctl->mode = bal < N_MAX_WEAK_ITERATIONS ? 1u : 2u;
        }
    }
    ctl->it = it;
    ctl->bal = bal;
    ctl->stamp++;
}

// One step of jet_refine's loop (jet:960-1030) after an iteration: counters,
// cut and block weight, best-partition tracking, ratio and mode of the next
// iteration.  Runs on one thread of the control launch.
__device__ __forceinline__ void control_step(const JetDev &d, s64 maxw, const u64 *tot, const JetCtl &in) {
    JetCtl cl = in;  // the control block as loaded by the caller; written back at the end
    JetCtl *ctl = &cl;
    unsigned long long rep[ST_COUNT];
#pragma unroll
    for (u32 s = 0; s < ST_COUNT; ++s) {
        rep[s] = ctl->st[s];
        ctl->st[s] = 0;
    }
    rep[ST_MISS] += tot[PT_MISS];
    rep[ST_MISS_DEG] += tot[PT_MISS_DEG];
    rep[ST_CAND_DEG] += tot[PT_CAND_DEG];
    rep[ST_GHASH] += tot[PT_GHASH];
    rep[ST_DELTA] += tot[PT_DELTA];
    rep[ST_MOVE_DEG] += tot[PT_MOVE_DEG];
    rep[ST_MHASH] += tot[PT_MHASH];
    rep[ST_CAND] += tot[PT_COUNT];
    rep[ST_MOVES] += tot[PT_COUNT + 1];
    const u32 mode = ctl->mode;
    const u32 n_moves = (u32) rep[ST_MOVES];
    ctl->last_moves = n_moves;
    account(d, ctl, rep, mode, ctl->save != 0);
    if (d.D) {
        map_control(d, ctl, rep, maxw);
        *d.ctl = cl;
        return;
    }
    u32 bal = ctl->bal, it = ctl->it, ri = ctl->ri;
    bool counted = true, end_ratio = false;
    if (mode == 0) {
        bal = 0;
        if (n_moves == 0) end_ratio = true;
    } else if (mode == 1) {
        if (n_moves == 0) {
            bal = N_MAX_WEAK_ITERATIONS;
            counted = false;
        } else {
            bal++;
        }
    } else {
        if (n_moves == 0) end_ratio = true;
        else bal++;
    }
    u32 save = 0;
    if (counted && !end_ratio) {
        ctl->iters++;
        const s64 cur_cut = ctl->cur_cut + (s64) rep[ST_DELTA] / 2;
        const s64 cur_maxw = maxw;
        ctl->cur_cut = cur_cut;
        ctl->cur_maxw = cur_maxw;
        if (d.trace) trace_record(d, ctl, rep, mode, cur_cut, cur_maxw);
        const s64 best_cut = ctl->best_cut, best_maxw = ctl->best_maxw;
        if (best_maxw > d.lmax && cur_maxw < best_maxw) {
            save = 1;
            ctl->best_cut = cur_cut;
            ctl->best_maxw = cur_maxw;
            it = 0;
        } else if (cur_cut < best_cut && (cur_maxw <= d.lmax || cur_maxw <= best_maxw)) {
            if ((double) cur_cut < PHI * (double) best_cut) it = 0;
            save = 1;
            ctl->best_cut = cur_cut;
            ctl->best_maxw = cur_maxw;
        }
    }
    ctl->save = save;
    if (end_ratio || it >= N_MAX_ITERATIONS) {
        ri++;
        bal = 0;
        it = 0;
    }
    if (ri >= d.n_ratios) {
        ctl->active = 0;
    } else {
        it += 1;
        ctl->mode = ctl->cur_maxw <= d.lmax ? 0u : (bal < N_MAX_WEAK_ITERATIONS ? 1u : 2u);
        ctl->ratio = d.ratios[ri];
    }
    ctl->ri = ri;
    ctl->it = it;
    ctl->bal = bal;
    ctl->stamp++;
    *d.ctl = cl;
}

// hubs: one workgroup per HUB_CHUNK edges of a moved hub (the chunk table
// lists (hub, chunk) pairs); chunk 0 also commits the move
constexpr u32 HUB_CHUNK = JB * ADJ_UNROLL * 4;
__global__ __launch_bounds__(JB) void apply_hub_kernel(JetDev d, const u32 *hubs, const u32 *chunk_hub,
                                                       const u32 *chunk_idx) {
    __shared__ s64 red[JB / WAVE];
    JetCtl *ctl = d.ctl;
    const u32 h = chunk_hub[blockIdx.x], ci = chunk_idx[blockIdx.x];
    const CtlHead hd = ctl_head(ctl);
    if (!live(hd)) return;
    const u32 stamp = hd.stamp;
    const u32 u = hubs[h];
    if (d.mark[u] != stamp) return;
    const u64 b = (u64) d.xadj[u] + (u64) ci * HUB_CHUNK, end = d.xadj[u + 1];
    const u64 e = b + HUB_CHUNK < end ? b + HUB_CHUNK : end;
    s64 acc = move_delta_range<ADJ_UNROLL>(d, u, b, e, threadIdx.x, JB, stamp);
    acc = block_sum<JB>(acc, red);
    if (threadIdx.x == 0) {
        if (acc) atomicAdd(&ctl->st[ST_DELTA], (unsigned long long) acc);
        if (ci == 0) {
            if (d.count_bytes) atomicAdd(&ctl->st[ST_MOVE_DEG], (unsigned long long) degree(d, u));
            move_commit(d, u);
            if (d.trace) atomicAdd(&ctl->st[ST_MHASH], (unsigned long long) mix64d(((u64) u << 32) | d.dest[u]));
        }
    }
}

__global__ void hub_chunk_count_kernel(const u32 *xadj, const u32 *hubs, u32 n_hubs, u32 *cnt) {
    for (u64 h = (u64) blockIdx.x * blockDim.x + threadIdx.x; h < n_hubs; h += (u64) gridDim.x * blockDim.x) {
        const u32 u = hubs[h];
        cnt[h] = (u32) (((u64) xadj[u + 1] - xadj[u] + HUB_CHUNK - 1) / HUB_CHUNK);
    }
}

__global__ void hub_chunk_fill_kernel(const u32 *cnt, const u32 *off, u32 n_hubs, u32 *chunk_hub, u32 *chunk_idx) {
    for (u64 h = (u64) blockIdx.x * blockDim.x + threadIdx.x; h < n_hubs; h += (u64) gridDim.x * blockDim.x)
        for (u32 c = 0; c < cnt[h]; ++c) {
            chunk_hub[off[h] + c] = (u32) h;
            chunk_idx[off[h] + c] = c;
        }
}

// G lanes per move: this workgroup's move region, then its share of the
// extra list
template<int G>
__device__ __forceinline__ void apply_one(const JetDev &d, u32 u, bool active, u32 lane, u32 stamp, bool do_hash,
                                          s64 &acc, u64 &degs, u64 &mh, s64 *bw) {
    const bool hub = active && degree(d, u) > d.hub_deg;
    if (active && !hub) {
        acc += move_delta<group_unroll<G>()>(d, u, lane, G, stamp);
        if (lane == 0) {
            degs += degree(d, u);
            move_commit(d, u, bw);
            if (do_hash) mh += mix64d(((u64) u << 32) | d.dest[u]);
        }
    }
}

// apply (perform_moves, jet:753-924 / promap:1104-1216): this workgroup's
// move region, then its contiguous share of the extra list, edge-balanced
constexpr u32 BW_LDS = 1024;  // blocks whose weight changes an apply workgroup sums in LDS

// lane-group form (levels without a skewed degree distribution): G lanes per
// move, no staging or prefix
template<int G>
__global__ __launch_bounds__(JB) void apply_groups_kernel(JetDev d) {
    __shared__ u64 s_red[JB / WAVE];
    JetCtl *ctl = d.ctl;
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    const u32 *region = d.moves + (u64) blockIdx.x * d.cap_mv;
    // issued with the control block (one round trip): the region length, the
    // group's first entry (a region of a dense level may hold fewer than
    // `groups` entries) and the extra list's length
    const u32 len = d.moves_n[blockIdx.x];
    const u32 first = gi < d.cap_mv ? region[gi] : 0u;
    const u32 lx = (u32) ctl->st[ST_MOVES];
    const CtlHead hd = ctl_head(ctl);
    if (!live(hd)) return;
    const u32 stamp = hd.stamp;
    const bool do_hash = d.trace != nullptr;
    s64 *bwt = d.bw;  // few moves per workgroup here: direct updates
    s64 acc = 0;
    u64 mh = 0, degs = 0;
    for (u32 i0 = 0; i0 < len; i0 += groups) {
        const u32 i = i0 + gi;
        apply_one<G>(d, i < len ? (i0 == 0 ? first : region[i]) : 0u, i < len, lane, stamp, do_hash, acc, degs, mh,
                     bwt);
    }
    for (u64 i0 = (u64) blockIdx.x * groups; i0 < lx; i0 += (u64) gridDim.x * groups) {
        const u64 i = i0 + gi;
        apply_one<G>(d, i < lx ? d.moves_x[i] : 0u, i < lx, lane, stamp, do_hash, acc, degs, mh, bwt);
    }
    const u64 a = block_total((u64) acc, s_red);
    degs = block_total(degs, s_red);
    mh = block_total(mh, s_red);
    if (threadIdx.x == 0) {
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_DELTA] = a;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MOVE_DEG] = degs;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MHASH] = mh;
    }
}

__global__ __launch_bounds__(JB) void apply_kernel(JetDev d) {
    __shared__ u64 s_red[JB / WAVE];
    __shared__ FlatLds L;
    __shared__ s64 s_bw[BW_LDS];
    JetCtl *ctl = d.ctl;
    const u32 *region = d.moves + (u64) blockIdx.x * d.cap_mv;
    // issued with the control block (one round trip)
    const u32 len = d.moves_n[blockIdx.x];
    const u32 lx = (u32) ctl->st[ST_MOVES];
    const CtlHead hd = ctl_head(ctl);
    if (!live(hd)) return;
    const u32 stamp = hd.stamp;
    const bool do_hash = d.trace != nullptr;
    const u32 xch = (lx + gridDim.x - 1) / gridDim.x;
    const u32 x0 = blockIdx.x * xch < lx ? blockIdx.x * xch : lx;
    const u32 x1 = x0 + xch < lx ? x0 + xch : lx;
    const u32 total = len + (x1 - x0);
    // block weight changes: summed per workgroup in LDS and added once per
    // block (a rebalancing iteration can move millions of vertices between
    // few blocks; per-move updates of those few addresses serialised)
    const bool lds_bw = d.k <= BW_LDS;
    if (lds_bw)
        for (u32 b = threadIdx.x; b < d.k; b += JB) s_bw[b] = 0;
    __syncthreads();
    s64 *bwt = lds_bw ? s_bw : d.bw;
    s64 acc = 0;
    u64 mh = 0, degs = 0;
    for (u32 c0 = 0; c0 < total; c0 += FLAT_CHUNK) {
        const u32 cnt = total - c0 < FLAT_CHUNK ? total - c0 : FLAT_CHUNK;
        const u32 t = threadIdx.x;
        u32 u = 0, deg = 0;
        bool hub = true;
        if (t < cnt) {
            const u32 i = c0 + t;
            u = i < len ? region[i] : d.moves_x[x0 + (i - len)];
            const u32 b = d.xadj[u], e = d.xadj[u + 1];
            hub = e - b > d.hub_deg;
            deg = hub ? 0u : e - b;
            L.u[t] = u;
            L.start[t] = b;
            L.a[t] = d.oldp[u];
            L.b[t] = d.dest[u];
        }
        const u32 E = flat_prefix(L, cnt, deg);
        for (u32 y0 = t; y0 < E; y0 += JB * ADJ_UNROLL) {
            u32 iv[ADJ_UNROLL], v[ADJ_UNROLL], mk[ADJ_UNROLL], op[ADJ_UNROLL], ds[ADJ_UNROLL], wk[ADJ_UNROLL];
            u32 wi[ADJ_UNROLL];
            u8 rf[ADJ_UNROLL];
            s64 w[ADJ_UNROLL];
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                const u32 y = y0 + q * JB;
                const bool ok = y < E;
                iv[q] = ok ? flat_find(L.off, cnt, y) : 0u;
                const u32 j = L.start[iv[q]] + (y - L.off[iv[q]]);
                v[q] = ok ? d.adj[j] : 0u;
                w[q] = ok ? ewt(d, j) : 0;
            }
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                const bool ok = y0 + q * JB < E;
                mk[q] = ok ? d.mark[v[q]] : 0u;
                op[q] = ok ? d.oldp[v[q]] : 0u;
                ds[q] = ok ? d.dest[v[q]] : 0u;
                wk[q] = ok ? d.work[v[q]] : 0u;
                rf[q] = ok && d.rflag ? d.rflag[v[q]] : (u8) 0;
                wi[q] = ok && d.wcnt ? d.wid[v[q]] : 0u;
            }
#pragma unroll
            for (u32 q = 0; q < ADJ_UNROLL; ++q) {
                if (y0 + q * JB >= E) continue;
                const u32 ou = L.a[iv[q]], nu = L.b[iv[q]];
                const bool mv = mk[q] == stamp;
                const u32 ov = mv ? op[q] : wk[q];
                const u32 nv = mv ? ds[q] : ov;
                if (d.D) {  // promap:1174-1207
                    if (ou != nu) {
                        const s64 x = w[q] * ((s64) d.D[(u64) nu * d.k + nv] - (s64) d.D[(u64) ou * d.k + ov]);
                        acc += nv != ov ? x : 2 * x;
                    }
                } else {
                    const s64 x = w[q] * ((s64) (nu != nv) - (s64) (ou != ov));
                    acc += mv ? x : 2 * x;
                }
                stale_of(d)[v[q]] = 3;
                row_touch(d, v[q], rf[q], wi[q], ou, nu);
            }
        }
        if (t < cnt && !hub) {
            degs += deg;
            move_commit(d, u, bwt);
            if (do_hash) mh += mix64d(((u64) u << 32) | L.b[t]);
        }
        __syncthreads();  // the next chunk restages the LDS
    }
    if (lds_bw)
        for (u32 b = threadIdx.x; b < d.k; b += JB)
            if (s_bw[b]) atomicAdd((unsigned long long *) &d.bw[b], (unsigned long long) s_bw[b]);
    const u64 a = block_total((u64) acc, s_red);
    degs = block_total(degs, s_red);
    mh = block_total(mh, s_red);
    if (threadIdx.x == 0) {
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_DELTA] = a;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MOVE_DEG] = degs;
        d.pstat[(u64) blockIdx.x * PT_COUNT + PT_MHASH] = mh;
    }
}

// one workgroup: the largest block weight and the sums of the per-workgroup
// partials (one barrier), then control_step.  The control block is copied to
// LDS with the first loads, so neither the live check nor control_step adds
// a memory round trip.
__global__ __launch_bounds__(JB) void control_kernel(JetDev d) {
    constexpr u32 NV = PT_COUNT + 3;  // partials, two list lengths, max block weight
    constexpr u32 CW = (sizeof(JetCtl) + 7) / 8;
    static_assert(CW <= JB && sizeof(JetCtl) % 8 == 0, "control block copy");
    __shared__ u64 s_part[JB / WAVE][NV];
    __shared__ u64 s_ctl[CW];
    if (threadIdx.x < CW) s_ctl[threadIdx.x] = ((const u64 *) d.ctl)[threadIdx.x];
    u64 t[NV];
#pragma unroll
    for (u32 j = 0; j < NV; ++j) t[j] = 0;
    s64 mx = INT64_MIN;
    for (u32 b = threadIdx.x; b < d.k; b += JB) mx = d.bw[b] > mx ? d.bw[b] : mx;
    for (u32 r = threadIdx.x; r < d.nreg; r += JB) {
#pragma unroll
        for (u32 j = 0; j < PT_COUNT; ++j) t[j] += d.pstat[(u64) r * PT_COUNT + j];
        t[PT_COUNT] += d.cand_n[r];
        t[PT_COUNT + 1] += d.moves_n[r];
    }
#pragma unroll
    for (u32 j = 0; j < NV - 1; ++j) t[j] = wave_sum(t[j]);
    mx = wave_max(mx);
    if (lane_id() == 0) {
#pragma unroll
        for (u32 j = 0; j < NV - 1; ++j) s_part[threadIdx.x / WAVE][j] = t[j];
        s_part[threadIdx.x / WAVE][NV - 1] = (u64) mx;
    }
    __syncthreads();
    const JetCtl &in = *(const JetCtl *) s_ctl;
    if (threadIdx.x == 0 && in.active && !in.stalled) {
        s64 m = INT64_MIN;
#pragma unroll
        for (u32 j = 0; j < NV - 1; ++j) t[j] = 0;
        for (u32 w = 0; w < JB / WAVE; ++w) {
#pragma unroll
            for (u32 j = 0; j < NV - 1; ++j) t[j] += s_part[w][j];
            m = (s64) s_part[w][NV - 1] > m ? (s64) s_part[w][NV - 1] : m;
        }
        control_step(d, m, t, in);
    }
}

// ---------------------------------------------------------------- BlockConn rows of wide vertices
// update_small / update_large (block_conn.h:208-466) reduced to what a
// decision reads.  For a wide vertex with a neighbour in the move list, P =
// the blocks of its neighbours after the moves (every edge weight is
// positive, so these are its positive entries), and
//   update_large (a move list longer than n / 10, jet:853): the row is
//     rebuilt; needed = |P|, size = needed + max(3, needed / 4) if that is
//     below k, else k; valid entries = P;
//   update_small: an entry whose weight falls to 0 becomes a tombstone unless
//     the size is k; a new block takes a tombstone or an empty slot inside the
//     size before the size grows, so size = max(size, |P|); valid entries = P,
//     or, in a row of size k, the old valid entries plus P.
// init_BlockConn (:48-196): size = |P|, valid entries = P.
enum RowMode : u32 { ROW_SMALL = 0, ROW_LARGE = 1, ROW_INIT = 2 };

template<int G>
__device__ __forceinline__ void group_sync() {
    if constexpr (G <= WAVE) wave_sync();
    else __syncthreads();
}

// one wide vertex u per group of G lanes; wm: the group's W LDS words.  The
// group walks u's adjacency (at init counting a counted vertex's neighbours
// per block in cl, the group's k LDS counters), or, after init, reads a
// counted vertex's counts.  A walk that needs no counts ends once every
// block has been seen (P is then all k blocks): a hub of an HM split
// (k <= 8) sees them all in its first rounds.
template<int G>
__device__ __forceinline__ void row_update(const JetDev &d, u32 *wm, u32 u, u32 lane, u32 rmode, bool active,
                                           u32 *red, u32 *cl = nullptr) {
    const bool counted = active && (d.rflag[u] & 4u);
    // the same branch for every group of a launch: at init all walk, after
    // init the counted vertices' groups read counts (a hub's workgroup and,
    // for k > 64, every group walk: none of them is counted)
    if (rmode == ROW_INIT || !counted) {
        if (active) {
            for (u32 i = lane; i < d.W; i += G) wm[i] = 0;
            if (counted)
                for (u32 b = lane; b < d.k; b += G) cl[b] = 0;
        }
        group_sync<G>();
        const bool can_stop = d.W <= 2 && !counted;
        if (active) {
            const u64 e = d.xadj[u + 1];
            // the loop runs on the group's base, so every lane of the group
            // runs the same rounds and meets the same barriers
            for (u64 j00 = d.xadj[u]; j00 < e; j00 += (u64) G * ADJ_UNROLL) {
                const u64 j0 = j00 + lane;
                u32 v[ADJ_UNROLL];
#pragma unroll
                for (u32 q = 0; q < ADJ_UNROLL; ++q) v[q] = j0 + (u64) q * G < e ? d.adj[j0 + (u64) q * G] : 0u;
#pragma unroll
                for (u32 q = 0; q < ADJ_UNROLL; ++q)
                    if (j0 + (u64) q * G < e) {
                        const u32 b = d.work[v[q]];
                        atomicOr(&wm[b >> 5], 1u << (b & 31));
                        if (counted) atomicAdd(&cl[b], 1u);
                    }
                if (can_stop) {
                    group_sync<G>();  // the group's bits of this round
                    const u32 seen = (u32) __popc(wm[0]) + (d.W > 1 ? (u32) __popc(wm[1]) : 0u);
                    group_sync<G>();  // every lane has read them before the next round adds any
                    if (seen == d.k) break;
                }
            }
        }
        group_sync<G>();
        if (counted) {
            const u64 base = (u64) d.wid[u] * d.k;
            for (u32 b = lane; b < d.k; b += G) d.wcnt[base + b] = cl[b];
        }
    } else if (lane == 0) {
        const u64 mask = row_count_mask(d, u);  // W <= 2
        wm[0] = (u32) mask;
        if (d.W > 1) wm[1] = (u32) (mask >> 32);
    }
    group_sync<G>();
    u32 pc = 0;
    if (active)
        for (u32 i = lane; i < d.W; i += G) pc += (u32) __popc(wm[i]);
    pc = group_sum<G>(pc, red);
    if (active) {
        const u32 r = d.wid[u];
        const u32 sz = d.wsz[r];
        const bool keep = rmode == ROW_SMALL && sz == d.k;
        for (u32 i = lane; i < d.W; i += G) {
            const u64 o = (u64) r * d.W + i;
            const u32 P = wm[i];
            d.wV[o] = keep ? (d.wV[o] | P) : P;
            d.wP[o] = P;
        }
        if (lane == 0) {
            u32 ns;
            if (rmode == ROW_INIT) {
                ns = pc;
            } else if (rmode == ROW_LARGE) {
                ns = pc + (pc / 4 < 3 ? 3 : pc / 4);
                if (ns >= d.k) ns = d.k;
            } else {
                ns = sz > pc ? sz : pc;
            }
            d.wsz[r] = ns;
            d.rflag[u] &= (u8) ~2u;
        }
    }
    group_sync<G>();
}

// Wide vertices of k <= 64 that are not hubs, one thread each, the block
// mask in a register: those of degree <= ROW_THREAD_DEG walk their adjacency,
// the counted ones read their counts (at init the row kernel counts them).
// On skewed graphs with small k most vertices are wide and a move list
// touches most of them; a lane group per vertex spent its time in the
// group's fixed costs, one vertex after another.
__global__ __launch_bounds__(JB) void row_thread_kernel(JetDev d, u32 init) {
    u32 rmode = ROW_INIT;
    if (!init) {
        const CtlHead hd = ctl_head(d.ctl);
        if (!live(hd)) return;
        rmode = d.ctl->last_moves > d.n / 10 ? ROW_LARGE : ROW_SMALL;
    }
    for (u64 v = (u64) blockIdx.x * JB + threadIdx.x; v < d.n; v += (u64) gridDim.x * JB) {
        const u8 f = d.rflag[v];
        if ((f & 3u) != 3u) continue;
        u64 mask = 0;
        if (f & 4u) {
            if (init) continue;  // the row kernel counts them
            mask = row_count_mask(d, (u32) v);
        } else {
            const u32 xb = d.xadj[v], xe = d.xadj[v + 1];
            if (xe - xb > ROW_THREAD_DEG) continue;  // a hub
            for (u32 j = xb; j < xe; j += 4) {
                u32 nb[4];
#pragma unroll
                for (u32 q = 0; q < 4; ++q) nb[q] = j + q < xe ? d.adj[j + q] : 0u;
#pragma unroll
                for (u32 q = 0; q < 4; ++q)
                    if (j + q < xe) mask |= 1ull << d.work[nb[q]];
            }
        }
        const u32 pc = (u32) __popcll(mask);
        const u32 r = d.wid[v];
        const u32 sz = d.wsz[r];
        const bool keep = rmode == ROW_SMALL && sz == d.k;
        for (u32 i = 0; i < d.W; ++i) {
            const u64 o = (u64) r * d.W + i;
            const u32 P = (u32) (mask >> (32 * i));
            d.wV[o] = keep ? (d.wV[o] | P) : P;
            d.wP[o] = P;
        }
        u32 ns;
        if (rmode == ROW_INIT) {
            ns = pc;
        } else if (rmode == ROW_LARGE) {
            ns = pc + (pc / 4 < 3 ? 3 : pc / 4);
            if (ns >= d.k) ns = d.k;
        } else {
            ns = sz > pc ? sz : pc;
        }
        d.wsz[r] = ns;
        d.rflag[v] = f & (u8) ~2u;
    }
}

// the flagged wide vertices that are not hubs, G lanes each; init: every
// wide vertex (flagged by the set-up), before the loop starts
template<int G>
__global__ __launch_bounds__(JB) void row_kernel(JetDev d, u32 init, u32 piece) {
    extern __shared__ __align__(16) u32 rsm[];
    __shared__ u32 s_nlist;
    u32 rmode = ROW_INIT;
    if (!init) {
        const CtlHead hd = ctl_head(d.ctl);
        if (!live(hd)) return;
        rmode = d.ctl->last_moves > d.n / 10 ? ROW_LARGE : ROW_SMALL;
    }
    const u32 groups = JB / G, lane = threadIdx.x % G, gi = threadIdx.x / G;
    u32 *list = rsm;
    u32 *wm = rsm + SWEEP_PIECE + gi * d.W;
    // the counted vertices' counts at init: k LDS counters per group (k <= 64)
    u32 *cl = init && d.W <= 2 ? rsm + SWEEP_PIECE + groups * d.W + gi * d.k : nullptr;
    for (u64 p0 = (u64) blockIdx.x * piece; p0 < d.n; p0 += (u64) gridDim.x * piece) {
        const u64 p1 = p0 + piece < d.n ? p0 + piece : d.n;
        if (threadIdx.x == 0) s_nlist = 0;
        __syncthreads();
        for (u64 v0 = p0; v0 < p1; v0 += JB) {
            const u64 vv = v0 + threadIdx.x;
            bool take = false;
            // k <= 64: the thread kernel takes the rest (the counted ones after init)
            if (vv < p1 && (d.rflag[vv] & 3u) == 3u) {
                const u32 dg = degree(d, (u32) vv);
                take = dg <= d.hub_deg && (d.W > 2 || (init && dg > ROW_THREAD_DEG));
            }
            append_local(take, (u32) vv, list, &s_nlist);
        }
        __syncthreads();
        const u32 nl = s_nlist;
        for (u32 r0 = 0; r0 < nl; r0 += groups) {
            const u32 idx = r0 + gi;
            row_update<G>(d, wm, idx < nl ? list[idx] : 0u, lane, rmode, idx < nl, nullptr, cl);
        }
        __syncthreads();
    }
}

// a flagged wide hub: one workgroup
__global__ __launch_bounds__(JB) void row_hub_kernel(JetDev d, const u32 *hubs, u32 init) {
    extern __shared__ __align__(16) u32 rsm[];
    __shared__ u32 red[JB / WAVE];
    u32 rmode = ROW_INIT;
    if (!init) {
        const CtlHead hd = ctl_head(d.ctl);
        if (!live(hd)) return;
        rmode = d.ctl->last_moves > d.n / 10 ? ROW_LARGE : ROW_SMALL;
    }
    const u32 u = hubs[blockIdx.x];
    if ((d.rflag[u] & 3u) != 3u) return;
    row_update<JB>(d, rsm, u, threadIdx.x, rmode, true, red);
}

// ---------------------------------------------------------------- single-workgroup form
// Small levels: one workgroup of BM threads runs whole iterations (sweep,
// minibucket prefixes, afterburner or classify, resolve, apply, control)
// with barriers where the multi-launch form has launch boundaries, up to
// max_iters iterations per launch.  Same device functions, same results.
// Lists: candidates in d.cand, moves in the extra move list, resolve entries
// in d.rlist, all counted in LDS.  A resolve list that does not fit stalls
// the loop: the iteration's lengths and statistics are published in st[] and
// region counts are zero, so the host's global path and the multi-launch
// apply and control finish that iteration.
constexpr int BM = 512;  // 256 registers per lane (1024 threads cap them at 128 and spill)
constexpr u32 BM_PIECE = 2048;
constexpr u32 BM_RCAP = 2048;
constexpr size_t BM_CONN_LDS = 40 * 1024;

__host__ __device__ inline size_t bm_sweep_lds_bytes(u32 k, u32 groups) {
    return conn_group_bytes(k) * groups + 4ull * BM_PIECE + 16 + 4ull * k;
}

// LDS copy of the per-vertex state of the single-workgroup form: s64
// gcache, gain1, icache (n each) and the k block weights, then u32 work,
// dest, dcache, mark, oldp, gid (n each), then the u8 locks and stale bytes
__host__ __device__ constexpr size_t bm_state_bytes(u32 n, u32 k) { return 50ull * n + 8ull * k; }
// the u32 row ids, then the u8 row flags
__host__ __device__ constexpr size_t bm_rows_bytes(u32 n) { return 5ull * n; }

// LS: the per-vertex state lives in LDS (from byte state_off of the dynamic
// LDS) for the whole launch: copied in at the start and back to device
// memory before the kernel returns (the host's global resolve path and the
// next launch read it there).  The device functions reach it through the
// same pointers in a local copy of the Jet state, as generic addresses.
// The control block is kept in LDS the same way in both forms.  With
// rows_off (LS only, 0: off) the BlockConn row ids and flags of the level's
// vertices follow at that byte: the apply and row steps read them for every
// neighbour of a move; the flags go back to device memory at the end.
template<int G, bool LS>
__global__ __launch_bounds__(BM) void jet_block_kernel(JetDev dg, u32 max_iters, u32 groups, u32 state_off,
                                                       u32 rows_off) {
    extern __shared__ __align__(16) char smem[];
    constexpr u32 CW = (sizeof(JetCtl) + 7) / 8;
    static_assert(CW <= BM && sizeof(JetCtl) % 8 == 0, "control block copy");
    __shared__ u64 s_ctl[CW];
    JetDev d = dg;
    if (threadIdx.x < CW) s_ctl[threadIdx.x] = ((const u64 *) dg.ctl)[threadIdx.x];
    d.ctl = (JetCtl *) s_ctl;
    if constexpr (LS) {
        const u32 n = dg.n, k = dg.k;
        s64 *gc8 = (s64 *) (smem + state_off);
        d.gcache = gc8;
        d.gain1 = gc8 + n;
        d.icache = gc8 + 2ull * n;
        d.bw = gc8 + 3ull * n;
        u32 *w4 = (u32 *) (gc8 + 3ull * n + k);
        d.work = w4;
        d.dest = w4 + n;
        d.dcache = w4 + 2ull * n;
        d.mark = w4 + 3ull * n;
        d.oldp = w4 + 4ull * n;
        d.gid = w4 + 5ull * n;
        d.lock = (u8 *) (w4 + 6ull * n);
        for (u32 v = threadIdx.x; v < n; v += BM) {
            d.gcache[v] = dg.gcache[v];
            d.gain1[v] = dg.gain1[v];
            d.icache[v] = dg.icache[v];
            d.work[v] = dg.work[v];
            d.dest[v] = dg.dest[v];
            d.dcache[v] = dg.dcache[v];
            d.mark[v] = dg.mark[v];
            d.oldp[v] = dg.oldp[v];
            d.gid[v] = dg.gid[v];
            d.lock[v] = dg.lock[v];
            d.lock[n + v] = dg.lock[n + v];  // stale bytes
        }
        for (u32 b = threadIdx.x; b < k; b += BM) d.bw[b] = dg.bw[b];
        if (rows_off) {
            u32 *wid = (u32 *) (smem + rows_off);
            u8 *rf = (u8 *) (wid + n);
            for (u32 v = threadIdx.x; v < n; v += BM) {
                wid[v] = dg.wid[v];
                rf[v] = dg.rflag[v];
            }
            d.wid = wid;
            d.rflag = rf;
        }
    }
    __syncthreads();
    __shared__ u32 s_live, s_mode, s_stamp, s_save;
    __shared__ double s_ratio;
    __shared__ u32 s_ncand, s_nlist, s_nmoves, s_nrl;
    __shared__ u64 s_acc[PT_COUNT];
    __shared__ u64 s_red[BM / WAVE];
    JetCtl *ctl = d.ctl;
    const u32 tid = threadIdx.x, lane = tid % G, gi = tid / G;
    const bool ingroup = gi < groups;
    const size_t cgb = conn_group_bytes(d.k);
    GroupConn gc = group_conn(smem, d.k, ingroup ? gi : 0);
    u32 *list = (u32 *) (smem + cgb * groups);
    RebShared *rs = (RebShared *) (smem + cgb * groups + 4ull * BM_PIECE);
    const ResolveLds rl = resolve_lds(smem, BM_RCAP);
    for (u32 r = tid; r < d.nreg; r += BM) {
        d.cand_n[r] = 0;
        d.moves_n[r] = 0;
    }
    for (u64 j = tid; j < (u64) d.nreg * PT_COUNT; j += BM) d.pstat[j] = 0;
    for (u32 iter = 0; iter < max_iters; ++iter) {
        if (tid == 0) {
            const CtlHead hd = ctl_head(ctl);
            s_live = live(hd);
            s_mode = hd.mode;
            s_stamp = hd.stamp;
            s_save = hd.save;
            s_ratio = hd.ratio;
            s_ncand = s_nmoves = s_nrl = 0;
            for (u32 j = 0; j < PT_COUNT; ++j) s_acc[j] = 0;
        }
        __syncthreads();
        if (!s_live) break;
        const u32 mode = s_mode, stamp = s_stamp;
        const bool save = s_save != 0, strong = mode == 2;
        const double ratio = s_ratio;
        const s64 mbw = mode ? max_block_weight(d, strong) : 0;
        // ---- sweep
        if (ingroup) {
            for (u32 i = lane; i < d.k; i += G) gc.cnt[i] = 0;
            if (lane == 0) *gc.tcount = 0;
        }
        if (mode) reb_prologue(d, strong, mbw, rs);
        __syncthreads();
        u64 nev = 0, degs = 0;
        for (u64 p0 = 0; p0 < d.n; p0 += BM_PIECE) {
            const u64 p1 = p0 + BM_PIECE < d.n ? p0 + BM_PIECE : d.n;
            if (tid == 0) s_nlist = 0;
            __syncthreads();
            for (u64 v0 = p0; v0 < p1; v0 += BM) {
                const u64 vv = v0 + tid;
                const u32 v = (u32) vv;
                bool heavy = false, cand = false;
                if (vv < p1) {
                    if (save) d.part[v] = d.work[v];
                    if (mode == 0) {
                        const u32 dc0 = d.dcache[v];
                        const u32 dc = (stale_of(d)[v] & 1u) ? NULL_PART : dc0;
                        if (dc == NULL_PART) {
                            heavy = true;
                        } else {
                            d.dest[v] = dc;
                            if (dc != NO_MOVE && d.lock[v] == 0) {
                                cand = true;
                                d.gain1[v] = d.gcache[v];
                            } else {
                                d.gain1[v] = GAIN_MIN;
                                d.lock[v] = 0;
                            }
                        }
                    } else if (reb_eligible(d, strong, rs, v, d.work[v])) {
                        const s64 ic0 = d.icache[v];
                        const s64 ic = (stale_of(d)[v] & 2u) ? -1 : ic0;
                        if (ic >= 0) cand = reb_interior(d, strong, rs, v, ic);
                        else heavy = true;
                    }
                }
                append_local(cand, v, d.cand, &s_ncand);
                append_local(heavy, v, list, &s_nlist);
            }
            __syncthreads();
            const u32 nl = s_nlist;
            // each round loads the adjacency offsets of the group's vertex of
            // the next round, so that they arrive with this round's edges
            u32 nxu = 0, nxb = 0, nxe = 0;
            if (ingroup && gi < nl) {
                nxu = list[gi];
                nxb = d.xadj[nxu];
                nxe = d.xadj[nxu + 1];
            }
            for (u32 r0 = 0; r0 < nl; r0 += groups) {
                const u32 idx = r0 + gi;
                const bool active = ingroup && idx < nl;
                const u32 u = active ? nxu : 0, ub = nxb, ue = nxe;
                if (ingroup && idx + groups < nl) {
                    nxu = list[idx + groups];
                    nxb = d.xadj[nxu];
                    nxe = d.xadj[nxu + 1];
                }
                if (active) {
                    if (lane == 0) {
                        nev++;
                        degs += ue - ub;
                    }
                    conn_accumulate_range<group_unroll<G>()>(d, gc, ub, ue, lane, G);
                }
                wave_sync();
                const u32 own = active ? d.work[u] : 0;
                const bool locked = active && d.lock[u] != 0;
                const bool listed = mode == 0
                                        ? lp_eval<G>(d, gc, u, own, locked, lane, ratio, nullptr, nullptr, active)
                                        : reb_eval<G>(d, strong, mbw, rs, gc, u, own, lane, nullptr, nullptr, active);
                wave_sync();
                if (active) conn_reset(gc, *gc.tcount, lane, G);
                wave_sync();
                if (active && lane == 0) *gc.tcount = 0;
                wave_sync();
                append_local(listed, u, d.cand, &s_ncand);
            }
            __syncthreads();
        }
        nev = block_total(nev, s_red);
        degs = block_total(degs, s_red);
        if (tid == 0) {
            s_acc[PT_MISS] = nev;
            s_acc[PT_MISS_DEG] = degs;
        }
        // ---- minibucket prefixes, one wave per block range
        if (mode && d.width <= WAVE) {
            // one chunk per block: a wave loads PU blocks' totals before
            // scanning any, so the loads overlap
            constexpr u32 NW = BM / WAVE, PU = 8;
            const u32 i = lane_id();
            for (u32 p0 = tid / WAVE; p0 < d.k; p0 += PU * NW) {
                s64 v[PU];
#pragma unroll
                for (u32 j = 0; j < PU; ++j) {
                    const u32 p = p0 + j * NW;
                    v[j] = p < d.k && i < d.width ? d.btot[(u64) p * d.width + i] : 0;
                }
#pragma unroll
                for (u32 j = 0; j < PU; ++j) {
                    const u32 p = p0 + j * NW;
                    if (p >= d.k) break;  // uniform in the wave
                    const s64 inc = wave_inclusive_scan(v[j]);
                    const u64 base = (u64) p * d.width;
                    if (i < d.width) {
                        d.Q[base + i] = inc - v[j];
                        d.R[base + i] = inc;
                        d.btot[base + i] = 0;
                    }
                    if (i == WAVE - 1) d.Tseg[p] = inc;  // lanes past the width add 0
                    if (i == 0 && p == d.k - 1) d.btot[(u64) d.k * d.width] = 0;
                }
            }
        } else if (mode) {
            for (u32 p = tid / WAVE; p < d.k; p += BM / WAVE) {
                const u64 base = (u64) p * d.width;
                s64 carry = 0;
                for (u32 c0 = 0; c0 < d.width; c0 += WAVE) {
                    const u32 i = c0 + lane_id();
                    const s64 v = i < d.width ? d.btot[base + i] : 0;
                    const s64 inc = wave_inclusive_scan(v);
                    if (i < d.width) {
                        d.Q[base + i] = carry + inc - v;
                        d.R[base + i] = carry + inc;
                        d.btot[base + i] = 0;
                    }
                    carry += __shfl(inc, WAVE - 1, WAVE);
                }
                if (lane_id() == 0) {
                    d.Tseg[p] = carry;
                    if (p == d.k - 1) d.btot[(u64) d.k * d.width] = 0;
                }
            }
        }
        __syncthreads();
        // ---- afterburner / classify
        const u32 nc = s_ncand;
        if (mode == 0) {
            u64 cdeg = 0, gh = 0;
            afterburner_list<G>(d, stamp, d.cand, nc, d.moves_x, &s_nmoves, cdeg, gh);
            cdeg = block_total(cdeg, s_red);
            gh = block_total(gh, s_red);
            if (tid == 0) {
                s_acc[PT_CAND_DEG] = cdeg;
                s_acc[PT_GHASH] = gh;
            }
        } else {
            for (u32 i0 = 0; i0 < nc; i0 += BM) {
                const u32 i = i0 + tid;
                const bool valid = i < nc;
                const u32 u = valid ? d.cand[i] : 0u;
                bool sure = false, boundary = false;
                u64 key = 0;
                if (valid) {
                    const u32 g = d.gid[u], ub = d.work[u];
                    const s64 limit = d.bw[ub] - d.lmax;
                    if (bucket_prefix(d, g, ub) < limit) {
                        if (!strong && d.R[g] < limit) {
                            sure = true;
                            d.mark[u] = stamp;
                            d.oldp[u] = ub;
                        } else {
                            boundary = true;
                            key = ((u64) g << 32) | u;
                        }
                    }
                }
                append_local(sure, u, d.moves_x, &s_nmoves);
                // rlist through a 32-bit LDS counter: entries are written by position
                const u64 bal = __ballot(boundary);
                if (bal) {
                    const u32 first = (u32) __builtin_ctzll(bal);
                    u32 base = 0;
                    if (lane_id() == first) base = atomicAdd(&s_nrl, (u32) __popcll(bal));
                    base = __shfl(base, (int) first, WAVE);
                    if (boundary) d.rlist[base + (u32) __popcll(bal & lanemask_lt())] = key;
                }
            }
        }
        __syncthreads();
        // ---- resolve
        const u32 L = s_nrl;
        if (mode) {
            const bool ok = resolve_body<BM_RCAP, BM>(d, strong, stamp, L, rl, &s_nmoves);
            if (!ok) {
                if (tid == 0) {
                    ctl->st[ST_MISS] = s_acc[PT_MISS];
                    ctl->st[ST_MISS_DEG] = s_acc[PT_MISS_DEG];
                    ctl->st[ST_CAND] = nc;
                    ctl->st[ST_MOVES] = s_nmoves;
                    ctl->st[ST_RLIST] = L;
                    ctl->stalled = 1;
                }
                break;
            }
        }
        __syncthreads();
        // ---- apply
        const u32 nm = s_nmoves;
        s64 acc = 0;
        u64 mdeg = 0, mh = 0;
        for (u32 i0 = 0; i0 < nm; i0 += BM / G) {
            const u32 i = i0 + gi;
            apply_one<G>(d, i < nm ? d.moves_x[i] : 0u, i < nm, lane, stamp, d.trace != nullptr, acc, mdeg, mh, d.bw);
        }
        const u64 a = block_total((u64) acc, s_red);
        mdeg = block_total(mdeg, s_red);
        mh = block_total(mh, s_red);
        __syncthreads();
        // ---- control (the barrier above completes the block weight updates)
        s64 mx = INT64_MIN;
        for (u32 b = tid; b < d.k; b += BM) mx = d.bw[b] > mx ? d.bw[b] : mx;
        mx = wave_max(mx);
        if (lane_id() == 0) ((s64 *) s_red)[tid / WAVE] = mx;
        __syncthreads();
        if (tid == 0) {
            s64 m = INT64_MIN;
            for (u32 w = 0; w < BM / WAVE; ++w) m = ((s64 *) s_red)[w] > m ? ((s64 *) s_red)[w] : m;
            u64 t[PT_COUNT + 2];
            for (u32 j = 0; j < PT_COUNT; ++j) t[j] = s_acc[j];
            t[PT_DELTA] = a;
            t[PT_MOVE_DEG] = mdeg;
            t[PT_MHASH] = mh;
            t[PT_COUNT] = nc;
            t[PT_COUNT + 1] = nm;
            ctl->st[ST_RLIST] = L;
            control_step(d, m, t, *d.ctl);
        }
        __syncthreads();
        // ---- rows of the wide vertices next to the moves (the group's
        // connectivity counters, all zero between sweeps, hold its mask words)
        if (d.rflag && nm) {
            const u32 rmode = nm > d.n / 10 ? ROW_LARGE : ROW_SMALL;
            u32 *wm = (u32 *) gc.cnt;
            for (u32 p0 = 0; p0 < d.n; p0 += BM_PIECE) {
                const u32 p1 = p0 + BM_PIECE < d.n ? p0 + BM_PIECE : d.n;
                if (tid == 0) s_nlist = 0;
                __syncthreads();
                for (u32 v0 = p0; v0 < p1; v0 += BM) {
                    const u32 v = v0 + tid;
                    append_local(v < p1 && (d.rflag[v] & 3u) == 3u, v, list, &s_nlist);
                }
                __syncthreads();
                const u32 nl = s_nlist;
                for (u32 r0 = 0; r0 < nl; r0 += groups) {
                    const u32 idx = r0 + gi;
                    const bool act = ingroup && idx < nl;
                    row_update<G>(d, wm, act ? list[idx] : 0u, lane, rmode, act, nullptr);
                }
                __syncthreads();
            }
            if (ingroup)
                for (u32 i = lane; i < d.W; i += G) wm[i] = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    if (threadIdx.x < CW) ((u64 *) dg.ctl)[threadIdx.x] = s_ctl[threadIdx.x];
    if constexpr (LS) {
        for (u32 v = threadIdx.x; v < dg.n; v += BM) {
            dg.gcache[v] = d.gcache[v];
            dg.gain1[v] = d.gain1[v];
            dg.icache[v] = d.icache[v];
            dg.work[v] = d.work[v];
            dg.dest[v] = d.dest[v];
            dg.dcache[v] = d.dcache[v];
            dg.mark[v] = d.mark[v];
            dg.oldp[v] = d.oldp[v];
            dg.gid[v] = d.gid[v];
            dg.lock[v] = d.lock[v];
            dg.lock[dg.n + v] = d.lock[dg.n + v];
            if (rows_off) dg.rflag[v] = d.rflag[v];
        }
        for (u32 b = threadIdx.x; b < dg.k; b += BM) dg.bw[b] = d.bw[b];
    }
}

// LDS the single-workgroup kernel declares itself (its __shared__
// variables), which the hardware adds to the dynamic size of a launch
template<int G, bool LS>
size_t jet_block_static_lds() {
    hipFuncAttributes fa{};
    HIP_CHECK(hipFuncGetAttributes(&fa, (const void *) jet_block_kernel<G, LS>));
    return fa.sharedSizeBytes;
}
inline size_t jet_block_static_lds_of(u32 G, bool ls);
inline size_t jet_block_static_lds(u32 G, bool ls) {
    return std::max<size_t>(1024, jet_block_static_lds_of(G, ls));  // at least 1 KiB of headroom
}
inline size_t jet_block_static_lds_of(u32 G, bool ls) {
    switch (G) {
        case 4: return ls ? jet_block_static_lds<4, true>() : jet_block_static_lds<4, false>();
        case 8: return ls ? jet_block_static_lds<8, true>() : jet_block_static_lds<8, false>();
        case 16: return ls ? jet_block_static_lds<16, true>() : jet_block_static_lds<16, false>();
        case 32: return ls ? jet_block_static_lds<32, true>() : jet_block_static_lds<32, false>();
        default: return ls ? jet_block_static_lds<64, true>() : jet_block_static_lds<64, false>();
    }
}

struct RatioTable {
    double r[MAX_RATIOS];
    u32 n;
};

__global__ void ctl_init_kernel(JetDev d, RatioTable rt, s64 cut, s64 maxw) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    JetCtl *ctl = d.ctl;
#pragma unroll
    for (u32 s = 0; s < ST_COUNT; ++s) ctl->st[s] = 0;
    ctl->cur_cut = ctl->best_cut = cut;
    ctl->cur_maxw = ctl->best_maxw = maxw;
    ctl->ratio = rt.r[0];
#pragma unroll
    for (u32 i = 0; i < 5; ++i) ctl->bytes[i] = 0.0;
    ctl->ri = 0;
    ctl->it = 1;
    // the mapping loop raises its balance counter before a rebalance iteration
    ctl->bal = (d.D && maxw > d.lmax) ? 1u : 0u;
    ctl->mode = maxw <= d.lmax ? 0u : 1u;
    ctl->zero_lp = 0;
    ctl->active = 1;
    ctl->stalled = 0;
    ctl->save = 0;
    ctl->iters = 0;
    ctl->n_trace = 0;
    ctl->stamp = 1;
    double *ratios = (double *) d.ratios;
#pragma unroll
    for (u32 i = 0; i < MAX_RATIOS; ++i) ratios[i] = rt.r[i];
}

// ---------------------------------------------------------------- global resolve path
// For resolve lists longer than the LDS capacity: a device-wide radix sort by
// (minibucket, id), offsets from a scan of the sorted weights, then the same
// selection, cookie cutter and adjust_scores as resolve_kernel.
__global__ void fb_vals_kernel(const u64 *keys, u32 n, u32 *vals) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64) gridDim.x * blockDim.x)
        vals[i] = (u32) keys[i];
}

struct WeightSrc {
    const u32 *sv;
    const s32 *vw;
    __device__ s64 operator()(u64 i) const { return vw ? (s64) vw[sv[i]] : 1; }
};
struct StoreS {
    s64 *S;
    __device__ void operator()(u64 i, s64 v) const { S[i] = v; }
};

// prefix at the first entry of every minibucket present
__global__ void fb_head_kernel(const u64 *sk, const s64 *S, u32 n, s64 *H) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64) gridDim.x * blockDim.x) {
        const u32 g = (u32) (sk[i] >> 32);
        if (i == 0 || (u32) (sk[i - 1] >> 32) != g) H[g] = S[i];
    }
}

__global__ void fb_select_kernel(JetDev d, const u64 *sk, const s64 *S, const s64 *H, u32 n, s64 *roff,
                                 int strong) {
    JetCtl *ctl = d.ctl;
    const u32 stamp = ctl->stamp;
    for (u64 i0 = (u64) blockIdx.x * blockDim.x; i0 < n; i0 += (u64) gridDim.x * blockDim.x) {
        const u64 i = i0 + threadIdx.x;
        bool sel = false;
        u32 u = 0;
        if (i < n) {
            u = (u32) sk[i];
            const u32 g = (u32) (sk[i] >> 32);
            const u32 ub = d.work[u];
            const s64 score = bucket_prefix(d, g, ub) + S[i] - H[g];
            const s64 limit = d.bw[ub] - d.lmax;
            if (score < limit) {
                sel = true;
                d.mark[u] = stamp;
                d.oldp[u] = ub;
                if (strong) {
                    roff[u] = score;
                    if (score + vwt(d, u) >= limit) d.evict_adjust[ub] = score + vwt(d, u);
                }
            }
        }
        append(sel, u, d.moves_x, &ctl->st[ST_MOVES]);
    }
}

__global__ void fb_cookie_kernel(JetDev d) {
    if (threadIdx.x < WAVE && blockIdx.x == 0) cookie_cutter(d, max_block_weight(d, true));
}

__global__ void fb_adjust_kernel(JetDev d, const s64 *roff, u32 *over_flag) {
    const u32 len = (u32) d.ctl->st[ST_MOVES];
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x) {
        const u32 u = d.moves_x[i];
        const u32 ub = d.work[u];
        const s64 pos = roff[u] + d.evict_adjust[ub];
        const u32 dst = chunk_of(d.evict_start, d.k, pos, vwt(d, u), ub);
        if (dst == NO_MOVE) over_flag[u] = 1;
        else d.dest[u] = dst;
    }
}

__global__ void fb_overflow_kernel(JetDev d, const s64 *roff, const u32 *olist, const u32 *olen, u32 *over_flag) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const u32 len = *olen;
    s64 counter = *d.counter;
    for (u32 q = 0; q < len; ++q) {
        const u32 u = olist[q];
        over_flag[u] = 0;
        const u32 ub = d.work[u];
        const s64 pos = roff[u] + d.evict_adjust[ub];
        const u32 id = overflow_chunk(d.evict_start, d.k, pos, vwt(d, u), counter);
        d.dest[u] = id < d.k ? id : ub;
    }
    *d.counter = counter;
}

__global__ void unstall_kernel(JetCtl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->stalled = 0;
}

struct OverSrc {
    const u32 *f;
    __device__ u32 operator()(u64 u) const { return f[u]; }
};
struct OverDst {
    const u32 *f;
    u32 *list;
    __device__ void operator()(u64 u, u32 pos) const {
        if (f[u]) list[pos] = (u32) u;
    }
};

// wide vertices (row capacity min(deg, k) = k): ranks and flags
struct WideSrc {
    const u32 *xadj;
    u32 k;
    __device__ u32 operator()(u64 u) const { return xadj[u + 1] - xadj[u] >= k ? 1u : 0u; }
};
struct WideDst {
    const u32 *xadj;
    u32 k;
    u32 *wid;
    u8 *rflag;
    u32 hub_deg;
    __device__ void operator()(u64 u, u32 pos) const {
        const u32 deg = xadj[u + 1] - xadj[u];
        const bool w = deg >= k;
        const bool counted = w && k <= 64 && deg > ROW_THREAD_DEG && deg <= hub_deg;
        wid[u] = w ? pos : ~0u;
        rflag[u] = w ? (counted ? 7 : 3) : 0;  // the init pass builds every wide row
    }
};

// ---------------------------------------------------------------- host driver
struct JetHost {
    Ctx &c;
    const DGraph &g;
    const GraphMeta &meta;
    JetDev d{};
    u32 G = 16, Gs = 16, Gb = 8, sweep_threads = JB;
    bool big_resolve = false;
    bool block_mode = false;
    u32 bm_groups = 1;
    size_t bm_lds = 0;
    bool bm_ls = false;  // single-workgroup form with the per-vertex state in LDS
    u32 bm_state_off = 0;
    u32 bm_rows_off = 0;  // and the row ids and flags after it (0: in device memory)
    size_t sweep_smem = 0, hub_smem = 0;
    s64 *H = nullptr, *roff = nullptr;
    u32 *over_flag = nullptr, *olist = nullptr, *olen = nullptr;
    bool flat = false;  // edge-balanced afterburner and apply (skewed degree distribution)
    u32 n_chunks = 0;  // hub chunks of the apply step
    u32 *chunk_hub = nullptr, *chunk_idx = nullptr;
    bool rows = false;  // the level has wide vertices (degree >= k)

    JetHost(Ctx &cc, const DGraph &gg, const GraphMeta &mm) : c(cc), g(gg), meta(mm) {}

    // pieces of at most SWEEP_PIECE vertices (the LDS list), and no larger than
    // a list region: on dense levels a region holds few vertices, and one
    // workgroup per 1024 vertices walked every wide vertex of a small level
    template<int GG>
    void launch_rows(u32 init) {
        const size_t smem = 4ull * SWEEP_PIECE + 4ull * (JB / GG) * d.W + (init && d.W <= 2 ? 4ull * (JB / GG) * d.k : 0);
        const u32 piece = std::min<u32>(SWEEP_PIECE, d.cap);
        launch(c, "jet_rows_group", 0.0, dim3(grid_for(c, div_up(d.n, piece), 1, 4)), dim3(JB), smem,
               row_kernel<GG>, d, init, piece);
    }
    // the rows of the wide vertices next to the last iteration's moves (or, at
    // init, of every wide vertex)
    void row_update_launches(u32 init) {
        if (!rows) return;
        if (d.W <= 2)
            launch(c, "jet_rows", 0.0, dim3(grid_for(c, d.n, JB, 8)), dim3(JB), 0, row_thread_kernel, d, init);
        if (d.W > 2 || (init && meta.max_deg > ROW_THREAD_DEG)) HEIPA_DISPATCH_G(G, launch_rows, init);
        if (meta.n_hubs)
            launch(c, "jet_rows_hub", 0.0, dim3(meta.n_hubs), dim3(JB), 4ull * d.W, row_hub_kernel, d,
                   (const u32 *) meta.hubs, init);
    }

    template<int GG>
    void launch_sweep() {
        launch(c, "jet_sweep", 0.0, dim3(d.nreg), dim3(sweep_threads), sweep_smem, sweep_kernel<GG>, d);
    }

    void apply() {
        if (n_chunks)
            launch(c, "jet_apply", 0.0, dim3(n_chunks), dim3(JB), 0, apply_hub_kernel, d, (const u32 *) meta.hubs,
                   (const u32 *) chunk_hub, (const u32 *) chunk_idx);
        if (flat) launch(c, "jet_apply", 0.0, dim3(d.nreg), dim3(JB), 0, apply_kernel, d);
        else HEIPA_DISPATCH_G(G, launch_apply);
        launch(c, "jet_control", 0.0, dim3(1), dim3(JB), 0, control_kernel, d);
        row_update_launches(0);
    }

    template<int GG>
    void launch_post() {
        launch(c, "jet_post", 0.0, dim3(d.nreg), dim3(JB), 0, post_kernel<GG, false>, d);
    }
    template<int GG>
    void launch_apply() {
        launch(c, "jet_apply", 0.0, dim3(d.nreg), dim3(JB), 0, apply_groups_kernel<GG>, d);
    }

    template<int GG>
    void launch_block(u32 max_iters) {
        if (bm_ls)
            launch(c, "jet_block", 0.0, dim3(1), dim3(BM),
                   bm_rows_off ? bm_rows_off + bm_rows_bytes(d.n) : bm_state_off + bm_state_bytes(d.n, d.k),
                   jet_block_kernel<GG, true>, d, max_iters, bm_groups, bm_state_off, bm_rows_off);
        else
            launch(c, "jet_block", 0.0, dim3(1), dim3(BM), bm_lds, jet_block_kernel<GG, false>, d, max_iters, bm_groups,
                   0u, 0u);
    }

    void iteration() {
        if (meta.n_hubs)
            launch(c, "jet_sweep", 0.0, dim3(meta.n_hubs), dim3(JB), hub_smem, hub_sweep_kernel, d,
                   (const u32 *) meta.hubs, 0);
        HEIPA_DISPATCH_G(Gs, launch_sweep);
        if (meta.n_hubs) {
            launch(c, "jet_sweep", 0.0, dim3(meta.n_hubs), dim3(JB), hub_smem, hub_sweep_kernel, d,
                   (const u32 *) meta.hubs, 1);
            launch(c, "jet_post", 0.0, dim3(meta.n_hubs), dim3(JB), 0, hub_post_kernel, d, (const u32 *) meta.hubs);
        }
        launch(c, "jet_buckets", 0.0, dim3(d.k), dim3(JB), 0, bucket_kernel, d);
        if (flat) launch(c, "jet_post", 0.0, dim3(d.nreg), dim3(JB), 0, post_kernel<4, true>, d);
        else HEIPA_DISPATCH_G(G, launch_post);
        if (big_resolve)
            launch(c, "jet_resolve", 0.0, dim3(1), dim3(RB), 0, resolve_kernel<RESOLVE_CAP>, d);
        else
            launch(c, "jet_resolve", 0.0, dim3(1), dim3(RB), 0, resolve_kernel<RESOLVE_CAP_SMALL>, d);
        apply();
    }

    // the stalled iteration's resolve on the global path, then its apply
    void fallback(const JetCtl &s) {
        BackScope bs(c.arena);
        const bool strong = s.mode == 2;
        const u32 L = (u32) s.st[ST_RLIST];
        u64 *k2 = bs.get<u64>(L);
        u32 *v1 = bs.get<u32>(L), *v2 = bs.get<u32>(L);
        s64 *S = bs.get<s64>(L);
        const char *name = "jet_rebalance_global";
        const dim3 grid(grid_for(c, L, 256));
        launch(c, name, 0.0, grid, dim3(256), 0, fb_vals_kernel, (const u64 *) d.rlist, L, v1);
        const bool flip = radix_sort_pairs<u64>(c, d.rlist, v1, k2, v2, L, 32 + bits_for((u64) d.t + 1));
        const u64 *sk = flip ? k2 : d.rlist;
        const u32 *sv = flip ? v2 : v1;
        scan<s64>(c, name, WeightSrc{sv, g.vw}, StoreS{S}, L, (s64 *) nullptr);
        launch(c, name, 0.0, grid, dim3(256), 0, fb_head_kernel, sk, (const s64 *) S, L, H);
        if (strong) fill<s64>(c, d.evict_adjust, d.k, 0);
        launch(c, name, 0.0, grid, dim3(256), 0, fb_select_kernel, d, sk, (const s64 *) S, (const s64 *) H, L, roff,
               (int) strong);
        if (strong) {
            launch(c, name, 0.0, dim3(1), dim3(64), 0, fb_cookie_kernel, d);
            launch(c, name, 0.0, dim3(grid_for(c, L, 256)), dim3(256), 0, fb_adjust_kernel, d, (const s64 *) roff,
                   over_flag);
            scan<u32>(c, name, OverSrc{over_flag}, OverDst{over_flag, olist}, g.n, olen);
            launch(c, name, 0.0, dim3(1), dim3(64), 0, fb_overflow_kernel, d, (const s64 *) roff,
                   (const u32 *) olist, (const u32 *) olen, over_flag);
        }
        launch(c, name, 0.0, dim3(1), dim3(64), 0, unstall_kernel, d.ctl);
        apply();
    }
};

} // namespace

JetResult jet_refine(Ctx &c, const DGraph &g, const GraphMeta &meta, u32 *part, u32 k, s64 lmax, bool ultra,
                     u32 level, s64 cut, s64 maxw, std::vector<heipa_jet_trace> *trace, const s32 *D) {
    JetResult res;
    res.cut = cut;
    res.maxw = maxw;
    if (g.n == 0) return res;
    BackScope bs(c.arena);
    JetHost h(c, g, meta);
    JetDev &d = h.d;
    const u32 n = g.n;
    d.n = n;
    d.k = k;
    d.lmax = lmax;
    d.opt = (g.gw + (s64) (k - 1)) / (s64) k;
    d.xadj = g.xadj;
    d.adj = g.adj;
    d.ew = g.ew;
    d.vw = g.vw;
    d.hub_deg = meta.hub_deg;
    d.D = D;
    d.race = c.race ? 1u : 0u;
    d.count_bytes = c.prof_on ? 1u : 0u;
    if (D) ultra = false;  // ProMap_jet_refine runs one filter ratio, which its LP does not read
    d.part = part;
    d.work = bs.get<u32>(n);
    d.dest = bs.get<u32>(n);
    d.dcache = bs.get<u32>(n);
    d.mark = bs.get<u32>(n);
    d.oldp = bs.get<u32>(n);
    d.gcache = bs.get<s64>(n);
    d.icache = bs.get<s64>(n);
    d.gain1 = bs.get<s64>(n);
    d.lock = bs.get<u8>(2ull * n);  // locks, then the stale bytes
    d.bw = bs.get<s64>(k);
    // minibuckets (jet:364-371, :605-612)
    {
        u32 s = MAX_SECTIONS;
        u32 section_size = (n + s * k) / (s * k);
        if (section_size < 4096) {
            section_size = 4096;
            s = (n + section_size * k) / (section_size * k);
        }
        d.sections = s;
        d.t = MAX_BUCKETS * k * s;
        d.width = MAX_BUCKETS * s;
    }
    d.gid = bs.get<u32>(n);
    d.Q = bs.get<s64>(d.t);
    d.R = bs.get<s64>(d.t);
    d.Tseg = bs.get<s64>(k);
    d.rlist = bs.get<u64>(n);
    d.evict_adjust = bs.get<s64>(k);
    d.evict_start = bs.get<s64>(k + 1);
    d.counter = bs.get<s64>(1);
    h.H = bs.get<s64>((u64) d.t + 1);
    h.roff = bs.get<s64>(n);
    h.over_flag = bs.get<u32>(n);
    h.olist = bs.get<u32>(n);
    h.olen = bs.get<u32>(4);
    // control block, followed by the trace ring when tracing
    const bool tracing = trace != nullptr;
    const size_t ctl_bytes = (sizeof(JetCtl) + 255) & ~(size_t) 255;
    const size_t snap_bytes = ctl_bytes + (tracing ? TRACE_RING * sizeof(heipa_jet_trace) : 0);
    const size_t slot_bytes = (snap_bytes + 255) & ~(size_t) 255;
    require(c.pinned_bytes >= 65536 + 2 * slot_bytes, "pinned staging buffer too small for the Jet control block");
    char *ctl_mem = bs.get<char>(snap_bytes);
    d.ctl = (JetCtl *) ctl_mem;
    d.trace = tracing ? (heipa_jet_trace *) (ctl_mem + ctl_bytes) : nullptr;
    d.ratios = bs.get<double>(MAX_RATIOS);
    RatioTable rt{};
    {
        const double ur[MAX_RATIOS] = {0.85, 0.80, 0.75, 0.70, 0.65, 0.60, 0.55, 0.50, 0.45,
                                       0.40, 0.35, 0.30, 0.25, 0.20, 0.15, 0.10, 0.05};
        if (ultra) {
            for (u32 i = 0; i < MAX_RATIOS; ++i) rt.r[i] = ur[i];
            rt.n = MAX_RATIOS;
        } else {
            rt.r[0] = level == 0 ? 0.25 : 0.75;
            rt.n = 1;
        }
        d.n_ratios = rt.n;
    }
    h.big_resolve = c.max_lds >= 160 * 1024;
    // single-workgroup form for small levels without hubs
    {
        const size_t cgb = conn_group_bytes(k);
        h.Gb = cap_group(meta.G, c.block_group_cap);
        h.bm_groups = (u32) std::max<size_t>(1, std::min<size_t>(BM / h.Gb, BM_CONN_LDS / cgb));
        h.bm_lds = std::max(bm_sweep_lds_bytes(k, h.bm_groups), resolve_lds_bytes(BM_RCAP));
        // one workgroup walks every edge each iteration: dense small levels
        // (coarse levels of skewed graphs) stay on the multi-launch form
        h.block_mode = !c.race && n <= c.block_jet_max_n && (u64) g.m <= 32ull * c.block_jet_max_n &&
                       meta.n_hubs == 0 && h.bm_lds <= std::min<size_t>(c.max_lds, 65536);
        if (h.block_mode) {
            // a launch's LDS is its dynamic size plus the kernel's own __shared__ variables
            h.bm_state_off = (u32) ((h.bm_lds + 15) & ~(size_t) 15);
            h.bm_ls = c.block_lds &&
                      h.bm_state_off + bm_state_bytes(n, k) + jet_block_static_lds(h.Gb, true) <= c.max_lds;
            require(h.bm_ls || h.bm_lds + jet_block_static_lds(h.Gb, false) <= c.max_lds,
                    "single-workgroup Jet does not fit the LDS");
        }
    }
    // copies of the minibucket totals: about one per 64 sweep workgroups, at
    // most 32 and 8 MiB; the single-workgroup form keeps one
    {
        const u64 per_copy = 8ull * ((u64) d.t + 1);
        const u64 wgs = (u64) div_up(n, 256);
        u32 s = 1;
        if (c.race)  // one copy: the fetch-add results are offsets in the minibucket
            s = 1;
        else if (c.bucket_copies)
            while (!h.block_mode && s * 2 <= c.bucket_copies) s *= 2;
        else
            while (!h.block_mode && s < 32 && (u64) s * 2 * 64 <= wgs && (u64) s * 2 * per_copy <= (8ull << 20)) s *= 2;
        d.nshard = s;
        d.btot = bs.get<s64>((u64) s * ((u64) d.t + 1));
    }
    d.rcap = std::min<u32>(c.resolve_cap, h.big_resolve ? RESOLVE_CAP : RESOLVE_CAP_SMALL);
    // lane-group geometry: G lanes per vertex; the sweep keeps one LDS
    // k-counter array per group, so wide k shrinks the groups per workgroup and,
    // if one wave's groups do not fit, widens the groups
    h.G = cap_group(meta.G, c.group_cap);
    h.flat = meta.max_deg > 8u * meta.G;  // skewed: a lane group per vertex would wait on the heaviest ones
    {
        const size_t cgb = conn_group_bytes(k);
        require(cgb <= SWEEP_LDS, "k too large for the LDS block-connectivity counters");
        u32 Gs = meta.G;
        while (Gs < WAVE && (size_t) (WAVE / Gs) * cgb > SWEEP_LDS) Gs <<= 1;
        u32 gpb = JB / Gs;
        while (gpb > WAVE / Gs && gpb * cgb > SWEEP_LDS) gpb >>= 1;
        h.Gs = Gs;
        h.sweep_threads = gpb * Gs;
        h.sweep_smem = cgb * gpb + 12ull * SWEEP_PIECE + 16 + 4ull * k;
        h.hub_smem = cgb + 16 + 4ull * k;
        require(h.sweep_smem <= c.max_lds && h.hub_smem <= 65536 && k < 4096,
                "k too large for the LDS block-connectivity counters");
    }
    // list regions: sweep workgroup r owns vertices [r * cap, (r + 1) * cap),
    // cap a multiple of the region unit; its post workgroup may add its
    // share of the extra candidate list (hubs) to its move region.  At most 4
    // regions per CU: every sweep workgroup is then resident at once (its LDS
    // allows about 6 per CU), and on the RGG 2^20 k = 64 level 0 the sweep
    // took 24 us against 28 us with 8 per CU (2 per CU: 31 us, 16: 30 us).
    // The unit is the sweep workgroup's thread count, halved on dense levels
    // while it holds more than region_edges adjacency entries on average: a
    // workgroup's lane groups take its vertices a few at a time, and on the
    // dense coarse levels of skewed graphs few large regions left most of the
    // chip idle while each workgroup walked hundreds of long adjacencies
    {
        u64 per_block = h.sweep_threads;
        const double avg_deg = (double) g.m / (double) n;
        while (c.region_edges && per_block > 8 && (double) per_block * avg_deg > (double) c.region_edges)
            per_block >>= 1;
        d.nreg = grid_for(c, n, (u32) per_block, 4);
        d.cap = (u32) (div_up(n, (u64) d.nreg * per_block) * per_block);
        d.nreg = div_up(n, d.cap);
        d.cap_mv = d.cap + div_up(meta.n_hubs, d.nreg);
        d.cand = bs.get<u32>((u64) d.nreg * d.cap);
        d.moves = bs.get<u32>((u64) d.nreg * d.cap_mv);
        d.cand_n = bs.get<u32>(d.nreg);
        d.moves_n = bs.get<u32>(d.nreg);
        d.pstat = bs.get<unsigned long long>((u64) d.nreg * PT_COUNT);
        d.cand_x = bs.get<u32>(meta.n_hubs + 1);
        d.moves_x = bs.get<u32>(n);
    }
    if (meta.n_hubs) {  // (hub, chunk) table of the hub apply step
        u32 *cnt = bs.get<u32>(meta.n_hubs), *off = bs.get<u32>(meta.n_hubs + 1);
        const dim3 hg(grid_for(c, meta.n_hubs, 256));
        launch(c, "jet_init", 0.0, hg, dim3(256), 0, hub_chunk_count_kernel, g.xadj, (const u32 *) meta.hubs,
               meta.n_hubs, cnt);
        exclusive_sum<u32>(c, "jet_init", cnt, off, meta.n_hubs);
        readback(c, &h.n_chunks, off + meta.n_hubs, 1);
        h.chunk_hub = bs.get<u32>(h.n_chunks);
        h.chunk_idx = bs.get<u32>(h.n_chunks);
        launch(c, "jet_init", 0.0, hg, dim3(256), 0, hub_chunk_fill_kernel, (const u32 *) cnt, (const u32 *) off,
               meta.n_hubs, h.chunk_hub, h.chunk_idx);
    }
    // BlockConn rows of the wide vertices (block_conn.h)
    if (meta.max_deg >= k) {
        h.rows = true;
        d.W = (k + 31) / 32;
        d.rflag = bs.get<u8>(n);
        d.wid = bs.get<u32>(n);
        u32 *cnt = bs.get<u32>(1);
        scan<u32>(c, "jet_init", WideSrc{g.xadj, k}, WideDst{g.xadj, k, d.wid, d.rflag, meta.hub_deg}, n, cnt);
        u32 nw = 0;
        readback(c, &nw, cnt, 1);
        d.wsz = bs.get<u32>(nw ? nw : 1);
        d.wV = bs.get<u32>((u64) (nw ? nw : 1) * d.W);
        d.wP = bs.get<u32>((u64) (nw ? nw : 1) * d.W);
        if (d.W <= 2) {  // counts of the counted vertices; nw * k <= m (a wide vertex has k edges)
            d.wcnt = bs.get<u32>((u64) (nw ? nw : 1) * k);
            fill<u32>(c, d.wcnt, (u64) (nw ? nw : 1) * k, 0u);
        }
        // the single-workgroup form keeps the row ids and flags in LDS too if
        // they fit beside its state
        if (h.block_mode && h.bm_ls) {
            const size_t off = (h.bm_state_off + bm_state_bytes(n, k) + 15) & ~(size_t) 15;
            if (off + bm_rows_bytes(n) + jet_block_static_lds(h.Gb, true) <= c.max_lds) h.bm_rows_off = (u32) off;
        }
    }
    copy_d2d(c, d.work, part, 4ull * n);
    fill<u32>(c, d.dcache, n, NULL_PART);
    fill<s64>(c, d.icache, n, -1);
    fill<u8>(c, d.lock, 2ull * n, (u8) 0);
    fill<u32>(c, d.mark, n, 0u);
    fill<u32>(c, h.over_flag, n, 0u);
    fill<s64>(c, d.btot, (u64) d.nshard * ((u64) d.t + 1), 0);
    block_weights_dev(c, g, d.work, k, d.bw);
    launch(c, "jet_init", 0.0, dim3(1), dim3(64), 0, ctl_init_kernel, d, rt, cut, maxw);
    h.row_update_launches(1);  // init_BlockConn (jet:951)

    // Batches of iterations; the host reads the control block (and the trace
    // ring) of batch i while batch i + 1 is already queued.  The launches of a
    // batch never change (all loop state is in device memory), so the
    // multi-launch form captures one batch of BATCH_MAX iterations as a HIP
    // graph and replays it: one host call per batch instead of six or more
    // kernel launches per iteration.
    struct GraphExec {
        hipGraphExec_t e = nullptr;
        ~GraphExec() {
            if (e) (void) hipGraphExecDestroy(e);
        }
    } gx;
    if (!h.block_mode && c.use_graphs && !c.prof_on) {
        hipGraph_t graph = nullptr;
        HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        try {
            for (u32 i = 0; i < BATCH_MAX; ++i) h.iteration();
        } catch (...) {
            (void) hipStreamEndCapture(c.stream, &graph);
            if (graph) (void) hipGraphDestroy(graph);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(c.stream, &graph));
        const hipError_t ie = hipGraphInstantiate(&gx.e, graph, nullptr, nullptr, 0);
        (void) hipGraphDestroy(graph);
        HIP_CHECK(ie);
    }
    char *slot[2] = {(char *) c.pinned + 65536, (char *) c.pinned + 65536 + slot_bytes};
    hipEvent_t ev[2] = {c.get_event(), c.get_event()};
    u32 B = 2, consumed = 0;
    u64 submits = 0;
    const u32 bm_iters = tracing ? TRACE_RING / 2 : (1u << 20);
    auto submit = [&](int sidx) {
        if (h.block_mode) {
            HEIPA_DISPATCH_G(h.Gb, h.template launch_block, bm_iters);
        } else if (gx.e) {
            HIP_CHECK(hipGraphLaunch(gx.e, c.stream));
        } else {
            for (u32 i = 0; i < B; ++i) h.iteration();
        }
        if (++submits > (u64) 1 << 20) throw Error(HEIPA_ERR_DEVICE, "jet_refine: iteration guard exceeded");
        copy_d2h(c, slot[sidx], ctl_mem, snap_bytes);
        HIP_CHECK(hipEventRecord(ev[sidx], c.stream));
    };
    JetCtl s{};
    int pend = 0;
    submit(0);
    for (;;) {
        submit(pend ^ 1);
        HIP_CHECK(hipEventSynchronize(ev[pend]));
        std::memcpy(&s, slot[pend], sizeof(JetCtl));
        if (tracing) {
            const heipa_jet_trace *ring = (const heipa_jet_trace *) (slot[pend] + ctl_bytes);
            for (u32 i = consumed; i < s.n_trace; ++i) trace->push_back(ring[i % TRACE_RING]);
            consumed = s.n_trace;
        }
        if (!s.active) break;
        if (s.stalled) {
            HIP_CHECK(hipEventSynchronize(ev[pend ^ 1]));
            h.fallback(s);
            submit(pend);
            continue;
        }
        pend ^= 1;
        B = std::min<u32>(B * 2, BATCH_MAX);
    }
    c.sync();
    c.event_pool.push_back(ev[0]);
    c.event_pool.push_back(ev[1]);
    if (s.save) copy_d2d(c, part, d.work, 4ull * n);
    if (h.block_mode) {
        prof_bytes(c, "jet_block", s.bytes[0] + s.bytes[1] + s.bytes[2] + s.bytes[3] + s.bytes[4]);
    } else {
        prof_bytes(c, "jet_sweep", s.bytes[0]);
        prof_bytes(c, "jet_buckets", s.bytes[1]);
        prof_bytes(c, "jet_post", s.bytes[2]);
        prof_bytes(c, "jet_resolve", s.bytes[3]);
        prof_bytes(c, "jet_apply", s.bytes[4]);
    }
    c.sync();
    res.cut = s.best_cut;
    res.maxw = s.best_maxw;
    res.iterations = s.iters;
    return res;
}

} // namespace heipa
