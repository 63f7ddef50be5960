// Jet refinement on the GPU, deterministic.
//
// Follows jet_refine (src/refinement/jet_label_propagation.h:926-1033) with
// jet_lp (:182-354), rebalance_weak (:592-751), rebalance_strong (:356-590) and
// perform_moves (:753-924).  Differences in mechanism, not in result:
//   * no BlockConn table (block_conn.h): a vertex's block connectivity is
//     rebuilt on demand in LDS — G lanes per vertex add their edges into a
//     dense per-group k-counter array with 64-bit LDS atomics; the first adder
//     of a block appends it to a per-group touched list.  The reference's
//     dest_cache/gain_cache (jet:194-237) decide which vertices are rebuilt.
//     Ties between blocks go to the smallest block id.
//   * the cut change of a move set is computed exactly in 64 bits from the old
//     and new blocks of each moved vertex and its neighbours; it equals the
//     reference's (c1 + c2) / 2 (jet:763-923).
//   * minibucket offsets: candidates are sorted by minibucket with a stable
//     radix sort over an id-ordered list, so a vertex's offset is the weight of
//     smaller ids in its minibucket; strong-rebalance overflow (jet:576-583) is
//     resolved serially in id order.
//   * gain_bucket (jet:147-180): float quotient as in the reference, then an
//     exact comparison against float thresholds of powers of 1.5.
// One host readback per iteration drives the reference's control flow.
#include "eval.hpp"

#include <algorithm>
#include <cmath>

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
constexpr int JB = 256;

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

// live counters / report (device)
enum Stat : u32 {
    ST_MISS = 0,
    ST_CAND,
    ST_MOVES,
    ST_OVER,
    ST_DELTA,  // s64 raw delta (twice the cut change)
    ST_MAXW,
    ST_MHASH,
    ST_GHASH,
    ST_MISS_DEG,  // sum of degrees of the rebuilt (missed) vertices
    ST_CAND_DEG,  // sum of degrees of the afterburner candidates
    ST_MOVE_DEG,  // sum of degrees of the moved vertices
    ST_COUNT
};

struct JetDev {
    u32 n, k;
    s64 lmax;
    const u32 *xadj, *adj;
    const s32 *ew, *vw;
    u32 *work, *dest, *dcache, *mark, *oldp;
    s64 *gcache, *gain1;
    u8 *lock;
    u32 *miss, *cand, *moves;
    s64 *bw;
    unsigned long long *st;  // Stat slots
    u32 stamp;
    u32 hub_deg;
};

__device__ __forceinline__ s64 vwt(const JetDev &d, u32 u) { return d.vw ? (s64) d.vw[u] : 1; }
__device__ __forceinline__ s64 ewt(const JetDev &d, u32 j) { return d.ew ? (s64) d.ew[j] : 1; }

// wave-aggregated append to a device list
__device__ __forceinline__ void append(bool pred, u32 val, u32 *list, unsigned long long *counter) {
    const u64 bal = __ballot(pred);
    if (!bal) return;
    const u32 first = (u32) __builtin_ctzll(bal);
    u32 base = 0;
    if (lane_id() == first) base = (u32) atomicAdd(counter, (unsigned long long) __popcll(bal));
    base = __shfl(base, (int) first, WAVE);
    if (pred) list[base + (u32) __popcll(bal & lanemask_lt())] = val;
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

__device__ __forceinline__ void conn_accumulate(const JetDev &d, const GroupConn &gc, u32 u, u32 lane, u32 G) {
    const u32 b0 = d.xadj[u], e = d.xadj[u + 1];
    for (u32 j = b0 + lane; j < e; j += G) {
        const u32 b = d.work[d.adj[j]];
        const unsigned long long old = atomicAdd(&gc.cnt[b], (unsigned long long) ewt(d, j));
        if (old == 0) gc.touched[atomicAdd(gc.tcount, 1u)] = b;
    }
}

__device__ __forceinline__ void conn_reset(const GroupConn &gc, u32 T, u32 lane, u32 G) {
    for (u32 i = lane; i < T; i += G) gc.cnt[gc.touched[i]] = 0;
}

// lexicographic best (conn desc, block asc) with a validity flag
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

// ---------------------------------------------------------------- LP kernels
// K1a: cached vertices take their cached decision and are filtered; misses are listed
__global__ void lp_cached_kernel(JetDev d) {
    for (u64 i0 = (u64) blockIdx.x * blockDim.x; i0 < d.n; i0 += (u64) gridDim.x * blockDim.x) {
        const u64 u = i0 + threadIdx.x;
        bool is_miss = false, is_cand = false;
        if (u < d.n) {
            const u32 dc = d.dcache[u];
            if (dc == NULL_PART) {
                // hub misses are handled by lp_hub_kernel
                is_miss = (d.xadj[u + 1] - d.xadj[u]) <= d.hub_deg;
            } else {
                d.dest[u] = dc;
                if (dc != NO_MOVE && d.lock[u] == 0) {
                    is_cand = true;
                    d.gain1[u] = d.gcache[u];
                } else {
                    d.gain1[u] = GAIN_MIN;
                    d.lock[u] = 0;
                }
            }
        }
        append(is_miss, (u32) u, d.miss, &d.st[ST_MISS]);
        append(is_cand, (u32) u, d.cand, &d.st[ST_CAND]);
    }
}

// shared evaluation of an LP miss (best_block :200-236 + filter :247-258)
template<int G>
__device__ __forceinline__ void lp_eval(const JetDev &d, const GroupConn &gc, u32 u, u32 lane, double ratio,
                                        Best *red, bool active, bool &cand_out) {
    const u32 T = active ? *gc.tcount : 0;
    const u32 own_b = active ? d.work[u] : 0;
    Best best{0, NO_MOVE};
    for (u32 i = lane; i < T; i += G) {
        const u32 b = gc.touched[i];
        if (b == own_b) continue;
        Best x{(s64) gc.cnt[b], b};
        if (best_better(x, best)) best = x;
    }
    best = best_reduce<G>(best, red);
    cand_out = false;
    if (active && lane == 0) {
        const s64 own = (s64) gc.cnt[own_b];
        u32 best_id = best.b;
        s64 gain = 0;
        if (best_id != NO_MOVE) {
            const s64 bc = best.c;
            if (bc >= own || ((double) own - (double) bc) < floor(ratio * (double) own)) gain = bc - own;
            else best_id = NO_MOVE;
        }
        d.gcache[u] = gain;
        d.dcache[u] = best_id;
        d.dest[u] = best_id;
        if (best_id != NO_MOVE && d.lock[u] == 0) {
            d.gain1[u] = gain;
            cand_out = true;
        } else {
            d.gain1[u] = GAIN_MIN;
            d.lock[u] = 0;
        }
    }
}

// K1b: G lanes per listed miss
template<int G>
__global__ void lp_miss_kernel(JetDev d, double ratio) {
    extern __shared__ __align__(16) char smem[];
    __shared__ Best red[JB / WAVE];
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    GroupConn gc = group_conn(smem, d.k, gi);
    for (u32 i = lane; i < d.k; i += G) gc.cnt[i] = 0;
    if (lane == 0) *gc.tcount = 0;
    __syncthreads();
    const u32 len = (u32) d.st[ST_MISS];
    u64 degs = 0;
    for (u64 i0 = (u64) blockIdx.x * groups; i0 < len; i0 += (u64) gridDim.x * groups) {
        const u64 i = i0 + gi;
        const bool active = i < len;
        const u32 u = active ? d.miss[i] : 0;
        if (active && lane == 0) degs += d.xadj[u + 1] - d.xadj[u];
        if (active) conn_accumulate(d, gc, u, lane, G);
        __syncthreads();
        bool cand;
        lp_eval<G>(d, gc, u, lane, ratio, red, active, cand);
        __syncthreads();
        if (active) conn_reset(gc, *gc.tcount, lane, G);
        __syncthreads();
        if (active && lane == 0) *gc.tcount = 0;
        append(cand, u, d.cand, &d.st[ST_CAND]);
    }
    degs = wave_sum(degs);
    if (lane_id() == 0 && degs) atomicAdd(&d.st[ST_MISS_DEG], (unsigned long long) degs);
}

// K1h: one workgroup per hub (hubs never enter the miss list)
__global__ void lp_hub_kernel(JetDev d, const u32 *hubs, double ratio) {
    extern __shared__ __align__(16) char smem[];
    __shared__ Best red[JB / WAVE];
    const u32 u = hubs[blockIdx.x];
    if (d.dcache[u] != NULL_PART) return;  // cached: done by lp_cached_kernel
    GroupConn gc = group_conn(smem, d.k, 0);
    for (u32 i = threadIdx.x; i < d.k; i += JB) gc.cnt[i] = 0;
    if (threadIdx.x == 0) *gc.tcount = 0;
    __syncthreads();
    conn_accumulate(d, gc, u, threadIdx.x, JB);
    __syncthreads();
    bool cand;
    lp_eval<JB>(d, gc, u, threadIdx.x, ratio, red, true, cand);
    append(cand, u, d.cand, &d.st[ST_CAND]);
    if (threadIdx.x == 0) atomicAdd(&d.st[ST_MISS_DEG], (unsigned long long) (d.xadj[u + 1] - d.xadj[u]));
}

// K2: afterburner (jet:269-328) over candidates, G lanes each; selected
// candidates are locked, stamped and appended to the move list
template<int G>
__device__ __forceinline__ s64 afterburner_change(const JetDev &d, u32 u, u32 lane, u32 stride) {
    const s64 ug = d.gain1[u];
    const u32 o = d.work[u], nn = d.dest[u];
    s64 change = 0;
    for (u32 j = d.xadj[u] + lane; j < d.xadj[u + 1]; j += stride) {
        const u32 v = d.adj[j];
        const s64 vg = d.gain1[v];
        const bool first = vg > ug || (vg == ug && v < u);
        if (!first) continue;
        const u32 vn = d.dest[v], vo = d.work[v];
        change += ewt(d, j) * ((s64) (vn == nn) - (s64) (vn == o) + (s64) (vo == o) - (s64) (vo == nn));
    }
    return change;
}

template<int G>
__global__ void afterburner_kernel(JetDev d, int do_hash) {
    __shared__ s64 red[JB / WAVE];
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    const u32 len = (u32) d.st[ST_CAND];
    u64 gh = 0, degs = 0;
    for (u64 i0 = (u64) blockIdx.x * groups; i0 < len; i0 += (u64) gridDim.x * groups) {
        const u64 i = i0 + gi;
        const bool active = i < len;
        const u32 u = active ? d.cand[i] : 0;
        const bool hub = active && (d.xadj[u + 1] - d.xadj[u]) > d.hub_deg;
        if (active && !hub && lane == 0) degs += d.xadj[u + 1] - d.xadj[u];
        s64 ch = (active && !hub) ? afterburner_change<G>(d, u, lane, G) : 0;
        ch = group_sum<G>(ch, red);
        bool sel = false;
        if (active && !hub && lane == 0) {
            const s64 ug = d.gain1[u];
            if (do_hash) gh += mix64d(mix64d(u) + (u64) ug);
            if (ug + ch >= 0) {
                d.lock[u] = 1;
                d.mark[u] = d.stamp;
                d.oldp[u] = d.work[u];
                sel = true;
            }
        }
        append(sel, u, d.moves, &d.st[ST_MOVES]);
    }
    degs = wave_sum(degs);
    if (lane_id() == 0 && degs) atomicAdd(&d.st[ST_CAND_DEG], (unsigned long long) degs);
    if (do_hash) {
        gh = wave_sum(gh);
        if (lane_id() == 0 && gh) atomicAdd(&d.st[ST_GHASH], (unsigned long long) gh);
    }
}

__global__ void afterburner_hub_kernel(JetDev d, const u32 *hubs, int do_hash) {
    __shared__ s64 red[JB / WAVE];
    const u32 u = hubs[blockIdx.x];
    const s64 ug = d.gain1[u];
    if (ug == GAIN_MIN) return;  // not a candidate
    s64 ch = afterburner_change<JB>(d, u, threadIdx.x, JB);
    ch = group_sum<JB>(ch, red);
    bool sel = false;
    if (threadIdx.x == 0) {
        atomicAdd(&d.st[ST_CAND_DEG], (unsigned long long) (d.xadj[u + 1] - d.xadj[u]));
        if (do_hash) atomicAdd(&d.st[ST_GHASH], (unsigned long long) mix64d(mix64d(u) + (u64) ug));
        if (ug + ch >= 0) {
            d.lock[u] = 1;
            d.mark[u] = d.stamp;
            d.oldp[u] = d.work[u];
            sel = true;
        }
    }
    append(sel, u, d.moves, &d.st[ST_MOVES]);
}

// ---------------------------------------------------------------- apply moves
// K3: exact cut change, cache invalidation (u and N(u)), partition and block
// weight update.  Non-moved neighbours are never written, moved neighbours'
// old block comes from oldp, so the concurrent writes to work[] are safe.
template<int G>
__device__ __forceinline__ s64 move_delta(const JetDev &d, u32 u, u32 lane, u32 stride) {
    const u32 ou = d.oldp[u], nu = d.dest[u];
    s64 delta = 0;
    for (u32 j = d.xadj[u] + lane; j < d.xadj[u + 1]; j += stride) {
        const u32 v = d.adj[j];
        const bool mv = d.mark[v] == d.stamp;
        const u32 ov = mv ? d.oldp[v] : d.work[v];
        const u32 nv = mv ? d.dest[v] : ov;
        const s64 x = ewt(d, j) * ((s64) (nu != nv) - (s64) (ou != ov));
        delta += mv ? x : 2 * x;
        d.dcache[v] = NULL_PART;
    }
    return delta;
}

template<int G>
__global__ void apply_kernel(JetDev d, int do_hash) {
    __shared__ s64 red[JB / WAVE];
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    const u32 len = (u32) d.st[ST_MOVES];
    s64 acc = 0;
    u64 mh = 0, degs = 0;
    for (u64 i0 = (u64) blockIdx.x * groups; i0 < len; i0 += (u64) gridDim.x * groups) {
        const u64 i = i0 + gi;
        const bool active = i < len;
        const u32 u = active ? d.moves[i] : 0;
        const bool hub = active && (d.xadj[u + 1] - d.xadj[u]) > d.hub_deg;
        if (active && !hub) {
            acc += move_delta<G>(d, u, lane, G);
            if (lane == 0) {
                degs += d.xadj[u + 1] - d.xadj[u];
                const u32 ou = d.oldp[u], nu = d.dest[u];
                const s64 w = vwt(d, u);
                d.dcache[u] = NULL_PART;
                d.work[u] = nu;
                atomicAdd((unsigned long long *) &d.bw[ou], (unsigned long long) (-w));
                atomicAdd((unsigned long long *) &d.bw[nu], (unsigned long long) w);
                if (do_hash) mh += mix64d(((u64) u << 32) | nu);
            }
        }
    }
    acc = block_sum<JB>(acc, red);
    if (threadIdx.x == 0 && acc) atomicAdd(&d.st[ST_DELTA], (unsigned long long) acc);
    degs = wave_sum(degs);
    if (lane_id() == 0 && degs) atomicAdd(&d.st[ST_MOVE_DEG], (unsigned long long) degs);
    if (do_hash) {
        mh = wave_sum(mh);
        if (lane_id() == 0 && mh) atomicAdd(&d.st[ST_MHASH], (unsigned long long) mh);
    }
}

__global__ void apply_hub_kernel(JetDev d, const u32 *hubs, int do_hash) {
    __shared__ s64 red[JB / WAVE];
    const u32 u = hubs[blockIdx.x];
    if (d.mark[u] != d.stamp) return;
    s64 acc = move_delta<JB>(d, u, threadIdx.x, JB);
    acc = block_sum<JB>(acc, red);
    if (threadIdx.x == 0) {
        if (acc) atomicAdd(&d.st[ST_DELTA], (unsigned long long) acc);
        atomicAdd(&d.st[ST_MOVE_DEG], (unsigned long long) (d.xadj[u + 1] - d.xadj[u]));
        const u32 ou = d.oldp[u], nu = d.dest[u];
        const s64 w = vwt(d, u);
        d.dcache[u] = NULL_PART;
        d.work[u] = nu;
        atomicAdd((unsigned long long *) &d.bw[ou], (unsigned long long) (-w));
        atomicAdd((unsigned long long *) &d.bw[nu], (unsigned long long) w);
        if (do_hash) atomicAdd(&d.st[ST_MHASH], (unsigned long long) mix64d(((u64) u << 32) | nu));
    }
}

// K4: max block weight into the report, copy of the counters
__global__ void finish_kernel(JetDev d, unsigned long long *report) {
    __shared__ s64 red[1024 / WAVE];
    s64 mx = INT64_MIN;
    for (u32 b = threadIdx.x; b < d.k; b += blockDim.x) mx = d.bw[b] > mx ? d.bw[b] : mx;
    mx = wave_max(mx);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        s64 m = INT64_MIN;
        for (u32 w = 0; w < blockDim.x / WAVE; ++w) m = red[w] > m ? red[w] : m;
        d.st[ST_MAXW] = (unsigned long long) m;
        for (u32 s = 0; s < ST_COUNT; ++s) report[s] = d.st[s];
    }
}

// ---------------------------------------------------------------- rebalancing
struct RebDev {
    u32 sections, width, t;
    s64 max_b_w, opt;
    u32 *under;         // underloaded blocks (weak), ascending
    u32 *n_under;
    u32 *gid;           // minibucket per vertex
    u32 *flag;          // candidate flag per vertex
    s64 *btot;          // bucket totals, t + 1
    s64 *max_vwgt;      // strong: max capacity; then the overflow counter
    s64 *evict_adjust;  // k
    s64 *evict_start;   // k + 1
};

__global__ void weak_prepare_kernel(JetDev d, RebDev r) {
    // underloaded blocks in ascending order (jet:614-632), one block of threads
    __shared__ u32 lds[1024 / WAVE];
    u32 carry = 0;
    for (u32 b0 = 0; b0 < d.k; b0 += blockDim.x) {
        const u32 b = b0 + threadIdx.x;
        const u32 f = b < d.k && d.bw[b] < r.max_b_w;
        u32 tot;
        const u32 e = block_exclusive_scan<1024>(f, lds, tot);
        if (f) r.under[carry + e] = b;
        carry += tot;
    }
    if (threadIdx.x == 0) *r.n_under = carry;
}

__global__ void strong_prepare_kernel(JetDev d, RebDev r) {
    // max capacity of blocks below max_b_w (jet:374-388); -1 if none
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s64 mx = -1;
        for (u32 b = 0; b < d.k; ++b)
            if (d.bw[b] < r.max_b_w) {
                const s64 cap = r.max_b_w - d.bw[b];
                if (cap > mx) mx = cap;
            }
        *r.max_vwgt = mx;
    }
}

// weak best block (jet:658-694) / strong score (jet:399-435): G lanes per vertex
template<int G, bool STRONG>
__device__ __forceinline__ void reb_eval(const JetDev &d, const RebDev &r, const GroupConn &gc, u32 u, u32 lane,
                                         Best *red, s64 *red2, bool active) {
    const u32 T = active ? *gc.tcount : 0;
    const u32 own_b = active ? d.work[u] : 0;
    if constexpr (!STRONG) {
        Best best{0, NO_MOVE};
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (d.bw[b] >= r.max_b_w) continue;
            Best x{(s64) gc.cnt[b], b};
            if (best_better(x, best)) best = x;
        }
        best = best_reduce<G>(best, red);
        if (active && lane == 0) {
            const s64 own = (s64) gc.cnt[own_b];
            u32 dst = own_b;
            s64 gain;
            if (best.b != NO_MOVE && best.c > 0) {
                dst = best.b;
                gain = best.c - own;
            } else {
                const u32 nu = *r.n_under;
                dst = nu ? r.under[u % nu] : own_b;
                gain = -own;
            }
            d.dest[u] = dst;
            if (dst != own_b) {
                const u32 bt = gain_bucket(gain, vwt(d, u));
                r.gid[u] = (MAX_BUCKETS * own_b + bt) * r.sections + (u % r.sections);
                r.flag[u] = 1;
                atomicAdd((unsigned long long *) &r.btot[r.gid[u]], (unsigned long long) vwt(d, u));
            }
        }
    } else {
        s64 sum = 0, count = 0;
        for (u32 i = lane; i < T; i += G) {
            const u32 b = gc.touched[i];
            if (b == own_b) continue;
            if (d.bw[b] < r.max_b_w) {
                sum += (s64) gc.cnt[b];
                count += 1;
            }
        }
        sum = group_sum<G>(sum, red2);
        count = group_sum<G>(count, red2);
        if (active && lane == 0) {
            const s64 own = (s64) gc.cnt[own_b];
            if (count == 0) count = 1;
            const s64 gain = sum / count - own;
            const s64 wu = vwt(d, u);
            const s64 lim = d.bw[own_b] - d.lmax;
            const u32 bt = gain_bucket(gain, wu < lim ? wu : lim);
            r.gid[u] = (MAX_BUCKETS * own_b + bt) * r.sections + (u % r.sections) + 1;
            r.flag[u] = 1;
            atomicAdd((unsigned long long *) &r.btot[r.gid[u]], (unsigned long long) wu);
        }
    }
}

template<bool STRONG>
__device__ __forceinline__ bool reb_eligible(const JetDev &d, const RebDev &r, u32 u) {
    const u32 ub = d.work[u];
    const s64 bwu = d.bw[ub];
    const s64 wu = vwt(d, u);
    if constexpr (!STRONG) {
        return bwu > d.lmax && (double) wu < 1.5 * (double) (bwu - r.opt);
    } else {
        const s64 mv = *r.max_vwgt;
        return mv >= 0 && bwu > d.lmax && wu <= 2 * mv && wu < 2 * (bwu - r.opt);
    }
}

template<int G, bool STRONG>
__global__ void reb_score_kernel(JetDev d, RebDev r) {
    extern __shared__ __align__(16) char smem[];
    __shared__ Best red[JB / WAVE];
    __shared__ s64 red2[JB / WAVE];
    const u32 groups = blockDim.x / G;
    const u32 gi = threadIdx.x / G, lane = threadIdx.x % G;
    GroupConn gc = group_conn(smem, d.k, gi);
    for (u32 i = lane; i < d.k; i += G) gc.cnt[i] = 0;
    if (lane == 0) *gc.tcount = 0;
    __syncthreads();
    for (u64 i0 = (u64) blockIdx.x * groups; i0 < d.n; i0 += (u64) gridDim.x * groups) {
        const u64 uu = i0 + gi;
        const u32 u = (u32) uu;
        bool active = uu < d.n;
        if (active) active = (d.xadj[u + 1] - d.xadj[u]) <= d.hub_deg && reb_eligible<STRONG>(d, r, u);
        if (active) conn_accumulate(d, gc, u, lane, G);
        __syncthreads();
        reb_eval<G, STRONG>(d, r, gc, u, lane, red, red2, active);
        __syncthreads();
        if (active) conn_reset(gc, *gc.tcount, lane, G);
        __syncthreads();
        if (active && lane == 0) *gc.tcount = 0;
    }
}

template<bool STRONG>
__global__ void reb_score_hub_kernel(JetDev d, RebDev r, const u32 *hubs) {
    extern __shared__ __align__(16) char smem[];
    __shared__ Best red[JB / WAVE];
    __shared__ s64 red2[JB / WAVE];
    const u32 u = hubs[blockIdx.x];
    if (!reb_eligible<STRONG>(d, r, u)) return;
    GroupConn gc = group_conn(smem, d.k, 0);
    for (u32 i = threadIdx.x; i < d.k; i += JB) gc.cnt[i] = 0;
    if (threadIdx.x == 0) *gc.tcount = 0;
    __syncthreads();
    conn_accumulate(d, gc, u, threadIdx.x, JB);
    __syncthreads();
    reb_eval<JB, STRONG>(d, r, gc, u, threadIdx.x, red, red2, true);
}

struct FlagSrc {
    const u32 *flag;
    __device__ u32 operator()(u64 u) const { return flag[u]; }
};
struct FlagDst {
    const u32 *flag;
    const u32 *gid;
    u32 *keys, *vals;
    __device__ void operator()(u64 u, u32 pos) const {
        if (flag[u]) {
            keys[pos] = gid[u];
            vals[pos] = (u32) u;
        }
    }
};
struct WeightSrc {
    const u32 *sv;
    const s32 *vw;
    __device__ s64 operator()(u64 i) const { return vw ? (s64) vw[sv[i]] : 1; }
};
struct StoreS {
    s64 *S;
    __device__ void operator()(u64 i, s64 v) const { S[i] = v; }
};

// weak filter (jet:718-744) / strong filter (jet:472-500): score = S[i] - P[block begin]
template<bool STRONG>
__global__ void reb_select_kernel(JetDev d, RebDev r, const u32 *sv, const s64 *S, const s64 *P, u32 len) {
    for (u64 i0 = (u64) blockIdx.x * blockDim.x; i0 < len; i0 += (u64) gridDim.x * blockDim.x) {
        const u64 i = i0 + threadIdx.x;
        bool sel = false;
        u32 u = 0;
        if (i < len) {
            u = sv[i];
            const u32 ub = d.work[u];
            const s64 score = S[i] - P[(u64) ub * r.width];
            const s64 limit = d.bw[ub] - d.lmax;
            if (score < limit) {
                sel = true;
                if (STRONG) {
                    if (score + vwt(d, u) >= limit) r.evict_adjust[ub] = score + vwt(d, u);
                    r.gid[u] = (u32) i;  // position in the sorted order, for adjust_scores
                } else {
                    d.mark[u] = d.stamp;
                    d.oldp[u] = ub;
                }
            }
        }
        append(sel, u, d.moves, &d.st[ST_MOVES]);
    }
}

// cookie cutter (jet:508-552), one thread (k is small)
__global__ void cookie_kernel(JetDev d, RebDev r, const s64 *P) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    s64 prefix = 0;
    for (u32 p = 0; p < d.k; ++p) {
        s64 add = r.evict_adjust[p];
        const u64 begin = (u64) MAX_BUCKETS * p * r.sections;
        if (add == 0) add = P[begin + (u64) MAX_BUCKETS * r.sections] - P[begin];
        r.evict_adjust[p] = P[begin] - prefix;
        prefix += add;
    }
    *r.max_vwgt = prefix;  // overflow counter start
    s64 run = 0;
    r.evict_start[0] = 0;
    for (u32 p = 0; p < d.k; ++p) {
        if (r.max_b_w > d.bw[p]) run += r.max_b_w - d.bw[p];
        r.evict_start[p + 1] = run;
    }
}

// adjust_scores (jet:556-586) first pass: vertices whose position falls well
// inside a chunk get it; the others are flagged for the serial pass
__global__ void adjust_kernel(JetDev d, RebDev r, const s64 *S, u32 *over_flag) {
    const u32 len = (u32) d.st[ST_MOVES];
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x) {
        const u32 u = d.moves[i];
        const u32 ub = d.work[u];
        const s64 pos = S[r.gid[u]] - r.evict_adjust[ub];
        const s64 wu = vwt(d, u);
        s64 id = 0;
        while (id <= (s64) d.k && r.evict_start[id] <= pos) id++;
        id--;
        d.mark[u] = d.stamp;
        d.oldp[u] = ub;
        if (id < (s64) d.k && wu / 2 <= r.evict_start[id + 1] - pos) {
            d.dest[u] = (u32) id;
        } else if (id < (s64) d.k) {
            over_flag[u] = 1;
        } else {
            d.dest[u] = ub;
        }
    }
}

// serial overflow resolution in vertex id order
__global__ void overflow_kernel(JetDev d, RebDev r, const s64 *S, const u32 *olist, const u32 *olen,
                                u32 *over_flag) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const u32 len = *olen;
    s64 counter = *r.max_vwgt;
    for (u32 q = 0; q < len; ++q) {
        const u32 u = olist[q];
        over_flag[u] = 0;
        const u32 ub = d.work[u];
        s64 pos = S[r.gid[u]] - r.evict_adjust[ub];
        const s64 wu = vwt(d, u);
        s64 id = 0;
        bool done = false;
        while (id < (s64) d.k) {
            while (id <= (s64) d.k && r.evict_start[id] <= pos) id++;
            id--;
            if (id < (s64) d.k && wu / 2 <= r.evict_start[id + 1] - pos) {
                d.dest[u] = (u32) id;
                done = true;
                break;
            }
            if (id < (s64) d.k) {
                pos = counter;
                counter += wu;
            }
        }
        if (!done) d.dest[u] = ub;
    }
    *r.max_vwgt = counter;
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

// ---------------------------------------------------------------- host driver
struct JetHost {
    Ctx &c;
    const DGraph &g;
    const GraphMeta &meta;
    JetDev d{};
    RebDev r{};
    unsigned long long *report = nullptr;  // device copy of the counters
    u32 G = 16;
    u32 groups_per_block = 16;
    size_t conn_smem = 0;
    bool trace = false;
    // rebalancing scratch
    u32 *skeys = nullptr, *svals = nullptr, *skeys2 = nullptr, *svals2 = nullptr;
    s64 *S = nullptr, *P = nullptr;
    u32 *over_flag = nullptr, *olist = nullptr, *olen = nullptr;

    JetHost(Ctx &cc, const DGraph &gg, const GraphMeta &mm) : c(cc), g(gg), meta(mm) {}

    void begin_iteration() {
        d.stamp++;
        fill<unsigned long long>(c, d.st, ST_COUNT, 0ull);
    }

    template<int GG>
    void launch_lp_miss(double ratio) {
        launch(c, "jet_lp_best_block", 0.0, dim3(grid_for(c, g.n, groups_per_block, 8)),
               dim3(groups_per_block * GG), conn_smem, lp_miss_kernel<GG>, d, ratio);
    }
    template<int GG>
    void launch_afterburner() {
        launch(c, "jet_afterburner", 0.0, dim3(grid_for(c, g.n, JB / GG, 8)), dim3(JB), 0, afterburner_kernel<GG>, d,
               (int) trace);
    }
    template<int GG>
    void launch_apply() {
        launch(c, "jet_apply", 0.0, dim3(grid_for(c, g.n, JB / GG, 8)), dim3(JB), 0, apply_kernel<GG>, d,
               (int) trace);
    }
    template<int GG>
    void launch_reb_score(bool strong) {
        if (strong)
            launch(c, "jet_rebalance_score", 0.0, dim3(grid_for(c, g.n, groups_per_block, 8)),
                   dim3(groups_per_block * GG), conn_smem, reb_score_kernel<GG, true>, d, r);
        else
            launch(c, "jet_rebalance_score", 0.0, dim3(grid_for(c, g.n, groups_per_block, 8)),
                   dim3(groups_per_block * GG), conn_smem, reb_score_kernel<GG, false>, d, r);
    }

    void apply_and_finish() {
        HEIPA_DISPATCH_G(G, launch_apply);
        if (meta.n_hubs)
            launch(c, "jet_apply", 0.0, dim3(meta.n_hubs), dim3(JB), 0, apply_hub_kernel, d, (const u32 *) meta.hubs,
                   (int) trace);
        launch(c, "jet_finish", 0.0, dim3(1), dim3(256), 0, finish_kernel, d, report);
    }

    void lp(double ratio) {
        begin_iteration();
        // reads dest_cache, lock, gain_cache; writes dest, gain1 (all n-sized)
        const double cached_bytes = 25.0 * g.n;
        launch(c, "jet_lp_cached", cached_bytes, dim3(grid_for(c, g.n, 256, 8)), dim3(256), 0, lp_cached_kernel, d);
        HEIPA_DISPATCH_G(G, launch_lp_miss, ratio);
        if (meta.n_hubs)
            launch(c, "jet_lp_best_block", 0.0, dim3(meta.n_hubs), dim3(JB), conn_group_bytes(g_k()), lp_hub_kernel,
                   d, (const u32 *) meta.hubs, ratio);
        HEIPA_DISPATCH_G(G, launch_afterburner);
        if (meta.n_hubs)
            launch(c, "jet_afterburner", 0.0, dim3(meta.n_hubs), dim3(JB), 0, afterburner_hub_kernel, d,
                   (const u32 *) meta.hubs, (int) trace);
        apply_and_finish();
    }

    u32 g_k() const { return d.k; }

    void sections(u32 k) {
        u32 s = MAX_SECTIONS;
        u32 section_size = (g.n + s * k) / (s * k);
        if (section_size < 4096) {
            section_size = 4096;
            s = (g.n + section_size * k) / (section_size * k);
        }
        r.sections = s;
        r.t = MAX_BUCKETS * k * s;
        r.width = MAX_BUCKETS * s;
    }

    // shared part of weak/strong: score, bucket prefix, stable sort, S
    u32 score_and_sort(bool strong, BackScope &bs) {
        r.btot = bs.get<s64>((u64) r.t + 1);
        P = bs.get<s64>((u64) r.t + 2);
        fill<s64>(c, r.btot, (u64) r.t + 1, 0);
        fill<u32>(c, r.flag, g.n, 0u);
        HEIPA_DISPATCH_G(G, launch_reb_score, strong);
        if (meta.n_hubs) {
            if (strong)
                launch(c, "jet_rebalance_score", 0.0, dim3(meta.n_hubs), dim3(JB), conn_group_bytes(d.k),
                       reb_score_hub_kernel<true>, d, r, (const u32 *) meta.hubs);
            else
                launch(c, "jet_rebalance_score", 0.0, dim3(meta.n_hubs), dim3(JB), conn_group_bytes(d.k),
                       reb_score_hub_kernel<false>, d, r, (const u32 *) meta.hubs);
        }
        exclusive_sum<s64>(c, "jet_rebalance_scan", r.btot, P, (u64) r.t + 1);
        u32 *ncand = olen + 1;
        scan<u32>(c, "jet_rebalance_compact", FlagSrc{r.flag}, FlagDst{r.flag, r.gid, skeys, svals}, g.n, ncand);
        u32 len;
        readback(c, &len, ncand, 1);
        if (len == 0) return 0;
        const bool flip = radix_sort_pairs<u32>(c, skeys, svals, skeys2, svals2, len, bits_for((u64) r.t + 1));
        const u32 *sv = flip ? svals2 : svals;
        scan<s64>(c, "jet_rebalance_scan", WeightSrc{sv, g.vw}, StoreS{S}, len, (s64 *) nullptr);
        if (strong) {
            fill<s64>(c, r.evict_adjust, d.k, 0);
            launch(c, "jet_rebalance_select", 0.0, dim3(grid_for(c, len, 256)), dim3(256), 0,
                   reb_select_kernel<true>, d, r, sv, (const s64 *) S, (const s64 *) P, len);
        } else {
            launch(c, "jet_rebalance_select", 0.0, dim3(grid_for(c, len, 256)), dim3(256), 0,
                   reb_select_kernel<false>, d, r, sv, (const s64 *) S, (const s64 *) P, len);
        }
        return len;
    }

    u32 weak() {
        begin_iteration();
        BackScope bs(c.arena);
        r.opt = (g.gw + (s64) (d.k - 1)) / (s64) d.k;
        s64 mbw = (s64) ((double) d.lmax * 0.99);
        if (mbw < d.lmax - 100) mbw = d.lmax - 100;
        r.max_b_w = mbw;
        sections(d.k);
        launch(c, "jet_rebalance_prepare", 0.0, dim3(1), dim3(1024), 0, weak_prepare_kernel, d, r);
        const u32 n_cand = score_and_sort(false, bs);
        apply_and_finish();
        return n_cand;
    }

    u32 strong() {
        begin_iteration();
        BackScope bs(c.arena);
        r.opt = (g.gw + (s64) (d.k - 1)) / (s64) d.k;
        r.max_b_w = std::max(r.opt + 1, (s64) ((double) d.lmax * 0.99));
        sections(d.k);
        launch(c, "jet_rebalance_prepare", 0.0, dim3(1), dim3(64), 0, strong_prepare_kernel, d, r);
        const u32 n_cand = score_and_sort(true, bs);
        if (n_cand) {
            launch(c, "jet_rebalance_cookie", 0.0, dim3(1), dim3(64), 0, cookie_kernel, d, r, (const s64 *) P);
            launch(c, "jet_rebalance_adjust", 0.0, dim3(grid_for(c, g.n, 256)), dim3(256), 0, adjust_kernel, d, r,
                   (const s64 *) S, over_flag);
            scan<u32>(c, "jet_rebalance_overflow", OverSrc{over_flag}, OverDst{over_flag, olist}, g.n, olen);
            launch(c, "jet_rebalance_overflow", 0.0, dim3(1), dim3(64), 0, overflow_kernel, d, r, (const s64 *) S,
                   (const u32 *) olist, (const u32 *) olen, over_flag);
        }
        apply_and_finish();
        return n_cand;
    }
};

// Algorithmic bytes of the data-dependent Jet kernels, from the counters of
// the iteration (w = 1 when edge weights are stored):
//   best block over misses M: 4(|M|+1)... per miss 8 (xadj) + 4 (own block)
//     + 25 (lock, gain/dest caches, dest, gain1), per edge 4(1+w) + 4 (block of
//     the neighbour, counted at most once per vertex: capped at 4n);
//   afterburner over candidates C: per candidate 8 + 8 (own gain, blocks),
//     per edge 4(1+w), neighbour gain/dest/block 16 per distinct vertex (cap 16n);
//   apply over moves M: per move 16, per edge 4(1+w), neighbour mark/block and
//     cache invalidation 16 per distinct vertex (cap 16n).
void account_bytes(Ctx &c, const DGraph &g, u32 mode, const unsigned long long *rep) {
    const double w = g.ew ? 2.0 : 1.0, n = g.n;
    if (mode == 0) {
        const double nm = (double) rep[ST_MISS], em = (double) rep[ST_MISS_DEG];
        prof_bytes(c, "jet_lp_best_block", 37.0 * nm + 4.0 * w * em + std::min(4.0 * em, 4.0 * n));
        const double nc = (double) rep[ST_CAND], ec = (double) rep[ST_CAND_DEG];
        prof_bytes(c, "jet_afterburner", 16.0 * nc + 4.0 * w * ec + std::min(16.0 * ec, 16.0 * n));
    }
    const double nmv = (double) rep[ST_MOVES], emv = (double) rep[ST_MOVE_DEG];
    prof_bytes(c, "jet_apply", 16.0 * nmv + 4.0 * w * emv + std::min(16.0 * emv, 16.0 * n));
}

} // namespace

JetResult jet_refine(Ctx &c, const DGraph &g, const GraphMeta &meta, u32 *part, u32 k, s64 lmax, bool ultra,
                     u32 level, s64 cut, s64 maxw, std::vector<heipa_jet_trace> *trace) {
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
    d.xadj = g.xadj;
    d.adj = g.adj;
    d.ew = g.ew;
    d.vw = g.vw;
    d.hub_deg = meta.hub_deg;
    d.work = bs.get<u32>(n);
    d.dest = bs.get<u32>(n);
    d.dcache = bs.get<u32>(n);
    d.mark = bs.get<u32>(n);
    d.oldp = bs.get<u32>(n);
    d.gcache = bs.get<s64>(n);
    d.gain1 = bs.get<s64>(n);
    d.lock = bs.get<u8>(n);
    d.miss = bs.get<u32>(n);
    d.cand = bs.get<u32>(n);
    d.moves = bs.get<u32>(n);
    d.bw = bs.get<s64>(k);
    d.st = bs.get<unsigned long long>(ST_COUNT);
    d.stamp = 0;
    h.report = bs.get<unsigned long long>(ST_COUNT);
    h.r.under = bs.get<u32>(k);
    h.r.n_under = bs.get<u32>(1);
    h.r.gid = bs.get<u32>(n);
    h.r.flag = bs.get<u32>(n);
    h.r.max_vwgt = bs.get<s64>(1);
    h.r.evict_adjust = bs.get<s64>(k);
    h.r.evict_start = bs.get<s64>(k + 1);
    h.skeys = bs.get<u32>(n);
    h.svals = bs.get<u32>(n);
    h.skeys2 = bs.get<u32>(n);
    h.svals2 = bs.get<u32>(n);
    h.S = bs.get<s64>(n);
    h.over_flag = bs.get<u32>(n);
    h.olist = bs.get<u32>(n);
    h.olen = bs.get<u32>(4);
    h.trace = trace != nullptr;
    // group geometry: G lanes per vertex, LDS per group for k counters
    h.G = meta.G;
    h.conn_smem = 0;
    {
        const size_t per_group = conn_group_bytes(k);
        require(per_group <= 64 * 1024, "k too large for the LDS block-connectivity counters");
        u32 gpb = JB / h.G;
        while (gpb > 1 && per_group * gpb > 48 * 1024) gpb >>= 1;
        h.groups_per_block = gpb;
        h.conn_smem = per_group * gpb;
    }
    copy_d2d(c, d.work, part, 4ull * n);
    fill<u32>(c, d.dcache, n, NULL_PART);
    fill<u8>(c, d.lock, n, (u8) 0);
    fill<u32>(c, d.mark, n, 0u);
    fill<u32>(c, h.over_flag, n, 0u);
    block_weights_dev(c, g, d.work, k, d.bw);

    std::vector<double> ratios;
    if (ultra) ratios = {0.85, 0.80, 0.75, 0.70, 0.65, 0.60, 0.55, 0.50, 0.45, 0.40, 0.35, 0.30, 0.25, 0.20, 0.15, 0.10, 0.05};
    else ratios = {level == 0 ? 0.25 : 0.75};
    s64 cur_cut = cut, cur_maxw = maxw, best_cut = cut, best_maxw = maxw;
    unsigned long long rep[ST_COUNT];
    u32 iters = 0;
    for (u32 ri = 0; ri < ratios.size(); ++ri) {
        u32 bal = 0, it = 0;
        while (it < N_MAX_ITERATIONS) {
            it += 1;
            u32 mode;
            u32 n_cand = 0;
            if (cur_maxw <= lmax) {
                mode = 0;
                h.lp(ratios[ri]);
            } else if (bal < N_MAX_WEAK_ITERATIONS) {
                mode = 1;
                n_cand = h.weak();
            } else {
                mode = 2;
                n_cand = h.strong();
            }
            readback(c, rep, h.report, ST_COUNT);
            const u32 n_moves = (u32) rep[ST_MOVES];
            if (c.prof_on) account_bytes(c, g, mode, rep);
            if (mode == 0) {
                n_cand = (u32) rep[ST_CAND];
                bal = 0;
                if (n_moves == 0) break;
            } else if (mode == 1) {
                if (n_moves == 0) {
                    bal = N_MAX_WEAK_ITERATIONS;
                    continue;
                }
                bal++;
            } else {
                if (n_moves == 0) break;
                bal++;
            }
            ++iters;
            cur_cut += (s64) rep[ST_DELTA] / 2;
            cur_maxw = (s64) rep[ST_MAXW];
            if (trace) {
                heipa_jet_trace t{};
                t.ratio_idx = ri;
                t.iteration = it;
                t.mode = mode;
                t.n_cand = n_cand;
                t.n_moves = n_moves;
                t.cut = cur_cut;
                t.maxw = cur_maxw;
                t.moves_hash = rep[ST_MHASH];
                t.gain_hash = mode == 0 ? rep[ST_GHASH] : 0;
                trace->push_back(t);
            }
            if (best_maxw > lmax && cur_maxw < best_maxw) {
                copy_d2d(c, part, d.work, 4ull * n);
                best_cut = cur_cut;
                best_maxw = cur_maxw;
                it = 0;
            } else if (cur_cut < best_cut && (cur_maxw <= lmax || cur_maxw <= best_maxw)) {
                if ((double) cur_cut < PHI * (double) best_cut) it = 0;
                copy_d2d(c, part, d.work, 4ull * n);
                best_cut = cur_cut;
                best_maxw = cur_maxw;
            }
        }
    }
    c.sync();
    res.cut = best_cut;
    res.maxw = best_maxw;
    res.iterations = iters;
    return res;
}

} // namespace heipa
