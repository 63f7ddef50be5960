// Two-hop heavy-edge matching on the GPU, deterministic.
//
// Follows two_hop_matcher_get_mapping (src/coarsening/two_hop_matching.h:573-650):
//   * initial pick (:271-297): uniform edge weights -> the neighbour at offset
//     xorshift(u ^ seed) % deg; otherwise the neighbour maximising
//     (edge weight, xorshift(v + xorshift(u ^ seed))).
//   * rounds of 4 commit sub-rounds (:313-333) + repick (:337-371) + compaction
//     (:373-388) until no unmatched vertex has a live candidate.
//   * leaf / twin / relative passes when more than 25% stay unmatched (:592-609).
//   * coarse ids by a scan over representatives in id order (:613-641).
//
// GPU design: the reference's p1 (mark receivers) is evaluated inline by the
// proposer, so a sub-round is two launches: `claim` (every proposer does an
// atomicMin of its id on its receiver's claim slot, so the smallest proposer id
// wins — the serial order of the reference's CAS loop) and `resolve`.  The
// repick and the compaction are one launch; the active list lives in device
// memory, its length too, and the host enqueues rounds without waiting,
// checking the length every few rounds.  Picks run G lanes per vertex with a
// lexicographic argmax reduction over shuffles; hubs get a workgroup.  The
// leaf/twin/relative passes group vertices by exact key with a stable radix
// sort (key, id) and pair consecutive entries, which is the reference's serial
// pairing order.
#include "eval.hpp"

namespace heipa {
namespace {

constexpr u32 SENT = 0xFFFFFFFFu;
constexpr int MB = 256;  // block size

__device__ __forceinline__ u32 xorshift32(u32 x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

__device__ __forceinline__ u32 rotl32(u32 x, int r) { return (x << r) | (x >> (32 - r)); }

// MurmurHash3_x86_32 of the 4 little-endian bytes of x with seed 0
// (Kokkos::pod_hash<u32>, two_hop_matching.h:42)
__device__ __forceinline__ u32 pod_hash(u32 x) {
    u32 k1 = x * 0xcc9e2d51u;
    k1 = rotl32(k1, 15);
    k1 *= 0x1b873593u;
    u32 h1 = k1;
    h1 = rotl32(h1, 13);
    h1 = h1 * 5u + 0xe6546b64u;
    h1 ^= 4u;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    h1 ^= h1 >> 16;
    return h1;
}

// proposer condition of commit sub-round r (two_hop_matching.h:318)
__device__ __forceinline__ bool proposes(u32 u, u32 v, u32 r) {
    return r > 0 ? pod_hash(u + r) < pod_hash(v + r) : u < v;
}

struct Cand {
    s32 w;
    u32 t;
    u32 v;
};

__device__ __forceinline__ bool cand_better(const Cand &a, const Cand &b) {
    if (a.v == SENT) return false;
    if (b.v == SENT) return true;
    if (a.w != b.w) return a.w > b.w;
    return a.t > b.t;
}

template<int WIDTH>
__device__ __forceinline__ Cand cand_reduce(Cand c) {
#pragma unroll
    for (int o = WIDTH / 2; o > 0; o >>= 1) {
        Cand d;
        d.w = __shfl_xor(c.w, o, WIDTH);
        d.t = __shfl_xor(c.t, o, WIDTH);
        d.v = __shfl_xor(c.v, o, WIDTH);
        if (cand_better(d, c)) c = d;
    }
    return c;
}

// lane-partial pick over edges b+lane, b+lane+stride, ...
template<bool INITIAL, bool UEW>
__device__ __forceinline__ Cand pick_partial(const DGraph &g, const u32 *vcmap, u32 b, u32 e, u32 lane, u32 stride,
                                             u32 r) {
    Cand best{0, 0, SENT};
    for (u32 j = b + lane; j < e; j += stride) {
        const u32 v = g.adj[j];
        if (!INITIAL && vcmap[v] != SENT) continue;
        Cand c{UEW ? 1 : g.ew[j], xorshift32(v + r), v};
        if (cand_better(c, best)) best = c;
    }
    return best;
}

// ---------------------------------------------------------------- kernels
__global__ void initial_pick_uniform_kernel(const u32 *xadj, const u32 *adj, u32 n, u32 seed, u32 *hn) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        const u32 b = xadj[u], d = xadj[u + 1] - b;
        if (d == 0) continue;
        hn[u] = adj[b + xorshift32((u32) u ^ seed) % d];
    }
}

// G lanes per vertex; with `list` the vertices are list[0..*len), else 0..n.
// INITIAL: every vertex picks.  Otherwise (repick + compaction): drop matched
// vertices, repick those whose candidate got matched, keep those with a
// candidate in `out_list`.
template<int G, bool INITIAL, bool UEW>
__global__ __launch_bounds__(MB) void pick_kernel(DGraph g, u32 *__restrict__ vcmap, u32 *__restrict__ hn, u32 seed,
                                                  const u32 *__restrict__ list, const u32 *__restrict__ len_ptr,
                                                  u32 hub_deg, u32 *__restrict__ out_list, u32 *__restrict__ out_len) {
    const u32 len = list ? *len_ptr : g.n;
    const u32 lg = threadIdx.x % G;
    const u64 groups = (u64) gridDim.x * (MB / G);
    for (u64 i0 = (u64) blockIdx.x * (MB / G); i0 < len; i0 += groups) {
        const u64 i = i0 + threadIdx.x / G;
        const bool active = i < len;
        u32 u = 0;
        bool keep = false;
        if (active) {
            u = list ? list[i] : (u32) i;
            const u32 b = g.xadj[u], e = g.xadj[u + 1];
            bool need;
            if (INITIAL) need = true;
            else need = vcmap[u] == SENT && hn[u] != SENT && vcmap[hn[u]] != SENT;
            const bool hub = e - b > hub_deg;
            // group-uniform decision: all G lanes see the same u
            if (need && !hub) {
                Cand c = pick_partial<INITIAL, UEW>(g, vcmap, b, e, lg, G, xorshift32(u ^ seed));
                c = cand_reduce<G>(c);
                if (lg == 0) hn[u] = c.v;
            }
            // membership after the repick; lane 0 reads back its own write, hubs
            // were repicked by the hub kernel launched before this one
            if (!INITIAL && lg == 0) keep = vcmap[u] == SENT && hn[u] != SENT;
        }
        if (!INITIAL) {
            const bool leader = active && lg == 0 && keep;
            const u64 bal = __ballot(leader);
            const u32 cnt = (u32) __popcll(bal);
            u32 base = 0;
            const u32 first = bal ? (u32) __builtin_ctzll(bal) : 0;
            if (cnt && lane_id() == first) base = atomicAdd(out_len, cnt);
            base = __shfl(base, (int) first, WAVE);
            if (leader) out_list[base + (u32) __popcll(bal & lanemask_lt())] = u;
        }
    }
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *ctr, u64 v) {
    v = wave_sum(v);
    if (lane_id() == 0 && v) atomicAdd(ctr, (unsigned long long) v);
}

// Repick and compaction (two_hop_matching.h:337-371).  Phase A: one thread
// per list entry tests whether the vertex is unmatched with a matched
// candidate and lists it in LDS; phase B: G lanes per listed vertex pick its
// best unmatched neighbour.  Every unmatched vertex that still has a candidate
// is kept for the next round; kept vertices collect in LDS and each workgroup
// reserves their place in out_list with one counter update per buffer fill
// (same-address atomics from every wave serialised the former form).  Hubs
// were repicked by pick_hub_kernel before this launch.  A workgroup takes
// pieces of `piece` <= MB entries: small pieces on dense levels, where one
// workgroup's lane groups otherwise walked many long adjacencies in turn.
constexpr u32 RP_KEEP = 2048;
template<int G, bool UEW>
__global__ __launch_bounds__(MB) void repick_kernel(DGraph g, u32 *__restrict__ vcmap, u32 *__restrict__ hn, u32 seed,
                                                    const u32 *__restrict__ list, const u32 *__restrict__ len_ptr,
                                                    u32 hub_deg, u32 *__restrict__ out_list, u32 *__restrict__ out_len,
                                                    unsigned long long *deg_acc, u32 piece) {
    __shared__ u32 s_work[MB];
    __shared__ u32 s_keep[RP_KEEP];
    __shared__ u32 s_nwork, s_nkeep, s_base;
    const u32 len = list ? *len_ptr : g.n;
    const u32 lg = threadIdx.x % G, gi = threadIdx.x / G;
    if (threadIdx.x == 0) s_nkeep = 0;
    auto flush = [&]() {
        __syncthreads();
        const u32 cnt = s_nkeep;
        if (threadIdx.x == 0 && cnt) s_base = atomicAdd(out_len, cnt);
        __syncthreads();
        for (u32 i = threadIdx.x; i < cnt; i += MB) out_list[s_base + i] = s_keep[i];
        __syncthreads();
        if (threadIdx.x == 0) s_nkeep = 0;
        __syncthreads();
    };
    for (u64 p0 = (u64) blockIdx.x * piece; p0 < len; p0 += (u64) gridDim.x * piece) {
        if (threadIdx.x == 0) s_nwork = 0;
        __syncthreads();
        const u64 i = p0 + threadIdx.x;
        u32 u = 0, h = SENT;
        bool unmatched = false, need = false;
        if (threadIdx.x < piece && i < len) {
            u = list ? list[i] : (u32) i;
            unmatched = vcmap[u] == SENT;
            h = hn[u];
            if (unmatched && h != SENT) need = vcmap[h] != SENT && g.xadj[u + 1] - g.xadj[u] <= hub_deg;
        }
        u32 pos = 0;
        {
            const u64 bal = __ballot(need);
            const u32 first = bal ? (u32) __builtin_ctzll(bal) : 0;
            u32 base = 0;
            if (bal && lane_id() == first) base = atomicAdd(&s_nwork, (u32) __popcll(bal));
            base = __shfl(base, (int) first, WAVE);
            pos = base + (u32) __popcll(bal & lanemask_lt());
            if (need) s_work[pos] = u;
        }
        __syncthreads();
        const u32 nw = s_nwork;
        u32 degs = 0;  // summed degree of this workgroup's repicked vertices (byte model, profiling only)
        for (u32 r0 = 0; r0 < nw; r0 += MB / G) {
            const u32 idx = r0 + gi;
            const bool act = idx < nw;
            const u32 w = act ? s_work[idx] : 0;
            Cand c{0, 0, SENT};
            if (act) c = pick_partial<false, UEW>(g, vcmap, g.xadj[w], g.xadj[w + 1], lg, G, xorshift32(w ^ seed));
            if (deg_acc && act && lg == 0) degs += g.xadj[w + 1] - g.xadj[w];
            c = cand_reduce<G>(c);
            if (act && lg == 0) {  // only this group reads or writes entry idx
                hn[w] = c.v;
                s_work[idx] = c.v;  // the new candidate, at the vertex's list position
            }
        }
        if (deg_acc) wave_add_u64(deg_acc, degs);
        __syncthreads();
        if (need) h = s_work[pos];
        const bool keep = unmatched && h != SENT;
        if (s_nkeep + MB > RP_KEEP) flush();  // uniform: s_nkeep was last written before a barrier
        {
            const u64 bal = __ballot(keep);
            const u32 first = bal ? (u32) __builtin_ctzll(bal) : 0;
            u32 base = 0;
            if (bal && lane_id() == first) base = atomicAdd(&s_nkeep, (u32) __popcll(bal));
            base = __shfl(base, (int) first, WAVE);
            if (keep) s_keep[base + (u32) __popcll(bal & lanemask_lt())] = u;
        }
        __syncthreads();
    }
    flush();
}

// hubs: one workgroup per hub
template<bool INITIAL, bool UEW>
__global__ __launch_bounds__(MB) void pick_hub_kernel(DGraph g, u32 *__restrict__ vcmap, u32 *__restrict__ hn,
                                                      u32 seed, const u32 *__restrict__ hubs) {
    const u32 u = hubs[blockIdx.x];
    bool need;
    if (INITIAL) need = true;
    else need = vcmap[u] == SENT && hn[u] != SENT && vcmap[hn[u]] != SENT;
    if (!need) return;
    const u32 b = g.xadj[u], e = g.xadj[u + 1];
    Cand c = pick_partial<INITIAL, UEW>(g, vcmap, b, e, threadIdx.x, MB, xorshift32(u ^ seed));
    c = cand_reduce<WAVE>(c);
    __shared__ Cand wbest[MB / WAVE];
    if (lane_id() == 0) wbest[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand best = wbest[0];
        for (int w = 1; w < MB / WAVE; ++w)
            if (cand_better(wbest[w], best)) best = wbest[w];
        hn[u] = best.v;
    }
}

// commit sub-round, claim phase (p1 + p2 of two_hop_matching.h:313-329);
// RACE: the first proposer to claim a receiver wins (the reference's CAS,
// :321-329), else the smallest proposer id (rule D1)
template<bool RACE>
__global__ void claim_kernel(u32 n, const u32 *__restrict__ vcmap, const u32 *__restrict__ hn, u32 *__restrict__ claim,
                             const u32 *__restrict__ list, const u32 *__restrict__ len_ptr, u32 r) {
    const u32 len = list ? *len_ptr : n;
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x) {
        const u32 u = list ? list[i] : (u32) i;
        const u32 v = hn[u];
        if (v == SENT || vcmap[u] != SENT) continue;
        if (!proposes(u, v, r)) continue;  // u is a receiver this sub-round
        // v must be a receiver: unmatched, with a candidate, and not proposing
        if (vcmap[v] != SENT) continue;
        const u32 hv = hn[v];
        if (hv == SENT || proposes(v, hv, r)) continue;
        if (RACE) atomicCAS(&claim[v], SENT, u);
        else atomicMin(&claim[v], u);
    }
}

// commit sub-round, resolve phase (p2 success + p3)
__global__ void resolve_kernel(u32 n, u32 *__restrict__ vcmap, u32 *__restrict__ claim, const u32 *__restrict__ list,
                               const u32 *__restrict__ len_ptr) {
    const u32 len = list ? *len_ptr : n;
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x) {
        const u32 u = list ? list[i] : (u32) i;
        const u32 c = claim[u];
        if (c == SENT) continue;
        const u32 cv = u < c ? u : c;
        vcmap[u] = cv;
        vcmap[c] = cv;
        claim[u] = SENT;
    }
}

// one counter update per workgroup (the grid is small): per-wave updates of
// one address serialise
__global__ __launch_bounds__(256) void count_unmapped_kernel(const u32 *vcmap, u32 n, u32 *out) {
    __shared__ u64 red[256 / WAVE];
    u32 c = 0;
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x)
        c += vcmap[u] == SENT;
    c = (u32) block_sum<256>((u64) c, red);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// ---- keyed pairing (leaf / twin / relative)
struct LeafPred {
    const u32 *vcmap, *xadj;
    __device__ u32 operator()(u64 u) const { return vcmap[u] == SENT && xadj[u + 1] - xadj[u] == 1; }
};
struct UnmappedPred {
    const u32 *vcmap;
    __device__ u32 operator()(u64 u) const { return vcmap[u] == SENT; }
};
template<typename Pred>
struct CompactStore {
    Pred pred;
    u32 *out;
    __device__ void operator()(u64 u, u32 pos) const {
        if (pred(u)) out[pos] = (u32) u;
    }
};

__global__ void leaf_keys_kernel(const u32 *list, const u32 *len_ptr, const u32 *xadj, const u32 *adj, u32 *keys) {
    const u32 len = *len_ptr;
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x)
        keys[i] = adj[xadj[list[i]]];
}

// twin digest: sum over neighbours of h*h + h, h = pod_hash(v) (two_hop_matching.h:455-468)
template<int G>
__global__ __launch_bounds__(MB) void twin_keys_kernel(DGraph g, const u32 *list, const u32 *len_ptr, u64 *keys) {
    const u32 len = *len_ptr;
    const u32 lg = threadIdx.x % G;
    const u64 groups = (u64) gridDim.x * (MB / G);
    for (u64 i0 = (u64) blockIdx.x * (MB / G); i0 < len; i0 += groups) {
        const u64 i = i0 + threadIdx.x / G;
        u64 h = 0;
        if (i < len) {
            const u32 u = list[i];
            for (u32 j = g.xadj[u] + lg; j < g.xadj[u + 1]; j += G) {
                const u64 y = pod_hash(g.adj[j]);
                h += y * y + y;
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) h += __shfl_xor(h, o, G);
        if (i < len && lg == 0) keys[i] = h;
    }
}

// relative key: neighbour of minimum degree, ties -> larger edge weight, then
// first in adjacency order (two_hop_matching.h:540-562); SENT when isolated
template<int G, bool UEW>
__global__ __launch_bounds__(MB) void relative_keys_kernel(DGraph g, const u32 *list, const u32 *len_ptr, u32 *keys) {
    const u32 len = *len_ptr;
    const u32 lg = threadIdx.x % G;
    const u64 groups = (u64) gridDim.x * (MB / G);
    for (u64 i0 = (u64) blockIdx.x * (MB / G); i0 < len; i0 += groups) {
        const u64 i = i0 + threadIdx.x / G;
        u32 bd = SENT, bj = SENT, bv = SENT;
        s32 bw = 0;
        if (i < len) {
            const u32 u = list[i];
            for (u32 j = g.xadj[u] + lg; j < g.xadj[u + 1]; j += G) {
                const u32 v = g.adj[j];
                const u32 d = g.xadj[v + 1] - g.xadj[v];
                const s32 w = UEW ? 1 : g.ew[j];
                if (d < bd || (d == bd && w > bw)) {  // within a lane j increases: keep first
                    bd = d;
                    bw = w;
                    bj = j;
                    bv = v;
                }
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            const u32 od = __shfl_xor(bd, o, G), oj = __shfl_xor(bj, o, G), ov = __shfl_xor(bv, o, G);
            const s32 ow = __shfl_xor(bw, o, G);
            bool take = false;
            if (oj != SENT) {
                if (bj == SENT) take = true;
                else if (od != bd) take = od < bd;
                else if (ow != bw) take = ow > bw;
                else take = oj < bj;
            }
            if (take) {
                bd = od;
                bw = ow;
                bj = oj;
                bv = ov;
            }
        }
        if (i < len && lg == 0) keys[i] = bv;
    }
}

template<typename K>
__global__ void head_flags_kernel(const K *keys, u64 len, u32 *head) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x)
        head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// gid = inclusive scan of head - 1; gstart[gid] = index of the group head
__global__ void group_start_kernel(const u32 *head, const u32 *gid_excl, u64 len, u32 *gstart) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (u64) gridDim.x * blockDim.x)
        if (head[i]) gstart[gid_excl[i]] = (u32) i;
}

__global__ void pair_kernel(const u32 *head, const u32 *gid_excl, const u32 *gstart, const u32 *vals, u64 len,
                            u32 *vcmap) {
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i + 1 < len; i += (u64) gridDim.x * blockDim.x) {
        // group of i: the last head at or before i
        const u32 g = gid_excl[i] + head[i] - 1;
        if (((u32) i - gstart[g]) & 1u) continue;
        if (head[i + 1]) continue;  // i + 1 starts another group
        const u32 a = vals[i], b = vals[i + 1];
        const u32 cv = a < b ? a : b;
        vcmap[a] = cv;
        vcmap[b] = cv;
    }
}

// ---- build_mapping (two_hop_matching.h:613-641)
struct RepSrc {
    u32 *vcmap;
    __device__ u32 operator()(u64 u) const {
        u32 x = vcmap[u];
        return x == SENT || x == (u32) u;
    }
};
struct RepDst {
    u32 *vcmap;
    u32 *mapping;
    __device__ void operator()(u64 u, u32 pos) const {
        u32 x = vcmap[u];
        if (x == SENT || x == (u32) u) {
            vcmap[u] = (u32) u;
            mapping[u] = pos;
        }
    }
};
__global__ void propagate_kernel(const u32 *vcmap, u32 n, u32 *mapping) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        const u32 r = vcmap[u];
        if (r != (u32) u) mapping[u] = mapping[r];
    }
}

// Algorithmic bytes (the pick pass over an active set A: 8 #A + 4 deg(A)(1+w)
// + 4 #A (hn) + 4 #N(A) (vcmap), w = 1 with edge weights): the initial pick
// reads every offset and edge once and writes hn, 4(n+1) + 4m(1+w) + 4n; the
// repick and claim / resolve passes are charged after each batch from the
// read-back list lengths (and, when profiling, the repicked degrees).
template<int G>
void launch_pick(Ctx &c, const DGraph &g, const GraphMeta &meta, bool initial, u32 *vcmap, u32 *hn, u32 seed,
                 const u32 *list, const u32 *len_ptr, u32 len_bound, u32 *out_list, u32 *out_len,
                 unsigned long long *deg_acc) {
    const bool uew = g.ew == nullptr;
    const dim3 block(MB);
    if (initial) {
        const double bytes = 4.0 * (g.n + 1) + 4.0 * g.m * (uew ? 1 : 2) + 4.0 * g.n;
        const dim3 grid(grid_for(c, len_bound, MB / G, 8));
        if (uew)
            launch(c, "match_initial_pick", bytes, grid, block, 0, pick_kernel<G, true, true>, g, vcmap, hn, seed,
                   list, len_ptr, meta.hub_deg, out_list, out_len);
        else
            launch(c, "match_initial_pick", bytes, grid, block, 0, pick_kernel<G, true, false>, g, vcmap, hn, seed,
                   list, len_ptr, meta.hub_deg, out_list, out_len);
    } else {
        // pieces halve while they hold more than region_edges adjacency
        // entries and the grid is not yet at its cap (past it, smaller pieces
        // only add per-piece barriers: on the RGG 2^20 levels with 16-lane
        // groups, 64 rather than 256 entries per piece made these launches
        // 2.4x slower)
        u32 piece = MB;
        const double avg_deg = g.n ? (double) g.m / (double) g.n : 0.0;
        while (c.region_edges && piece > 8 && (double) piece * avg_deg > (double) c.region_edges &&
               (u64) div_up(len_bound, piece) < (u64) c.num_cus * 2)
            piece >>= 1;
        const dim3 grid(grid_for(c, len_bound, piece, 2));
        if (uew)
            launch(c, "match_repick", 0.0, grid, block, 0, repick_kernel<G, true>, g, vcmap, hn, seed, list, len_ptr,
                   meta.hub_deg, out_list, out_len, deg_acc, piece);
        else
            launch(c, "match_repick", 0.0, grid, block, 0, repick_kernel<G, false>, g, vcmap, hn, seed, list,
                   len_ptr, meta.hub_deg, out_list, out_len, deg_acc, piece);
    }
}

template<int G>
void launch_twin_keys(Ctx &c, const DGraph &g, const u32 *list, const u32 *len_ptr, u32 len, u64 *keys) {
    launch(c, "match_twin_keys", 0.0, dim3(grid_for(c, len, MB / G, 8)), dim3(MB), 0, twin_keys_kernel<G>, g, list,
           len_ptr, keys);
}
template<int G>
void launch_relative_keys(Ctx &c, const DGraph &g, const u32 *list, const u32 *len_ptr, u32 len, u32 *keys) {
    if (g.ew)
        launch(c, "match_relative_keys", 0.0, dim3(grid_for(c, len, MB / G, 8)), dim3(MB), 0,
               relative_keys_kernel<G, false>, g, list, len_ptr, keys);
    else
        launch(c, "match_relative_keys", 0.0, dim3(grid_for(c, len, MB / G, 8)), dim3(MB), 0,
               relative_keys_kernel<G, true>, g, list, len_ptr, keys);
}

u32 count_unmapped(Ctx &c, const u32 *vcmap, u32 n, u32 *scratch) {
    fill<u32>(c, scratch, 1, 0u);
    launch(c, "match_count_unmapped", 4.0 * n, dim3(grid_for(c, n, 256, 2)), dim3(256), 0, count_unmapped_kernel, vcmap,
           n, scratch);
    u32 h;
    readback(c, &h, scratch, 1);
    return h;
}

// groups `vals` by exact key (stable radix sort) and pairs consecutive entries
template<typename K>
void pair_by_key(Ctx &c, K *keys, u32 *vals, u32 len, u32 bits, u32 *vcmap) {
    if (len < 2) return;
    BackScope bs(c.arena);
    K *ktmp = bs.get<K>(len);
    u32 *vtmp = bs.get<u32>(len);
    const bool flip = radix_sort_pairs<K>(c, keys, vals, ktmp, vtmp, len, bits);
    const K *sk = flip ? ktmp : keys;
    const u32 *sv = flip ? vtmp : vals;
    u32 *head = bs.get<u32>(len + 1);
    u32 *gid = bs.get<u32>(len + 1);
    u32 *gstart = bs.get<u32>(len + 1);
    const dim3 grid(grid_for(c, len, 256)), block(256);
    launch(c, "match_pair", 0.0, grid, block, 0, head_flags_kernel<K>, sk, (u64) len, head);
    exclusive_sum<u32>(c, "match_pair", head, gid, len);
    launch(c, "match_pair", 0.0, grid, block, 0, group_start_kernel, (const u32 *) head, (const u32 *) gid, (u64) len,
           gstart);
    launch(c, "match_pair", 0.0, grid, block, 0, pair_kernel, (const u32 *) head, (const u32 *) gid,
           (const u32 *) gstart, sv, (u64) len, vcmap);
}

} // namespace

u32 match(Ctx &c, const DGraph &g, const GraphMeta &meta, u32 seed, u32 *mapping, heipa_match_stats *stats) {
    const u32 n = g.n;
    heipa_match_stats st{};
    if (n == 0) {
        if (stats) *stats = st;
        return 0;
    }
    BackScope bs(c.arena);
    u32 *vcmap = bs.get<u32>(n);
    u32 *hn = bs.get<u32>(n);
    u32 *claim = bs.get<u32>(n);
    u32 *lists[2] = {bs.get<u32>(n), bs.get<u32>(n)};
    u32 *lens = bs.get<u32>(64);  // ring of list lengths, one per round
    fill<u32>(c, vcmap, n, SENT);
    fill<u32>(c, hn, n, SENT);
    fill<u32>(c, claim, n, SENT);
    fill<u32>(c, lens, 64, 0u);
    const dim3 blk(256);

    // initial pick
    // profiling: summed degree of the repicked vertices, for the byte model
    unsigned long long *deg_acc = nullptr;
    if (c.prof_on) {
        deg_acc = bs.get<unsigned long long>(1);
        fill<unsigned long long>(c, deg_acc, 1, 0ull);
    }
    const double wb = g.ew ? 2.0 : 1.0;
    if (g.ew == nullptr) {
        // offsets, one gathered neighbour and hn per vertex
        launch(c, "match_initial_pick", 4.0 * (n + 1) + 8.0 * n, dim3(grid_for(c, n, 256)), blk, 0,
               initial_pick_uniform_kernel, g.xadj, g.adj, n, seed, hn);
    } else {
        HEIPA_DISPATCH_G(meta.G, launch_pick, c, g, meta, true, vcmap, hn, seed, (const u32 *) nullptr,
                         (const u32 *) nullptr, n, (u32 *) nullptr, (u32 *) nullptr,
                         (unsigned long long *) nullptr);
        if (meta.n_hubs)
            launch(c, "match_initial_pick", 0.0, dim3(meta.n_hubs), dim3(MB), 0, pick_hub_kernel<true, false>, g,
                   vcmap, hn, seed, (const u32 *) meta.hubs);
    }

    // rounds: list for round 0 is implicit (all vertices)
    const u32 *cur_list = nullptr;
    const u32 *cur_len = nullptr;
    u32 len_bound = n;
    u32 round = 0;
    u32 batch = 2;
    u32 *pinned = c.pinned_as<u32>();
    u32 exact_rounds = 0;
    while (true) {
        const u32 first_round = round;
        for (u32 bi = 0; bi < batch; ++bi, ++round) {
            const u32 round_seed = seed ^ (round * 0x9e3779b1u);
            const dim3 grid(grid_for(c, len_bound, 256, 8));
            for (u32 r = 0; r < 4; ++r) {
                launch(c, "match_claim", 0.0, grid, blk, 0, c.race ? claim_kernel<true> : claim_kernel<false>, n,
                       (const u32 *) vcmap, (const u32 *) hn, claim, cur_list, cur_len, r);
                launch(c, "match_resolve", 0.0, grid, blk, 0, resolve_kernel, n, vcmap, claim, cur_list, cur_len);
            }
            if (meta.n_hubs)
                launch(c, "match_repick", 0.0, dim3(meta.n_hubs), dim3(MB), 0,
                       g.ew ? pick_hub_kernel<false, false> : pick_hub_kernel<false, true>, g, vcmap, hn, round_seed,
                       (const u32 *) meta.hubs);
            u32 *next_list = lists[round & 1];
            u32 *next_len = lens + (round % 64);
            fill<u32>(c, next_len, 1, 0u);
            HEIPA_DISPATCH_G(meta.G, launch_pick, c, g, meta, false, vcmap, hn, round_seed, cur_list, cur_len,
                             len_bound, next_list, next_len, deg_acc);
            cur_list = next_list;
            cur_len = next_len;
        }
        // lengths produced by this batch's rounds (ring slots are contiguous
        // unless the batch wraps the ring)
        for (u32 rr = first_round; rr < round; ++rr) copy_d2h(c, pinned + (rr - first_round), lens + (rr % 64), 4);
        c.sync();
        u32 len = 0;
        bool done = false;
        u32 in_len = len_bound;  // input list length of round first_round
        for (u32 rr = first_round; rr < round; ++rr) {
            // bytes of round rr over its input list of length L: four claim
            // passes (list, hn and vcmap of u, hn and vcmap of the candidate:
            // 20 L), four resolve passes (list, claim read and reset: 12 L), the
            // repick's phase A (list, vcmap, hn, vcmap of the candidate: 16 L)
            const double L = (double) in_len;
            prof_bytes(c, "match_claim", 80.0 * L);
            prof_bytes(c, "match_resolve", 48.0 * L);
            prof_bytes(c, "match_repick", 16.0 * L);
            len = pinned[rr - first_round];
            in_len = len;
            if (len == 0) {
                exact_rounds = rr + 1;  // the loop of the reference stops after this round
                done = true;
                break;
            }
        }
        if (done) break;
        len_bound = len;
        batch = batch < 8 ? batch * 2 : 8;
        require(round < 20000, "matching did not converge (is the graph symmetric?)");
    }
    st.rounds = exact_rounds;
    if (deg_acc) {  // the repicked vertices' edges (and weights) and their neighbours' vcmap
        unsigned long long dsum = 0;
        readback(c, &dsum, deg_acc, 1);
        prof_bytes(c, "match_repick", 4.0 * wb * (double) dsum + 4.0 * std::min<double>((double) dsum, n));
    }

    u32 *scr = bs.get<u32>(4);
    u32 unmapped = count_unmapped(c, vcmap, n, scr);
    st.after_hem = st.after_leaf = st.after_twin = st.after_relative = unmapped;
    u32 *plist = lists[0];
    u32 *plen = scr + 1;
    if ((double) unmapped / (double) n > 0.25) {  // leaf
        st.ran_leaf = 1;
        LeafPred pred{vcmap, g.xadj};
        scan<u32>(c, "match_leaf_list", pred, CompactStore<LeafPred>{pred, plist}, n, plen);
        u32 len;
        readback(c, &len, plen, 1);
        if (len > 1) {
            BackScope bs2(c.arena);
            u32 *keys = bs2.get<u32>(len);
            launch(c, "match_leaf_keys", 0.0, dim3(grid_for(c, len, 256)), blk, 0, leaf_keys_kernel,
                   (const u32 *) plist, (const u32 *) plen, g.xadj, g.adj, keys);
            pair_by_key<u32>(c, keys, plist, len, bits_for(n - 1), vcmap);
        }
        unmapped = count_unmapped(c, vcmap, n, scr);
    }
    st.after_leaf = st.after_twin = st.after_relative = unmapped;
    if ((double) unmapped / (double) n > 0.25) {  // twin
        st.ran_twin = 1;
        UnmappedPred pred{vcmap};
        scan<u32>(c, "match_twin_list", pred, CompactStore<UnmappedPred>{pred, plist}, n, plen);
        const u32 len = unmapped;
        if (len > 1) {
            BackScope bs2(c.arena);
            u64 *keys = bs2.get<u64>(len);
            HEIPA_DISPATCH_G(meta.G, launch_twin_keys, c, g, (const u32 *) plist, (const u32 *) plen, len, keys);
            pair_by_key<u64>(c, keys, plist, len, 64, vcmap);
        }
        unmapped = count_unmapped(c, vcmap, n, scr);
    }
    st.after_twin = st.after_relative = unmapped;
    if ((double) unmapped / (double) n > 0.25) {  // relative
        st.ran_relative = 1;
        UnmappedPred pred{vcmap};
        scan<u32>(c, "match_relative_list", pred, CompactStore<UnmappedPred>{pred, plist}, n, plen);
        const u32 len = unmapped;
        if (len > 1) {
            BackScope bs2(c.arena);
            u32 *keys = bs2.get<u32>(len);
            HEIPA_DISPATCH_G(meta.G, launch_relative_keys, c, g, (const u32 *) plist, (const u32 *) plen, len, keys);
            pair_by_key<u32>(c, keys, plist, len, 32, vcmap);
        }
        st.after_relative = count_unmapped(c, vcmap, n, scr);
    }

    // build_mapping
    u32 *nc_dev = scr + 2;
    scan<u32>(c, "match_build_mapping", RepSrc{vcmap}, RepDst{vcmap, mapping}, n, nc_dev, 0u, 12.0 * n);
    launch(c, "match_build_mapping", 12.0 * n, dim3(grid_for(c, n, 256)), blk, 0, propagate_kernel,
           (const u32 *) vcmap, n, mapping);
    u32 nc;
    readback(c, &nc, nc_dev, 1);
    if (stats) *stats = st;
    return nc;
}

} // namespace heipa
