// Graph contraction and partition projection.
//
// Semantics of from_Graph_Mapping (src/datastructures/graph.h:127-390): coarse
// vertex weight = sum of member weights; the coarse adjacency of c is the
// aggregation by M(v) of all fine edges (u,v) with M(u) = c != M(v), weights
// summed; g_weight is inherited and the coarse graph always stores weights.
// Canonical form: every coarse row is sorted by neighbour id (the reference's
// row order is its hash-slot order, a race).
//
// GPU design: instead of global hash tables with CAS, each coarse row is
// produced by one worker that gathers its members' adjacencies, maps them,
// sorts by coarse id and merges duplicates:
//   * ub(c) = sum of member degrees <= 64: W lanes per row (W = 8..64), bitonic
//     sort in registers over cross-lane shuffles, segmented sum, ballot
//     compaction;
//   * ub(c) <= 4096: one workgroup per row, bitonic sort in LDS;
//   * larger (hubs): all such rows at once through a global stable radix sort
//     of (row, coarse id) keys.
// Rows are written to a staging buffer at offset prefix(ub), then the unique
// counts are prefix-summed into the coarse xadj and the rows copied compact.
//
// Algorithmic bytes: read 4(n+1) + 4m(1+w) + 4n (mapping) + 4n*vw;
// write 4(nc+1) + 8mc + 4nc.
#include "eval.hpp"

namespace heipa {
namespace {

constexpr u32 PAD = 0xFFFFFFFFu;
constexpr u32 BLK_ROW_MAX = 4096;
constexpr int CB = 256;

// members of coarse vertices (counting sort by M(u)) + coarse weights + ub
template<bool UVW>
__global__ void member_count_kernel(DGraph g, const u32 *__restrict__ map, u32 *__restrict__ mcnt,
                                    unsigned long long *__restrict__ cw, u32 *__restrict__ ub) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < g.n; u += (u64) gridDim.x * blockDim.x) {
        const u32 c = map[u];
        atomicAdd(&mcnt[c], 1u);
        atomicAdd(&cw[c], (unsigned long long) (UVW ? 1 : g.vw[u]));
        atomicAdd(&ub[c], g.xadj[u + 1] - g.xadj[u]);
    }
}

__global__ void member_scatter_kernel(u32 n, const u32 *__restrict__ map, const u32 *__restrict__ moff,
                                      u32 *__restrict__ mfill, u32 *__restrict__ mem) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        const u32 c = map[u];
        mem[moff[c] + atomicAdd(&mfill[c], 1u)] = (u32) u;
    }
}

__global__ void cw_finish_kernel(u32 nc, const unsigned long long *__restrict__ cw, s32 *__restrict__ cvw,
                                 u32 *__restrict__ overflow) {
    for (u64 c = (u64) blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (u64) gridDim.x * blockDim.x) {
        const unsigned long long w = cw[c];
        if (w > (unsigned long long) INT32_MAX) *overflow = 1;
        cvw[c] = (s32) w;
    }
}

// tiers: 0..3 -> W = 8,16,32,64; 4 -> workgroup; 5 -> global sort
__device__ __forceinline__ u32 tier_of(u32 ub) {
    if (ub <= 8) return 0;
    if (ub <= 16) return 1;
    if (ub <= 32) return 2;
    if (ub <= 64) return 3;
    if (ub <= BLK_ROW_MAX) return 4;
    return 5;
}

// Each workgroup takes a contiguous chunk of coarse vertices: it counts them
// per tier in LDS, reserves each tier's space with one counter update, then
// writes its entries (per-wave updates of the six counters serialised).
__global__ __launch_bounds__(256) void classify_kernel(u32 nc, const u32 *__restrict__ ub, u32 *__restrict__ tier_cnt,
                                                       u32 *__restrict__ tier_lists, u32 stride) {
    __shared__ u32 s_cnt[6], s_base[6];
    const u64 chunk = ((u64) nc + gridDim.x - 1) / gridDim.x;
    const u64 c0 = (u64) blockIdx.x * chunk;
    const u64 c1 = c0 + chunk < nc ? c0 + chunk : nc;
    if (threadIdx.x < 6) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    for (u64 i0 = c0; i0 < c1; i0 += blockDim.x) {
        const u64 c = i0 + threadIdx.x;
        const u32 t = c < c1 ? tier_of(ub[c]) : 99u;
        for (u32 tt = 0; tt < 6; ++tt) {
            const u64 bal = __ballot(t == tt);
            if (bal && lane_id() == 0) atomicAdd(&s_cnt[tt], (u32) __popcll(bal));
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const u32 cnt = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = cnt ? atomicAdd(&tier_cnt[threadIdx.x], cnt) : 0u;
        s_cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    for (u64 i0 = c0; i0 < c1; i0 += blockDim.x) {
        const u64 c = i0 + threadIdx.x;
        const u32 t = c < c1 ? tier_of(ub[c]) : 99u;
        for (u32 tt = 0; tt < 6; ++tt) {
            const bool mine = t == tt;
            const u64 bal = __ballot(mine);
            if (!bal) continue;
            const u32 first = (u32) __builtin_ctzll(bal);
            u32 base = 0;
            if (lane_id() == first) base = atomicAdd(&s_cnt[tt], (u32) __popcll(bal));
            base = __shfl(base, (int) first, WAVE);
            if (mine) tier_lists[(u64) tt * stride + s_base[tt] + base + (u32) __popcll(bal & lanemask_lt())] = (u32) c;
        }
    }
}

// locate entry `idx` of the concatenated member adjacencies of coarse vertex c
__device__ __forceinline__ bool row_entry(const DGraph &g, const u32 *mem, u32 mb, u32 me, u32 idx, u32 &j) {
    for (u32 t = mb; t < me; ++t) {
        const u32 u = mem[t];
        const u32 b = g.xadj[u], d = g.xadj[u + 1] - b;
        if (idx < d) {
            j = b + idx;
            return true;
        }
        idx -= d;
    }
    return false;
}

template<int W, bool UEW>
__global__ __launch_bounds__(CB) void row_small_kernel(DGraph g, const u32 *__restrict__ map,
                                                       const u32 *__restrict__ moff, const u32 *__restrict__ mem,
                                                       const u32 *__restrict__ uboff, const u32 *__restrict__ list,
                                                       u32 len, u32 *__restrict__ tkey, s32 *__restrict__ tw,
                                                       u32 *__restrict__ cnt, u32 *__restrict__ overflow) {
    const u32 l = threadIdx.x % W;
    const u64 groups = (u64) gridDim.x * (CB / W);
    const u32 sub_base = (lane_id() / W) * W;  // first wave lane of this group
    for (u64 i0 = (u64) blockIdx.x * (CB / W); i0 < len; i0 += groups) {
        const u64 i = i0 + threadIdx.x / W;
        const bool active = i < len;
        u32 c = 0, key = PAD;
        s64 w = 0;
        if (active) {
            c = list[i];
            u32 j;
            if (row_entry(g, mem, moff[c], moff[c + 1], l, j)) {
                const u32 x = map[g.adj[j]];
                if (x != c) {
                    key = x;
                    w = UEW ? 1 : g.ew[j];
                }
            }
        }
        // bitonic sort ascending by key across the W lanes
#pragma unroll
        for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                const u32 pk = __shfl_xor(key, jj, W);
                const s64 pw = __shfl_xor(w, jj, W);
                const bool up = (l & k) == 0;
                const bool lower = (l & jj) == 0;
                const bool take = (lower == up) ? (pk < key) : (pk > key);
                if (take) {
                    key = pk;
                    w = pw;
                }
            }
        }
        // segmented inclusive sum of w over runs of equal keys
        const u32 prev = __shfl_up(key, 1, W);
        const bool head = key != PAD && (l == 0 || prev != key);
        s64 acc = w;
        int seg = head ? 1 : 0;
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            const s64 y = __shfl_up(acc, o, W);
            const int f = __shfl_up(seg, o, W);
            if (l >= (u32) o && !seg) {
                acc += y;
                seg = f;
            }
        }
        const u32 next = __shfl_down(key, 1, W);
        const bool tail = key != PAD && (l == W - 1 || next != key);
        const u64 heads = __ballot(head);
        const u64 gmask = (W == 64 ? ~0ull : ((1ull << W) - 1ull)) << sub_base;
        const u64 hg = heads & gmask;
        if (active) {
            if (tail) {
                const u32 rank = (u32) __popcll(hg & ((2ull << lane_id()) - 1ull)) - 1u;
                tkey[uboff[c] + rank] = key;
                if (acc > INT32_MAX) *overflow = 1;
                tw[uboff[c] + rank] = (s32) acc;
            }
            if (l == 0) cnt[c] = (u32) __popcll(hg);
        }
    }
}

template<bool UEW>
__global__ __launch_bounds__(CB) void row_block_kernel(DGraph g, const u32 *__restrict__ map,
                                                       const u32 *__restrict__ moff, const u32 *__restrict__ mem,
                                                       const u32 *__restrict__ uboff, const u32 *__restrict__ ub,
                                                       const u32 *__restrict__ list, u32 *__restrict__ tkey,
                                                       s32 *__restrict__ tw, u32 *__restrict__ cnt,
                                                       u32 *__restrict__ overflow) {
    __shared__ u32 skey[BLK_ROW_MAX];
    __shared__ s32 sw[BLK_ROW_MAX];
    __shared__ u32 red[CB / WAVE];
    const u32 c = list[blockIdx.x];
    const u32 len = ub[c];
    // P is a power of two and a multiple of the block size, so every thread's
    // slot in the dedupe loop below holds a sorted key or PAD
    u32 P = CB;
    while (P < len) P <<= 1;
    // gather: walk the members' adjacency ranges in order
    {
        u32 base = 0;
        for (u32 t = moff[c]; t < moff[c + 1]; ++t) {
            const u32 u = mem[t];
            const u32 b = g.xadj[u], d = g.xadj[u + 1] - b;
            for (u32 x = threadIdx.x; x < d; x += CB) {
                const u32 y = map[g.adj[b + x]];
                skey[base + x] = y == c ? PAD : y;
                sw[base + x] = UEW ? 1 : g.ew[b + x];
            }
            base += d;
        }
        for (u32 x = len + threadIdx.x; x < P; x += CB) {
            skey[x] = PAD;
            sw[x] = 0;
        }
    }
    __syncthreads();
    for (u32 k = 2; k <= P; k <<= 1) {
        for (u32 jj = k >> 1; jj > 0; jj >>= 1) {
            for (u32 i = threadIdx.x; i < P; i += CB) {
                const u32 p = i ^ jj;
                if (p > i) {
                    const bool up = (i & k) == 0;
                    const u32 a = skey[i], b = skey[p];
                    if (up ? (a > b) : (a < b)) {
                        skey[i] = b;
                        skey[p] = a;
                        const s32 t = sw[i];
                        sw[i] = sw[p];
                        sw[p] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    // unique heads, ranks by block scan, run sums by the head thread
    u32 carry = 0;
    for (u32 b0 = 0; b0 < P; b0 += CB) {
        const u32 i = b0 + threadIdx.x;
        const u32 key = skey[i];
        const bool head = key != PAD && (i == 0 || skey[i - 1] != key);
        u32 tot;
        const u32 rank = block_exclusive_scan<CB>(head ? 1u : 0u, red, tot);
        if (head) {
            s64 s = 0;
            for (u32 x = i; x < P && skey[x] == key; ++x) s += sw[x];
            if (s > INT32_MAX) *overflow = 1;
            tkey[uboff[c] + carry + rank] = key;
            tw[uboff[c] + carry + rank] = (s32) s;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) cnt[c] = carry;
}

// ---- hub rows through a global sort of (row index << 32 | coarse id)
__global__ void hub_gather_kernel(DGraph g, const u32 *__restrict__ map, const u32 *__restrict__ moff,
                                  const u32 *__restrict__ mem, const u32 *__restrict__ list,
                                  const u32 *__restrict__ hoff, u32 nh, u64 *__restrict__ keys,
                                  u32 *__restrict__ vals) {
    // one block per hub row
    const u32 t = blockIdx.x;
    if (t >= nh) return;
    const u32 c = list[t];
    u32 base = hoff[t];
    for (u32 q = moff[c]; q < moff[c + 1]; ++q) {
        const u32 u = mem[q];
        const u32 b = g.xadj[u], d = g.xadj[u + 1] - b;
        for (u32 x = threadIdx.x; x < d; x += blockDim.x) {
            const u32 y = map[g.adj[b + x]];
            keys[base + x] = ((u64) t << 32) | (y == c ? PAD : y);
            vals[base + x] = g.ew ? (u32) g.ew[b + x] : 1u;
        }
        base += d;
    }
}

__global__ void hub_merge_kernel(const u64 *__restrict__ keys, const u32 *__restrict__ vals,
                                 const u32 *__restrict__ list, const u32 *__restrict__ hoff, u32 nh,
                                 const u32 *__restrict__ uboff, u32 *__restrict__ tkey, s32 *__restrict__ tw,
                                 u32 *__restrict__ cnt, u32 *__restrict__ overflow) {
    // one block per hub row; the row's sorted entries are [hoff[t], hoff[t+1])
    __shared__ u32 red[CB / WAVE];
    const u32 t = blockIdx.x;
    if (t >= nh) return;
    const u32 c = list[t];
    const u32 b = hoff[t], e = hoff[t + 1];
    u32 carry = 0;
    for (u32 b0 = b; b0 < e; b0 += CB) {
        const u32 i = b0 + threadIdx.x;
        const u32 key = i < e ? (u32) keys[i] : PAD;
        const bool head = key != PAD && (i == b || (u32) keys[i - 1] != key);
        u32 tot;
        const u32 rank = block_exclusive_scan<CB>(head ? 1u : 0u, red, tot);
        if (head) {
            s64 s = 0;
            for (u32 x = i; x < e && (u32) keys[x] == key; ++x) s += (s32) vals[x];
            if (s > INT32_MAX) *overflow = 1;
            tkey[uboff[c] + carry + rank] = key;
            tw[uboff[c] + carry + rank] = (s32) s;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) cnt[c] = carry;
}

__global__ void hub_offsets_kernel(const u32 *__restrict__ list, u32 nh, const u32 *__restrict__ ub,
                                   u32 *__restrict__ hub_ub) {
    for (u64 t = (u64) blockIdx.x * blockDim.x + threadIdx.x; t < nh; t += (u64) gridDim.x * blockDim.x)
        hub_ub[t] = ub[list[t]];
}

// compact copy of the staged rows, W lanes per row
__global__ __launch_bounds__(CB) void copy_rows_kernel(u32 nc, const u32 *__restrict__ uboff,
                                                       const u32 *__restrict__ cxadj, const u32 *__restrict__ tkey,
                                                       const s32 *__restrict__ tw, u32 *__restrict__ cadj,
                                                       s32 *__restrict__ cew, u32 *__restrict__ bad) {
    constexpr int W = 16;
    const u32 l = threadIdx.x % W;
    const u64 groups = (u64) gridDim.x * (CB / W);
    for (u64 c = (u64) blockIdx.x * (CB / W) + threadIdx.x / W; c < nc; c += groups) {
        const u32 src = uboff[c], dst = cxadj[c], len = cxadj[c + 1] - dst;
        if (len > uboff[c + 1] - src) {
            *bad = 1;
            continue;
        }
        for (u32 x = l; x < len; x += W) {
            const u32 key = tkey[src + x];
            if (key >= nc || key == (u32) c) *bad = 1;
            cadj[dst + x] = key;
            cew[dst + x] = tw[src + x];
        }
    }
}

__global__ void project_kernel(const u32 *__restrict__ map, u32 n, const u32 *__restrict__ cpart,
                               u32 *__restrict__ fpart) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x)
        fpart[u] = cpart[map[u]];
}

template<int W>
void launch_row_small(Ctx &c, const DGraph &g, const u32 *map, const u32 *moff, const u32 *mem, const u32 *uboff,
                      const u32 *list, u32 len, u32 *tkey, s32 *tw, u32 *cnt, u32 *overflow) {
    if (!len) return;
    const dim3 grid(grid_for(c, len, CB / W, 8)), block(CB);
    if (g.ew)
        launch(c, "contract_rows", 0.0, grid, block, 0, row_small_kernel<W, false>, g, map, moff, mem, uboff,
               list, len, tkey, tw, cnt, overflow);
    else
        launch(c, "contract_rows", 0.0, grid, block, 0, row_small_kernel<W, true>, g, map, moff, mem, uboff,
               list, len, tkey, tw, cnt, overflow);
}

} // namespace

CoarseGraph contract(Ctx &c, const DGraph &fine, const GraphMeta &meta, const u32 *map, u32 nc) {
    (void) meta;
    CoarseGraph out;
    out.front_depth_before = c.arena.front_depth();
    const u32 n = fine.n;
    // long-lived coarse arrays (front stack)
    s32 *cvw = (s32 *) c.arena.push_front(4ull * (nc ? nc : 1));
    u32 *cxadj = (u32 *) c.arena.push_front(4ull * (nc + 1));
    BackScope bs(c.arena);
    u32 *mcnt = bs.get<u32>(nc + 1);
    u32 *moff = bs.get<u32>(nc + 1);
    u32 *mfill = bs.get<u32>(nc + 1);
    u32 *mem = bs.get<u32>(n);
    auto *cw = bs.get<unsigned long long>(nc + 1);
    u32 *ub = bs.get<u32>(nc + 1);
    u32 *uboff = bs.get<u32>(nc + 1);
    u32 *cnt = bs.get<u32>(nc + 1);
    u32 *scal = bs.get<u32>(16);  // [0..5] tier counts, [8] overflow
    fill<u32>(c, mcnt, nc + 1, 0u);
    fill<u32>(c, mfill, nc + 1, 0u);
    fill<unsigned long long>(c, cw, nc + 1, 0ull);
    fill<u32>(c, ub, nc + 1, 0u);
    fill<u32>(c, scal, 16, 0u);
    const dim3 blk(256);
    const dim3 gn(grid_for(c, n, 256)), gc(grid_for(c, nc, 256));
    const double read_bytes = 4.0 * (n + 1) + 4.0 * fine.m * (fine.ew ? 2 : 1) + 4.0 * n + (fine.vw ? 4.0 * n : 0.0);
    if (fine.vw)
        launch(c, "contract_members", 0.0, gn, blk, 0, member_count_kernel<false>, fine, map, mcnt, cw, ub);
    else
        launch(c, "contract_members", 0.0, gn, blk, 0, member_count_kernel<true>, fine, map, mcnt, cw, ub);
    exclusive_sum<u32>(c, "contract_members", mcnt, moff, nc);
    launch(c, "contract_members", 0.0, gn, blk, 0, member_scatter_kernel, n, map, (const u32 *) moff, mfill, mem);
    launch(c, "contract_members", 0.0, gc, blk, 0, cw_finish_kernel, nc, (const unsigned long long *) cw, cvw,
           scal + 8);
    exclusive_sum<u32>(c, "contract_offsets", ub, uboff, nc);
    // staging rows
    const u32 mtot = fine.m;
    u32 *tkey = bs.get<u32>(mtot);
    s32 *tw = bs.get<s32>(mtot);
    const u32 stride = nc;
    u32 *tier_lists = bs.get<u32>((u64) 6 * stride);
    launch(c, "contract_classify", 0.0, dim3(grid_for(c, nc, 256, 2)), blk, 0, classify_kernel, nc, (const u32 *) ub,
           scal, tier_lists, stride);
    u32 tc[6];
    readback(c, tc, scal, 6);
    launch_row_small<8>(c, fine, map, moff, mem, uboff, tier_lists + 0ull * stride, tc[0], tkey, tw, cnt, scal + 8);
    launch_row_small<16>(c, fine, map, moff, mem, uboff, tier_lists + 1ull * stride, tc[1], tkey, tw, cnt, scal + 8);
    launch_row_small<32>(c, fine, map, moff, mem, uboff, tier_lists + 2ull * stride, tc[2], tkey, tw, cnt, scal + 8);
    launch_row_small<64>(c, fine, map, moff, mem, uboff, tier_lists + 3ull * stride, tc[3], tkey, tw, cnt, scal + 8);
    if (tc[4]) {
        if (fine.ew)
            launch(c, "contract_rows", 0.0, dim3(tc[4]), dim3(CB), 0, row_block_kernel<false>, fine, map,
                   (const u32 *) moff, (const u32 *) mem, (const u32 *) uboff, (const u32 *) ub,
                   (const u32 *) (tier_lists + 4ull * stride), tkey, tw, cnt, scal + 8);
        else
            launch(c, "contract_rows", 0.0, dim3(tc[4]), dim3(CB), 0, row_block_kernel<true>, fine, map,
                   (const u32 *) moff, (const u32 *) mem, (const u32 *) uboff, (const u32 *) ub,
                   (const u32 *) (tier_lists + 4ull * stride), tkey, tw, cnt, scal + 8);
    }
    if (tc[5]) {
        const u32 nh = tc[5];
        const u32 *hl = tier_lists + 5ull * stride;
        BackScope hs(c.arena);
        u32 *hub_ub = hs.get<u32>(nh + 1);
        u32 *hoff = hs.get<u32>(nh + 1);
        launch(c, "contract_hubs", 0.0, dim3(grid_for(c, nh, 256)), blk, 0, hub_offsets_kernel, hl, nh,
               (const u32 *) ub, hub_ub);
        exclusive_sum<u32>(c, "contract_hubs", hub_ub, hoff, nh);
        u32 tot;
        readback(c, &tot, hoff + nh, 1);
        u64 *keys = hs.get<u64>(tot);
        u32 *vals = hs.get<u32>(tot);
        u64 *k2 = hs.get<u64>(tot);
        u32 *v2 = hs.get<u32>(tot);
        launch(c, "contract_hubs", 0.0, dim3(nh), dim3(CB), 0, hub_gather_kernel, fine, map, (const u32 *) moff,
               (const u32 *) mem, hl, (const u32 *) hoff, nh, keys, vals);
        const bool flip = radix_sort_pairs<u64>(c, keys, vals, k2, v2, tot, 32 + bits_for(nh));
        launch(c, "contract_hubs", 0.0, dim3(nh), dim3(CB), 0, hub_merge_kernel, (const u64 *) (flip ? k2 : keys),
               (const u32 *) (flip ? v2 : vals), hl, (const u32 *) hoff, nh, (const u32 *) uboff, tkey, tw, cnt,
               scal + 8);
    }
    exclusive_sum<u32>(c, "contract_offsets", cnt, cxadj, nc);
    u32 hm[2];
    copy_d2h(c, c.pinned, cxadj + nc, 4);
    copy_d2h(c, (char *) c.pinned + 4, scal + 8, 4);
    c.sync();
    std::memcpy(hm, c.pinned, 8);
    const u32 mc = hm[0];
    if (mc > fine.m) throw Error(HEIPA_ERR_DEVICE, "contraction produced more coarse edges than fine edges");
    // row building: member adjacencies 4m(1+w), coarse ids of the neighbours
    // (distinct fine vertices, at most 4n), member offsets and xadj 12n,
    // per-row metadata 16nc, unique entries written 8mc
    prof_bytes(c, "contract_rows", 4.0 * fine.m * (fine.ew ? 2 : 1) + std::min(4.0 * fine.m, 4.0 * n) + 12.0 * n +
                                       16.0 * nc + 8.0 * mc);
    if (hm[1]) throw Error(HEIPA_ERR_OVERFLOW, "coarse weight exceeds the s32 range of the graph format");
    u32 *cadj = (u32 *) c.arena.push_front(4ull * (mc ? mc : 1));
    s32 *cew = (s32 *) c.arena.push_front(4ull * (mc ? mc : 1));
    (void) read_bytes;
    launch(c, "contract_copy", 16.0 * mc + 8.0 * (nc + 1), dim3(grid_for(c, nc, CB / 16, 8)), dim3(CB), 0,
           copy_rows_kernel, nc, (const u32 *) uboff, (const u32 *) cxadj, (const u32 *) tkey, (const s32 *) tw, cadj,
           cew, scal + 9);
    u32 bad;
    readback(c, &bad, scal + 9, 1);
    if (bad) throw Error(HEIPA_ERR_DEVICE, "contraction produced an inconsistent coarse graph");
    out.g.n = nc;
    out.g.m = mc;
    out.g.gw = fine.gw;
    out.g.xadj = cxadj;
    out.g.adj = cadj;
    out.g.ew = cew;
    out.g.vw = cvw;
    return out;
}

void project(Ctx &c, const u32 *map, u32 fine_n, const u32 *cpart, u32 *fpart) {
    launch(c, "project", 12.0 * fine_n, dim3(grid_for(c, fine_n, 256 * 4)), dim3(256), 0, project_kernel, map, fine_n,
           cpart, fpart);
}

} // namespace heipa
