// Evaluators: edge cut and J(C,D,Pi), block weights, and graph metadata.
//
// edge cut  = sum over directed (u,v) of w * [P(u) != P(v)], halved
//             (src/utility/edge_cut.h:71-90)
// J         = sum over directed (u,v) of w * D[P(u)][P(v)], not halved
//             (src/utility/comm_cost.h:36-56)
// Both accumulate in 64 bits (the reference's s32 sums overflow on large
// weighted graphs).  One streaming pass over the CSR; the reference's COO
// source array edges_u is not needed: G lanes walk one vertex's adjacency,
// hubs get a workgroup each.  D lives in LDS when k*k <= 16384.
//
// Algorithmic bytes per launch: 4(n+1) + 4m(1+w) + 4n.
#include "eval.hpp"

namespace heipa {

// one update per workgroup over a small grid (per-wave updates of one address serialise)
__global__ __launch_bounds__(256) void max_degree_kernel(const u32 *xadj, u32 n, u32 *out) {
    __shared__ u32 red[256 / WAVE];
    u32 mx = 0;
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        u32 d = xadj[u + 1] - xadj[u];
        mx = d > mx ? d : mx;
    }
    mx = wave_max(mx);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (u32 w = 1; w < 256 / WAVE; ++w) mx = red[w] > mx ? red[w] : mx;
        atomicMax(out, mx);
    }
}

__global__ void hub_count_kernel(const u32 *xadj, u32 n, u32 hub_deg, u32 *count, u32 *list) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        if (xadj[u + 1] - xadj[u] > hub_deg) {
            u32 p = atomicAdd(count, 1u);
            if (list) list[p] = (u32) u;
        }
    }
}

GraphMeta make_meta(Ctx &c, const DGraph &g, bool back) {
    GraphMeta meta;
    const double avg = g.n ? (double) g.m / (double) g.n : 0.0;
    meta.G = group_for_avg(avg);
    meta.hub_deg = std::max<u32>(512, 32 * meta.G);
    if (g.n == 0) return meta;
    u32 *scr;
    {
        BackScope bs(c.arena);
        scr = bs.get<u32>(2);
        fill<u32>(c, scr, 2, 0u);
        launch(c, "graph_meta", 4.0 * (g.n + 1), dim3(grid_for(c, g.n, 256, 2)), dim3(256), 0, max_degree_kernel,
               g.xadj, g.n, scr);
        launch(c, "graph_meta", 4.0 * (g.n + 1), dim3(grid_for(c, g.n, 256)), dim3(256), 0, hub_count_kernel,
               g.xadj, g.n, meta.hub_deg, scr + 1, (u32 *) nullptr);
        u32 h[2];
        readback(c, h, scr, 2);
        meta.max_deg = h[0];
        meta.n_hubs = h[1];
    }
    if (meta.n_hubs) {
        meta.hubs = (u32 *) (back ? c.arena.push_back(4ull * (meta.n_hubs + 1)) : c.arena.push_front(4ull * (meta.n_hubs + 1)));
        u32 *cnt = meta.hubs + meta.n_hubs;
        fill<u32>(c, cnt, 1, 0u);
        launch(c, "graph_meta", 4.0 * (g.n + 1), dim3(grid_for(c, g.n, 256)), dim3(256), 0, hub_count_kernel,
               g.xadj, g.n, meta.hub_deg, cnt, meta.hubs);
    }
    return meta;
}

namespace {

constexpr int EVAL_BLOCK = 256;
constexpr u32 D_LDS_MAX = 16384;  // k*k entries kept in LDS

// MODE 0: edge cut, MODE 1: J with D
template<int G, bool UEW, int MODE, bool D_LDS>
__global__ __launch_bounds__(EVAL_BLOCK) void eval_kernel(DGraph g, const u32 *__restrict__ part,
                                                          const s32 *__restrict__ D, u32 k, u32 hub_deg,
                                                          unsigned long long *out) {
    extern __shared__ s32 sD[];
    const s32 *Dp = D;
    if constexpr (MODE == 1 && D_LDS) {
        for (u32 i = threadIdx.x; i < k * k; i += blockDim.x) sD[i] = D[i];
        __syncthreads();
        Dp = sD;
    }
    const u32 lg = threadIdx.x % G;
    const u64 groups = (u64) gridDim.x * (EVAL_BLOCK / G);
    s64 acc = 0;
    for (u64 u = (u64) blockIdx.x * (EVAL_BLOCK / G) + threadIdx.x / G; u < g.n; u += groups) {
        const u32 b = g.xadj[u], e = g.xadj[u + 1];
        if (e - b > hub_deg) continue;
        const u32 pu = part[u];
        for (u32 j = b + lg; j < e; j += G) {
            const u32 pv = part[g.adj[j]];
            const s64 w = UEW ? 1 : (s64) g.ew[j];
            if constexpr (MODE == 0) acc += pv != pu ? w : 0;
            else acc += w * (s64) Dp[(u64) pu * k + pv];
        }
    }
    __shared__ s64 red[EVAL_BLOCK / WAVE];
    acc = block_sum<EVAL_BLOCK>(acc, red);
    if (threadIdx.x == 0 && acc != 0) atomicAdd(out, (unsigned long long) acc);
}

template<bool UEW, int MODE>
__global__ __launch_bounds__(EVAL_BLOCK) void eval_hub_kernel(DGraph g, const u32 *__restrict__ part,
                                                              const s32 *__restrict__ D, u32 k,
                                                              const u32 *__restrict__ hubs,
                                                              unsigned long long *out) {
    const u32 u = hubs[blockIdx.x];
    const u32 b = g.xadj[u], e = g.xadj[u + 1];
    const u32 pu = part[u];
    s64 acc = 0;
    for (u32 j = b + threadIdx.x; j < e; j += EVAL_BLOCK) {
        const u32 pv = part[g.adj[j]];
        const s64 w = UEW ? 1 : (s64) g.ew[j];
        if constexpr (MODE == 0) acc += pv != pu ? w : 0;
        else acc += w * (s64) D[(u64) pu * k + pv];
    }
    __shared__ s64 red[EVAL_BLOCK / WAVE];
    acc = block_sum<EVAL_BLOCK>(acc, red);
    if (threadIdx.x == 0 && acc != 0) atomicAdd(out, (unsigned long long) acc);
}

template<int G>
void eval_launch(Ctx &c, const DGraph &g, const GraphMeta &meta, const u32 *part, const s32 *D, u32 k, int mode,
                 unsigned long long *acc) {
    const bool uew = g.ew == nullptr;
    const double bytes = 4.0 * (g.n + 1) + 4.0 * g.m * (uew ? 1 : 2) + 4.0 * g.n;
    const u32 per_block = EVAL_BLOCK / G;
    const dim3 grid(grid_for(c, g.n, per_block, 8)), block(EVAL_BLOCK);
    const bool dl = mode == 1 && (u64) k * k <= D_LDS_MAX;
    const size_t shmem = dl ? sizeof(s32) * k * k : 0;
    const char *name = mode == 0 ? "edge_cut" : "comm_cost";
#define EV(UEWv, MODEv, DLv) launch(c, name, bytes, grid, block, shmem, eval_kernel<G, UEWv, MODEv, DLv>, g, part, D, k, meta.hub_deg, acc)
    if (mode == 0) {
        if (uew) EV(true, 0, false);
        else EV(false, 0, false);
    } else if (dl) {
        if (uew) EV(true, 1, true);
        else EV(false, 1, true);
    } else {
        if (uew) EV(true, 1, false);
        else EV(false, 1, false);
    }
#undef EV
    if (meta.n_hubs) {
        if (mode == 0) {
            if (uew) launch(c, name, 0.0, dim3(meta.n_hubs), block, 0, eval_hub_kernel<true, 0>, g, part, D, k, (const u32 *) meta.hubs, acc);
            else launch(c, name, 0.0, dim3(meta.n_hubs), block, 0, eval_hub_kernel<false, 0>, g, part, D, k, (const u32 *) meta.hubs, acc);
        } else {
            if (uew) launch(c, name, 0.0, dim3(meta.n_hubs), block, 0, eval_hub_kernel<true, 1>, g, part, D, k, (const u32 *) meta.hubs, acc);
            else launch(c, name, 0.0, dim3(meta.n_hubs), block, 0, eval_hub_kernel<false, 1>, g, part, D, k, (const u32 *) meta.hubs, acc);
        }
    }
}

// block weights: LDS histogram per workgroup, then global 64-bit atomics
template<bool UVW>
__global__ __launch_bounds__(256) void block_weights_kernel(u32 n, const u32 *__restrict__ part,
                                                            const s32 *__restrict__ vw, u32 k,
                                                            unsigned long long *bw) {
    extern __shared__ unsigned long long hist[];
    for (u32 i = threadIdx.x; i < k; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x)
        atomicAdd(&hist[part[u]], (unsigned long long) (UVW ? 1 : vw[u]));
    __syncthreads();
    for (u32 i = threadIdx.x; i < k; i += blockDim.x)
        if (hist[i]) atomicAdd(&bw[i], hist[i]);
}

// block weights for k beyond the LDS histogram: global 64-bit atomics
template<bool UVW>
__global__ __launch_bounds__(256) void block_weights_global_kernel(u32 n, const u32 *__restrict__ part,
                                                                   const s32 *__restrict__ vw,
                                                                   unsigned long long *bw) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        const s64 w = UVW ? 1 : vw[u];
        if (w) atomicAdd(&bw[part[u]], (unsigned long long) w);
    }
}

// input validation: xadj monotone within [0, m], neighbour ids < n, edge
// weights > 0, vertex weights >= 0, no self loops; flags bit 0 = bad offsets,
// 1 = bad id, 2 = bad edge weight, 3 = bad vertex weight, 4 = self loop.
// The self-loop scan runs here: a thread scans an adjacency of up to
// VAL_SMALL entries, longer ones are listed in LDS and scanned by a whole wave
// each (a hub's adjacency on one thread would serialise the launch).
constexpr u32 VAL_SMALL = 32;
__global__ __launch_bounds__(256) void validate_vertices_kernel(DGraph g, u32 *flags) {
    __shared__ u32 s_big[256];
    __shared__ u32 s_nbig;
    for (u64 u0 = (u64) blockIdx.x * blockDim.x; u0 < g.n; u0 += (u64) gridDim.x * blockDim.x) {
        if (threadIdx.x == 0) s_nbig = 0;
        __syncthreads();
        const u64 u = u0 + threadIdx.x;
        u32 f = 0;
        if (u < g.n) {
            const u32 b = g.xadj[u], e = g.xadj[u + 1];
            if (b > e || e > g.m) {
                f |= 1u;
            } else if (e - b <= VAL_SMALL) {
                for (u32 j = b; j < e; ++j) f |= g.adj[j] == (u32) u ? 16u : 0u;
            } else {
                s_big[atomicAdd(&s_nbig, 1u)] = (u32) u;
            }
            if (g.vw && g.vw[u] < 0) f |= 8u;
        }
        __syncthreads();
        const u32 nb = s_nbig;
        for (u32 q = threadIdx.x / WAVE; q < nb; q += blockDim.x / WAVE) {
            const u32 v = s_big[q];
            const u32 e = g.xadj[v + 1];
            for (u32 j = g.xadj[v] + lane_id(); j < e; j += WAVE) f |= g.adj[j] == v ? 16u : 0u;
        }
        if (f) atomicOr(flags, f);
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (g.xadj[0] != 0 || g.xadj[g.n] != g.m)) atomicOr(flags, 1u);
}

__global__ void validate_edges_kernel(DGraph g, u32 *flags) {
    u32 f = 0;
    for (u64 j = (u64) blockIdx.x * blockDim.x + threadIdx.x; j < g.m; j += (u64) gridDim.x * blockDim.x) {
        f |= g.adj[j] >= g.n ? 2u : 0u;
        if (g.ew && g.ew[j] <= 0) f |= 4u;
    }
    if (f) atomicOr(flags, f);
}

// one update per workgroup over a small grid
__global__ __launch_bounds__(256) void max_part_kernel(const u32 *part, u32 n, u32 *out) {
    __shared__ u32 red[256 / WAVE];
    u32 mx = 0;
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x)
        mx = part[u] > mx ? part[u] : mx;
    mx = wave_max(mx);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (u32 w = 1; w < 256 / WAVE; ++w) mx = red[w] > mx ? red[w] : mx;
        atomicMax(out, mx);
    }
}

} // namespace

void validate_graph(Ctx &c, const DGraph &g) {
    BackScope bs(c.arena);
    u32 *flags = bs.get<u32>(1);
    fill<u32>(c, flags, 1, 0u);
    launch(c, "validate", 4.0 * (g.n + 1), dim3(grid_for(c, g.n + 1, 256)), dim3(256), 0, validate_vertices_kernel, g,
           flags);
    if (g.m)
        launch(c, "validate", 4.0 * g.m * (g.ew ? 2 : 1), dim3(grid_for(c, g.m, 256)), dim3(256), 0,
               validate_edges_kernel, g, flags);
    u32 f;
    readback(c, &f, flags, 1);
    if (f & 1u) throw Error(HEIPA_ERR_INVALID, "invalid graph: xadj is not a monotone offset array ending at m");
    if (f & 2u) throw Error(HEIPA_ERR_INVALID, "invalid graph: neighbour id >= n");
    if (f & 4u) throw Error(HEIPA_ERR_INVALID, "invalid graph: edge weights must be positive");
    if (f & 8u) throw Error(HEIPA_ERR_INVALID, "invalid graph: vertex weights must be non-negative");
    if (f & 16u) throw Error(HEIPA_ERR_INVALID, "invalid graph: self loops are not allowed");
}

void validate_part(Ctx &c, const u32 *part, u32 n, u32 k) {
    if (n == 0) return;
    BackScope bs(c.arena);
    u32 *mx = bs.get<u32>(1);
    fill<u32>(c, mx, 1, 0u);
    launch(c, "validate", 4.0 * n, dim3(grid_for(c, n, 256, 2)), dim3(256), 0, max_part_kernel, part, n, mx);
    u32 h;
    readback(c, &h, mx, 1);
    if (h >= k) throw Error(HEIPA_ERR_INVALID, "partition holds a block id >= k");
}

s64 edge_cut(Ctx &c, const DGraph &g, const GraphMeta &meta, const u32 *part) {
    BackScope bs(c.arena);
    auto *acc = bs.get<unsigned long long>(1);
    fill<unsigned long long>(c, acc, 1, 0ull);
    HEIPA_DISPATCH_G(meta.G, eval_launch, c, g, meta, part, (const s32 *) nullptr, 0u, 0, acc);
    unsigned long long h;
    readback(c, &h, acc, 1);
    return (s64) h / 2;
}

s64 comm_cost(Ctx &c, const DGraph &g, const GraphMeta &meta, const u32 *part, const s32 *D, u32 k) {
    BackScope bs(c.arena);
    auto *acc = bs.get<unsigned long long>(1);
    fill<unsigned long long>(c, acc, 1, 0ull);
    HEIPA_DISPATCH_G(meta.G, eval_launch, c, g, meta, part, D, k, 1, acc);
    unsigned long long h;
    readback(c, &h, acc, 1);
    return (s64) h;
}

void block_weights_dev(Ctx &c, const DGraph &g, const u32 *part, u32 k, s64 *bw_dev) {
    fill<s64>(c, bw_dev, k, 0);
    const double bytes = 4.0 * g.n * (g.vw ? 2 : 1);
    const size_t shmem = sizeof(unsigned long long) * k;
    if (shmem > 64 * 1024) {  // no LDS histogram: one global atomic per vertex
        if (g.vw)
            launch(c, "block_weights", bytes, dim3(grid_for(c, g.n, 256)), dim3(256), 0,
                   block_weights_global_kernel<false>, g.n, part, g.vw, (unsigned long long *) bw_dev);
        else
            launch(c, "block_weights", bytes, dim3(grid_for(c, g.n, 256)), dim3(256), 0,
                   block_weights_global_kernel<true>, g.n, part, g.vw, (unsigned long long *) bw_dev);
        return;
    }
    if (g.vw)
        launch(c, "block_weights", bytes, dim3(grid_for(c, g.n, 256 * 8)), dim3(256), shmem,
               block_weights_kernel<false>, g.n, part, g.vw, k, (unsigned long long *) bw_dev);
    else
        launch(c, "block_weights", bytes, dim3(grid_for(c, g.n, 256 * 8)), dim3(256), shmem,
               block_weights_kernel<true>, g.n, part, g.vw, k, (unsigned long long *) bw_dev);
}

} // namespace heipa
