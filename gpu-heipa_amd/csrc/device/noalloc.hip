// Edge cut without allocation: the variant of heipa_edge_cut for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  No counterpart in the reference.
#include "eval.hpp"

#include <mutex>

namespace heipa {

// ---- edge cut without allocation (heipa_edge_cut_noalloc)
//
// The only device memory beyond the caller's inputs and output is the
// file-scope array `scratch`: one partial sum and one flag word per
// workgroup, NA_MAX_GROUPS of each (32 KiB per device, held until the process
// ends).  Every launch writes all the slots it later reads, so nothing is
// cleared first (no memset node when the caller captures the stream).  The
// graph checks of validate_graph run inside the same pass: an offset outside
// [0, m] or a neighbour id >= n is never dereferenced, and a malformed graph
// makes the result -1.  Calls that overlap on one device share `scratch` and
// race.
constexpr u32 NA_BLOCK = 256;
constexpr u32 NA_MAX_GROUPS = 2048;  // workgroups per launch = partial-sum slots
constexpr u32 NA_SMALL = 32;         // longer adjacencies are walked by a whole wave

static __device__ __align__(256) unsigned long long scratch[2 * NA_MAX_GROUPS];

namespace {

// workgroup r writes slots[r] (sum of w over cut directed edges of its
// vertices) and slots[NA_MAX_GROUPS + r] (validation flags, bit values as in
// validate_graph)
template<bool UEW>
__global__ __launch_bounds__(NA_BLOCK) void na_cut_partial_kernel(DGraph g, const u32 *__restrict__ part,
                                                                  unsigned long long *slots) {
    __shared__ u32 s_big[NA_BLOCK];
    __shared__ u32 s_nbig, s_flags;
    __shared__ s64 red[NA_BLOCK / WAVE];
    if (threadIdx.x == 0) s_flags = 0;
    s64 acc = 0;
    u32 f = 0;
    for (u64 u0 = (u64) blockIdx.x * NA_BLOCK; u0 < g.n; u0 += (u64) gridDim.x * NA_BLOCK) {
        if (threadIdx.x == 0) s_nbig = 0;
        __syncthreads();
        const u64 u = u0 + threadIdx.x;
        if (u < g.n) {
            const u32 b = g.xadj[u], e = g.xadj[u + 1];
            if (b > e || e > g.m) {
                f |= 1u;
            } else if (e - b <= NA_SMALL) {
                const u32 pu = part[u];
                for (u32 j = b; j < e; ++j) {
                    const u32 v = g.adj[j];
                    const s64 w = UEW ? 1 : (s64) g.ew[j];
                    if (v >= g.n) { f |= 2u; continue; }
                    f |= v == (u32) u ? 16u : 0u;
                    f |= w <= 0 ? 4u : 0u;
                    acc += part[v] != pu ? w : 0;
                }
            } else {
                s_big[atomicAdd(&s_nbig, 1u)] = (u32) u;
            }
            if (g.vw && g.vw[u] < 0) f |= 8u;
        }
        __syncthreads();
        const u32 nb = s_nbig;
        for (u32 q = threadIdx.x / WAVE; q < nb; q += NA_BLOCK / WAVE) {
            const u32 v0 = s_big[q];
            const u32 e = g.xadj[v0 + 1], pu = part[v0];  // b <= e <= m was checked when v0 was listed
            for (u32 j = g.xadj[v0] + lane_id(); j < e; j += WAVE) {
                const u32 v = g.adj[j];
                const s64 w = UEW ? 1 : (s64) g.ew[j];
                if (v >= g.n) { f |= 2u; continue; }
                f |= v == v0 ? 16u : 0u;
                f |= w <= 0 ? 4u : 0u;
                acc += part[v] != pu ? w : 0;
            }
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (g.xadj[0] != 0 || g.xadj[g.n] != g.m)) f |= 1u;
    if (f) atomicOr(&s_flags, f);
    acc = block_sum<NA_BLOCK>(acc, red);  // its barriers also order the LDS flag updates
    if (threadIdx.x == 0) {
        slots[blockIdx.x] = (unsigned long long) acc;
        slots[NA_MAX_GROUPS + blockIdx.x] = s_flags;
    }
}

// one workgroup: sums the `groups` partials and writes the halved cut, or -1
// when a workgroup flagged the graph
__global__ __launch_bounds__(NA_BLOCK) void na_cut_final_kernel(const unsigned long long *slots, u32 groups,
                                                                int64_t *out) {
    __shared__ s64 red[NA_BLOCK / WAVE];
    __shared__ u32 s_flags;
    if (threadIdx.x == 0) s_flags = 0;
    __syncthreads();
    s64 acc = 0;
    u32 f = 0;
    for (u32 r = threadIdx.x; r < groups; r += NA_BLOCK) {
        acc += (s64) slots[r];
        f |= (u32) slots[NA_MAX_GROUPS + r];
    }
    if (f) atomicOr(&s_flags, f);
    acc = block_sum<NA_BLOCK>(acc, red);
    if (threadIdx.x == 0) *out = s_flags ? (int64_t) -1 : acc / 2;
}

std::mutex na_mutex;
void *na_scratch[64];  // device address of `scratch` per device index

} // namespace

static int edge_cut_noalloc(const DGraph &g, const u32 *part, int64_t *cut_dev, hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return HEIPA_ERR_DEVICE;
    unsigned long long *slots;
    {
        // once per device, with that device current: the host-side name
        // `scratch` is not a device address; this is
        std::lock_guard<std::mutex> lk(na_mutex);
        if (!na_scratch[dev]) {
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return HEIPA_ERR_DEVICE;
            na_scratch[dev] = ptr;
        }
        slots = (unsigned long long *) na_scratch[dev];
    }
    const u32 groups = (u32) std::min<u64>(NA_MAX_GROUPS, ((u64) g.n + NA_BLOCK - 1) / NA_BLOCK);
    if (groups) {
        if (g.ew) hipLaunchKernelGGL(na_cut_partial_kernel<false>, dim3(groups), dim3(NA_BLOCK), 0, stream, g, part, slots);
        else hipLaunchKernelGGL(na_cut_partial_kernel<true>, dim3(groups), dim3(NA_BLOCK), 0, stream, g, part, slots);
    }
    hipLaunchKernelGGL(na_cut_final_kernel, dim3(1), dim3(NA_BLOCK), 0, stream, (const unsigned long long *) slots,
                       groups, cut_dev);
    return hipGetLastError() == hipSuccess ? HEIPA_OK : HEIPA_ERR_DEVICE;
}

} // namespace heipa

extern "C" int heipa_edge_cut_noalloc(const heipa_graph *g, const uint32_t *part, int64_t *cut_dev, void *stream) {
    using namespace heipa;
    if (!g || !cut_dev) return HEIPA_ERR_INVALID;
    if (g->n && (!g->xadj || !part || (g->m && !g->adjncy))) return HEIPA_ERR_INVALID;
    if (!(g->flags & HEIPA_UNIFORM_EW) && g->m && !g->adjwgt) return HEIPA_ERR_INVALID;
    if (!(g->flags & HEIPA_UNIFORM_VW) && g->n && !g->vwgt) return HEIPA_ERR_INVALID;
    return edge_cut_noalloc(to_dgraph(g), part, cut_dev, (hipStream_t) stream);
}
