// Multilevel drivers: the plain Solver level loop, the IM process-mapping
// solver and the HM recursion.
//
// IM (src/datastructures/promap_solver.h:146-347, --config IM): the same
// level loop with global multisection of the coarsest graph
// (global_multisection.h:139-287, host), the comm cost J as the tracked
// objective and the mapping form of Jet at every finer level.
//
// Solver (src/datastructures/solver.h:144-246 device ctor, :340-556): coarsen
// while n > 8k (two-hop matching with seed 12345, then contraction), partition
// the coarsest graph on the host, then project and Jet-refine level by level.
// lmax = ceil((1 + eps) * W / k) in f64 (solver.h:410).
//
// HM (src/initial_partitioning/hierarchical_multisection.h:38-307): split with
// k = hierarchy[level] under the adaptive imbalance of
// global_multisection.h:39-48, extract every non-empty block as a child graph
// (edge order preserved) and recurse; leaves write
// offset + local block into the global partition.
#include "eval.hpp"
#include "../host/host_graph.hpp"

#include <chrono>
#include <cmath>
#include <exception>
#include <system_error>
#include <thread>

namespace heipa {
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

__global__ void mapping_hash_kernel(const u32 *map, u32 n, unsigned long long *out) {
    u64 h = 0;
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x) {
        u64 z = ((u64) u << 32) | map[u];
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        h += z ^ (z >> 31);
    }
    h = wave_sum(h);
    if (lane_id() == 0 && h) atomicAdd(out, (unsigned long long) h);
}

u64 mapping_hash(Ctx &c, const u32 *map, u32 n) {
    BackScope bs(c.arena);
    auto *acc = bs.get<unsigned long long>(1);
    fill<unsigned long long>(c, acc, 1, 0ull);
    launch(c, "mapping_hash", 0.0, dim3(grid_for(c, n, 256)), dim3(256), 0, mapping_hash_kernel, map, n, acc);
    unsigned long long h;
    readback(c, &h, acc, 1);
    return h;
}

s64 max_block_weight(Ctx &c, const DGraph &g, const u32 *part, u32 k) {
    BackScope bs(c.arena);
    s64 *bw = bs.get<s64>(k);
    block_weights_dev(c, g, part, k, bw);
    std::vector<s64> h(k);
    readback(c, h.data(), bw, k);
    s64 m = 0;
    for (s64 x: h) m = std::max(m, x);
    return m;
}

// ---- HM child extraction
struct ChildSrc {
    const u32 *tmp;
    u32 id;
    __device__ u32 operator()(u64 u) const { return tmp[u] == id; }
};
struct ChildDst {
    const u32 *tmp;
    u32 id;
    u32 *loc;
    u32 *c_n_to_o;
    const u32 *n_to_o;
    s32 *c_vw;
    const s32 *vw;
    __device__ void operator()(u64 u, u32 pos) const {
        if (tmp[u] == id) {
            loc[u] = pos;
            c_n_to_o[pos] = n_to_o[u];
            c_vw[pos] = vw ? vw[u] : 1;
        }
    }
};

__global__ void child_weight_kernel(const s32 *c_vw, u32 cn, unsigned long long *out) {
    u64 s = 0;
    for (u64 i = (u64) blockIdx.x * blockDim.x + threadIdx.x; i < cn; i += (u64) gridDim.x * blockDim.x)
        s += (u64) (s64) c_vw[i];
    s = wave_sum(s);
    if (lane_id() == 0 && s) atomicAdd(out, (unsigned long long) s);
}

// per child vertex: the number of its neighbours inside the child.  A thread
// counts a vertex of degree up to CD_SMALL; larger ones are listed in LDS and
// counted by a whole wave each (a hub's adjacency on one thread serialised).
constexpr u32 CD_SMALL = 32;
__global__ __launch_bounds__(256) void child_degree_kernel(DGraph g, const u32 *tmp, u32 id,
                                                           const u32 *c_n_to_o_local, u32 cn, u32 *cdeg) {
    __shared__ u32 s_big[256];
    __shared__ u32 s_nbig;
    for (u64 i0 = (u64) blockIdx.x * blockDim.x; i0 < cn; i0 += (u64) gridDim.x * blockDim.x) {
        if (threadIdx.x == 0) s_nbig = 0;
        __syncthreads();
        const u64 i = i0 + threadIdx.x;
        if (i < cn) {
            const u32 u = c_n_to_o_local[i];
            const u32 b = g.xadj[u], e = g.xadj[u + 1];
            if (e - b <= CD_SMALL) {
                u32 d = 0;
#pragma unroll 8
                for (u32 j = b; j < e; ++j) d += tmp[g.adj[j]] == id;
                cdeg[i] = d;
            } else {
                s_big[atomicAdd(&s_nbig, 1u)] = (u32) i;
            }
        }
        __syncthreads();
        const u32 nb = s_nbig;
        for (u32 q = threadIdx.x / WAVE; q < nb; q += blockDim.x / WAVE) {
            const u32 ii = s_big[q];
            const u32 u = c_n_to_o_local[ii];
            const u32 e = g.xadj[u + 1];
            u32 d = 0;
#pragma unroll 8
            for (u32 j = g.xadj[u] + lane_id(); j < e; j += WAVE) d += tmp[g.adj[j]] == id;
            d = wave_sum(d);
            if (lane_id() == 0) cdeg[ii] = d;
        }
        __syncthreads();
    }
}

// fill child rows in adjacency order, one wave per child vertex
__global__ void child_fill_kernel(DGraph g, const u32 *tmp, u32 id, const u32 *loc, const u32 *local_ids, u32 cn,
                                  const u32 *cxadj, u32 *cadj, s32 *cew) {
    const u64 waves = (u64) gridDim.x * (blockDim.x / WAVE);
    for (u64 i = (u64) blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; i < cn; i += waves) {
        const u32 u = local_ids[i];
        u32 out = cxadj[i];
        const u32 b = g.xadj[u], e = g.xadj[u + 1];
        for (u32 j0 = b; j0 < e; j0 += WAVE) {
            const u32 j = j0 + lane_id();
            bool keep = false;
            u32 v = 0;
            if (j < e) {
                v = g.adj[j];
                keep = tmp[v] == id;
            }
            const u64 bal = __ballot(keep);
            if (keep) {
                const u32 p = out + (u32) __popcll(bal & lanemask_lt());
                cadj[p] = loc[v];
                cew[p] = g.ew ? g.ew[j] : 1;
            }
            out += (u32) __popcll(bal);
        }
    }
}

struct LocalIdDst {
    const u32 *tmp;
    u32 id;
    u32 *local_ids;
    __device__ void operator()(u64 u, u32 pos) const {
        if (tmp[u] == id) local_ids[pos] = (u32) u;
    }
};

__global__ void write_leaf_kernel(const u32 *n_to_o, const u32 *tmp, u32 n, u32 offset, u32 *global) {
    for (u64 u = (u64) blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (u64) gridDim.x * blockDim.x)
        global[n_to_o[u]] = offset + tmp[u];
}

struct HMParams {
    std::vector<u32> hier, index_vec, k_rem;
    double eps;
    s64 gW;
    u32 gk;
    u64 seed;
    bool ultra;
};

// Builds child `id` of g (the vertices with tmp == id, their internal edges
// in adjacency order) on the front stack of c's arena; false: the child is
// empty (nothing pushed).
bool build_child(Ctx &c, const DGraph &g, const u32 *tmp, const u32 *n_to_o, u32 id, DGraph &child,
                 u32 *&c_n_to_o) {
    u32 cn = 0, cm = 0;
    s64 cw = 0;
    u32 *local_ids, *cxadj, *cadj;
    s32 *c_vw, *cew;
    {
        BackScope bs(c.arena);
        u32 *loc = bs.get<u32>(g.n);
        u32 *scal = bs.get<u32>(4);
        // child size (compaction count) first
        scan<u32>(c, "hm_child_index", ChildSrc{tmp, id}, LocalIdDst{tmp, id, loc}, g.n, scal);
        readback(c, &cn, scal, 1);
        if (cn == 0) return false;
        c_vw = (s32 *) c.arena.push_front(4ull * cn);
        cxadj = (u32 *) c.arena.push_front(4ull * (cn + 1));
        c_n_to_o = (u32 *) c.arena.push_front(4ull * cn);
        local_ids = bs.get<u32>(cn);
        copy_d2d(c, local_ids, loc, 4ull * cn);
        scan<u32>(c, "hm_child_index", ChildSrc{tmp, id},
                  ChildDst{tmp, id, loc, c_n_to_o, n_to_o, c_vw, g.vw}, g.n, scal);
        auto *acc = (unsigned long long *) bs.get<unsigned long long>(1);
        fill<unsigned long long>(c, acc, 1, 0ull);
        launch(c, "hm_child_weight", 4.0 * cn, dim3(grid_for(c, cn, 256)), dim3(256), 0, child_weight_kernel,
               (const s32 *) c_vw, cn, acc);
        u32 *cdeg = bs.get<u32>(cn + 1);
        launch(c, "hm_child_degree", 0.0, dim3(grid_for(c, cn, 256)), dim3(256), 0, child_degree_kernel, g,
               (const u32 *) tmp, id, (const u32 *) local_ids, cn, cdeg);
        exclusive_sum<u32>(c, "hm_child_degree", cdeg, cxadj, cn);
        unsigned long long wsum;
        readback(c, &wsum, acc, 1);
        cw = (s64) wsum;
        readback(c, &cm, cxadj + cn, 1);
        cadj = (u32 *) c.arena.push_front(4ull * (cm ? cm : 1));
        cew = (s32 *) c.arena.push_front(4ull * (cm ? cm : 1));
        launch(c, "hm_child_fill", 0.0, dim3(grid_for(c, cn, 4, 8)), dim3(256), 0, child_fill_kernel, g,
               (const u32 *) tmp, id, (const u32 *) loc, (const u32 *) local_ids, cn, (const u32 *) cxadj, cadj,
               cew);
        // n_to_o of the child: global ids (ChildDst wrote n_to_o[u])
    }
    child.n = cn;
    child.m = cm;
    child.gw = cw;
    child.xadj = cxadj;
    child.adj = cadj;
    child.ew = cew;
    child.vw = c_vw;
    return true;
}

void hm_rec(Ctx &c, const DGraph &g, const u32 *n_to_o, u32 level, std::vector<u32> &ident, const HMParams &p,
            u32 *global);

// Arena bytes of one child subtree's solve (the pool's job estimate for a
// multisection of what remains of the hierarchy).
u64 hm_child_bytes(const DGraph &g, const HMParams &p, u32 level) {
    u64 k = 1;
    for (u32 i = 0; i <= level; ++i) k = std::max<u64>(k, p.hier[i]);
    const u64 n = g.n, m = g.m;
    const u64 sections = std::max<u64>(1, std::min<u64>(128, n / (4096 * k) + 1));
    // the pool's measured envelope for an HM job (a worker that runs out
    // hands its children back to the calling context)
    return 80 * n + 88 * m + 50ull * k * sections * 32 + (2ull << 20);
}

// The subtrees below the first split are independent (disjoint vertices,
// disjoint entries of the output, deterministic solves), so they run at the
// same time on worker contexts, one host thread each; the result does not
// depend on their order.  The children are built here first, on c's
// stream; a worker context that cannot be opened leaves the caller on the
// sequential path (false, nothing changed but the children built).
bool hm_split_parallel(Ctx &c, const DGraph &g, const u32 *tmp, const u32 *n_to_o, u32 level, const HMParams &p,
                       u32 *global) {
    struct Kid {
        DGraph g;
        u32 *n_to_o;
        u32 id;
    };
    const u32 k = p.hier[level];
    // In a pool (one subtree worker per job), the subtrees run on the pool's
    // idle worker contexts beside this one; with none idle, on this one alone.
    std::vector<Ctx *> lent;
    const bool lending = c.hm_workers <= 1 && c.borrow;
    if (lending) {
        lent = c.borrow(k - 1);
        if (lent.empty()) return false;
    }
    struct GiveBack {
        Ctx &c;
        std::vector<Ctx *> &lent;
        ~GiveBack() {
            if (!lent.empty()) c.give_back(lent);
        }
    } give_back{c, lent};
    std::vector<Kid> kids;
    for (u32 id = 0; id < k; ++id) {
        Kid kd{};
        kd.id = id;
        if (build_child(c, g, tmp, n_to_o, id, kd.g, kd.n_to_o)) kids.push_back(kd);
    }
    auto sequential = [&] {
        for (const Kid &kd: kids) {
            std::vector<u32> ident{kd.id};
            hm_rec(c, kd.g, kd.n_to_o, level - 1, ident, p, global);
        }
        return true;
    };
    if (kids.size() < 2) return sequential();
    c.sync();  // the children are complete before other streams read them
    // W: the contexts that solve the subtrees, kid i on W[i mod nw]
    std::vector<Ctx *> W;
    if (lending) {
        W.push_back(&c);
        for (Ctx *x: lent) W.push_back(x);
        while (W.size() > kids.size()) {  // more lent than children: the rest go back now
            c.give_back({W.back()});
            lent.pop_back();
            W.pop_back();
        }
    } else {
        const size_t nw = std::min<size_t>(kids.size(), c.hm_workers);
        u64 need = 0;
        for (const Kid &kd: kids) need = std::max(need, hm_child_bytes(kd.g, p, level - 1));
        if (c.hm_worker_bytes) need = c.hm_worker_bytes;
        try {
            if (c.workers.size() < nw) c.workers.resize(nw);
            for (size_t i = 0; i < nw; ++i) {
                auto &w = c.workers[i];
                if (w && w->arena.capacity() >= need) continue;
                if (w) ctx_close(*w);
                w = std::make_unique<Ctx>();
                ctx_open(*w, c.device, need);
            }
        } catch (const std::exception &) {  // device or host memory for the workers: one after another here
            for (auto &w: c.workers)
                if (w) ctx_close(*w);
            c.workers.clear();
            return sequential();
        }
        for (size_t i = 0; i < nw; ++i) W.push_back(c.workers[i].get());
    }
    const size_t nw = W.size();
    // each worker's children one after another on its own stream
    std::vector<std::exception_ptr> errors(nw);
    std::vector<u8> done(kids.size(), 0);
    auto run = [&](size_t wi) {
        Ctx &w = *W[wi];
        const size_t fd = w.arena.front_depth(), bd = w.arena.back_depth();
        try {
            HIP_CHECK(hipSetDevice(w.device));
            if (&w != &c) ctx_copy_settings(w, c);
            for (size_t i = wi; i < kids.size(); i += nw) {
                std::vector<u32> ident{kids[i].id};
                hm_rec(w, kids[i].g, kids[i].n_to_o, level - 1, ident, p, global);
                w.sync();
                done[i] = 1;
            }
        } catch (...) {
            errors[wi] = std::current_exception();
            // the worker is reused by later calls: leave its arena as it was
            (void) hipStreamSynchronize(w.stream);
            w.arena.release_back(bd);
            w.arena.release_front(fd);
        }
    };
    std::vector<std::thread> threads;
    for (size_t wi = lending ? 1 : 0; wi < nw; ++wi) {
        try {
            threads.emplace_back(run, wi);
        } catch (const std::system_error &) {
            run(wi);  // no thread available: this worker's children run here
        }
    }
    if (lending) run(0);  // this context's own share, on the calling thread
    for (auto &t: threads) t.join();
    for (size_t wi = 0; wi < nw; ++wi)
        if (W[wi] != &c) ctx_merge_prof(c, *W[wi]);
    HIP_CHECK(hipSetDevice(c.device));
    // A worker whose arena (sized by the estimate) ran out hands its
    // unfinished children back: they run one after another on c, whose arena
    // is the caller's.  A child solved again rewrites the same output entries
    // with the same values.  Any other error is the call's.
    for (size_t wi = 0; wi < nw; ++wi) {
        if (!errors[wi]) continue;
        try {
            std::rethrow_exception(errors[wi]);
        } catch (const Error &e) {
            if (e.code != HEIPA_ERR_OOM || W[wi] == &c) throw;
        }
        for (size_t i = wi; i < kids.size(); i += nw) {
            if (done[i]) continue;
            std::vector<u32> ident{kids[i].id};
            hm_rec(c, kids[i].g, kids[i].n_to_o, level - 1, ident, p, global);
        }
    }
    return true;
}

void hm_rec(Ctx &c, const DGraph &g, const u32 *n_to_o, u32 level, std::vector<u32> &ident, const HMParams &p,
            u32 *global) {
    const u32 l = (u32) p.hier.size();
    const u32 k = p.hier[level];
    const size_t depth0 = c.arena.front_depth();
    u32 *tmp = (u32 *) c.arena.push_front(4ull * (g.n ? g.n : 1));
    const double imb = adaptive_imbalance(p.eps, p.gW, p.gk, g.gw, p.k_rem[l - 1 - ident.size()], l - ident.size());
    if (k == 1) fill<u32>(c, tmp, g.n, 0u);
    else solve(c, g, k, imb, p.seed, p.ultra, tmp, nullptr);
    c.hkey = 0;  // the caller's graph key names the input graph only, not the children
    if (ident.size() == l - 1) {
        u32 offset = 0;
        for (u32 i = 0; i + 1 < l; ++i) offset += ident[i] * p.index_vec[p.index_vec.size() - 1 - i];
        launch(c, "hm_write_leaf", 12.0 * g.n, dim3(grid_for(c, g.n, 256)), dim3(256), 0, write_leaf_kernel, n_to_o,
               (const u32 *) tmp, g.n, offset, global);
        c.arena.release_front(depth0);
        return;
    }
    if (ident.empty() && level > 0 && (c.hm_workers > 1 || c.borrow) &&
        hm_split_parallel(c, g, tmp, n_to_o, level, p, global)) {
        c.arena.release_front(depth0);
        return;
    }
    for (u32 id = 0; id < k; ++id) {
        const size_t cdepth = c.arena.front_depth();
        DGraph child;
        u32 *c_n_to_o = nullptr;
        if (!build_child(c, g, tmp, n_to_o, id, child, c_n_to_o)) continue;
        ident.push_back(id);
        hm_rec(c, child, c_n_to_o, level - 1, ident, p, global);
        ident.pop_back();
        c.arena.release_front(cdepth);
    }
    c.arena.release_front(depth0);
}

// ---- resident hierarchies (HierCache): the entry of the context's graph key
std::shared_ptr<HierEntry> hier_entry(Ctx &c, const DGraph &g0) {
    if (!c.hcache || !c.hkey) return nullptr;
    HierCache &hc = *c.hcache;
    std::lock_guard<std::mutex> lk(hc.m);
    std::shared_ptr<HierEntry> &e = hc.entries[c.hkey];
    if (!e || e->n0 != g0.n || e->m0 != g0.m || e->gw0 != g0.gw) {  // a new graph under this key
        if (e) hc.bytes -= e->bytes;
        e = std::make_shared<HierEntry>();
        e->n0 = g0.n;
        e->m0 = g0.m;
        e->gw0 = g0.gw;
    }
    e->last_use = ++hc.clock;
    return e;
}

bool hier_level(HierEntry &e, size_t l, HierLevel &out) {
    std::lock_guard<std::mutex> lk(e.m);
    if (l >= e.levels.size()) return false;
    out = e.levels[l];
    return true;
}

// Copies coarsening step l (the map of level l, the graph of level l + 1)
// into device memory the cache owns, unless another solve appended it first;
// then evicts least recently used entries no solve holds while the cache is
// over its budget.  No device memory for it: the step is not kept.
void hier_append(Ctx &c, HierEntry &e, size_t l, u32 fine_n, const u32 *map, const DGraph &cg) {
    {
        std::lock_guard<std::mutex> lk(e.m);
        if (e.levels.size() != l) return;
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t) 255; };
    const size_t b_map = al(4ull * fine_n), b_x = al(4ull * (cg.n + 1)), b_a = al(4ull * cg.m);
    const size_t b_v = al(4ull * cg.n);
    const size_t nb = b_map + b_x + 2 * b_a + b_v;
    void *mem = nullptr;
    if (hipMalloc(&mem, nb) != hipSuccess) {
        (void) hipGetLastError();
        return;
    }
    char *p = (char *) mem;
    HierLevel lv;
    lv.fine_n = fine_n;
    lv.map = (u32 *) p;
    p += b_map;
    u32 *x = (u32 *) p;
    p += b_x;
    u32 *a = (u32 *) p;
    p += b_a;
    s32 *w = (s32 *) p;
    p += b_a;
    s32 *v = (s32 *) p;
    copy_d2d(c, lv.map, map, 4ull * fine_n);
    copy_d2d(c, x, cg.xadj, 4ull * (cg.n + 1));
    if (cg.m) copy_d2d(c, a, cg.adj, 4ull * cg.m);
    if (cg.m && cg.ew) copy_d2d(c, w, cg.ew, 4ull * cg.m);
    if (cg.vw) copy_d2d(c, v, cg.vw, 4ull * cg.n);
    c.sync();
    lv.g = cg;
    lv.g.xadj = x;
    lv.g.adj = a;
    lv.g.ew = cg.ew ? w : nullptr;
    lv.g.vw = cg.vw ? v : nullptr;
    {
        std::lock_guard<std::mutex> lk(e.m);
        if (e.levels.size() != l) {
            (void) hipFree(mem);
            return;
        }
        e.levels.push_back(lv);
        e.mem.push_back(mem);
        e.bytes += nb;
    }
    HierCache &hc = *c.hcache;
    std::vector<std::shared_ptr<HierEntry>> dropped;  // destroyed (freed) outside the lock
    {
        std::lock_guard<std::mutex> lk(hc.m);
        hc.bytes += nb;
        hc.levels_built++;
        while (hc.bytes > hc.budget) {
            auto victim = hc.entries.end();
            for (auto it = hc.entries.begin(); it != hc.entries.end(); ++it)
                if (it->second.use_count() == 1 && (victim == hc.entries.end() ||
                                                    it->second->last_use < victim->second->last_use))
                    victim = it;
            if (victim == hc.entries.end()) break;  // every entry is in use
            hc.bytes -= victim->second->bytes;
            dropped.push_back(std::move(victim->second));
            hc.entries.erase(victim);
        }
    }
}

} // namespace

SolveResult solve(Ctx &c, const DGraph &g0, u32 k, double imbalance, u64 seed, bool ultra, u32 *part,
                  std::vector<heipa_level_trace> *levels, const MapSpec *map) {
    SolveResult res;
    require(k >= 1, "k must be >= 1");
    c.prof_l0 = false;
    if (map) ultra = false;
    res.lmax = (s64) std::ceil((1.0 + imbalance) * ((double) g0.gw / (double) k));
    const size_t depth0 = c.arena.front_depth();
    const u32 n0 = g0.n;
    u32 *pa = (u32 *) c.arena.push_front(4ull * (n0 ? n0 : 1));
    u32 *pb = (u32 *) c.arena.push_front(4ull * (n0 ? n0 : 1));
    std::vector<DGraph> graphs{g0};
    std::vector<GraphMeta> metas{make_meta(c, g0, false)};
    std::vector<u32 *> maps;
    const u32 max_n = 8 * k;
    double t0 = now_ms();
    std::shared_ptr<HierEntry> ent = hier_entry(c, g0);
    while (graphs.back().n > max_n) {
        const size_t l = graphs.size() - 1;
        HierLevel lv;
        if (ent && hier_level(*ent, l, lv)) {  // a resident level of this graph
            maps.push_back(lv.map);
            graphs.push_back(lv.g);
            metas.push_back(make_meta(c, lv.g, false));
            std::lock_guard<std::mutex> lk(c.hcache->m);
            c.hcache->levels_reused++;
            continue;
        }
        const DGraph &g = graphs.back();
        u32 *map = (u32 *) c.arena.push_front(4ull * g.n);
        double ta = now_ms();
        c.prof_l0 = c.prof_on && graphs.size() == 1;
        const u32 nc = match(c, g, metas.back(), 12345u, map, nullptr);
        c.sync();
        double tb = now_ms();
        res.ms[0] += tb - ta;
        CoarseGraph cg = contract(c, g, metas.back(), map, nc);
        c.prof_l0 = false;
        const u32 fine_n = g.n;
        maps.push_back(map);
        graphs.push_back(cg.g);
        metas.push_back(make_meta(c, cg.g, false));
        c.sync();
        res.ms[1] += now_ms() - tb;
        if (ent) hier_append(c, *ent, l, fine_n, map, cg.g);
    }
    (void) t0;
    // initial partition of the coarsest graph on the host
    double ti = now_ms();
    const DGraph &cg = graphs.back();
    {
        std::vector<u32> hx(cg.n + 1), ha(cg.m), hp(cg.n);
        std::vector<s32> hw, hv;
        readback(c, hx.data(), cg.xadj, cg.n + 1);
        readback(c, ha.data(), cg.adj, cg.m);
        if (cg.ew) {
            hw.resize(cg.m);
            readback(c, hw.data(), cg.ew, cg.m);
        }
        if (cg.vw) {
            hv.resize(cg.n);
            readback(c, hv.data(), cg.vw, cg.n);
        }
        if (map)  // promap_solver.h:336-347
            global_multisection(cg.n, hx.data(), ha.data(), cg.ew ? hw.data() : nullptr, cg.vw ? hv.data() : nullptr,
                                map->hier.data(), (u32) map->hier.size(), imbalance, seed, hp.data());
        else
            initial_partition(cg.n, hx.data(), ha.data(), cg.ew ? hw.data() : nullptr, cg.vw ? hv.data() : nullptr,
                              k, imbalance, seed, hp.data());
        copy_h2d(c, pa, hp.data(), 4ull * cg.n);
        c.sync();
    }
    s64 cut = map ? comm_cost(c, cg, metas.back(), pa, map->D, k) : edge_cut(c, cg, metas.back(), pa);
    s64 maxw = max_block_weight(c, cg, pa, k);
    res.initial_cut = cut;
    res.ms[2] += now_ms() - ti;
    if (levels) {
        levels->assign(graphs.size(), heipa_level_trace{});
        for (size_t l = 0; l < graphs.size(); ++l) {
            auto &L = (*levels)[l];
            L.n = graphs[l].n;
            L.m = graphs[l].m;
            L.nc = l + 1 < graphs.size() ? graphs[l + 1].n : 0;
            if (l < maps.size()) L.mapping_hash = mapping_hash(c, maps[l], graphs[l].n);
        }
        levels->back().cut = cut;
        levels->back().maxw = maxw;
    }
    u32 *cur = pa, *other = pb;
    for (size_t l = maps.size(); l-- > 0;) {
        double ta = now_ms();
        project(c, maps[l], graphs[l].n, cur, other);
        std::swap(cur, other);
        c.sync();
        double tb = now_ms();
        res.ms[3] += tb - ta;
        c.prof_l0 = c.prof_on && l == 0;
        JetResult jr = jet_refine(c, graphs[l], metas[l], cur, k, res.lmax, ultra, (u32) l, cut, maxw, nullptr,
                                  map ? map->D : nullptr);
        c.prof_l0 = false;
        cut = jr.cut;
        maxw = jr.maxw;
        res.jet_iterations += jr.iterations;
        res.ms[4] += now_ms() - tb;
        if (levels) {
            (*levels)[l].cut = cut;
            (*levels)[l].maxw = maxw;
        }
    }
    copy_d2d(c, part, cur, 4ull * n0);
    c.sync();
    res.cut = cut;
    res.maxw = maxw;
    res.levels = (u32) graphs.size();
    c.arena.release_front(depth0);
    return res;
}

void hm_partition(Ctx &c, const DGraph &g, const std::vector<u32> &hier, double imbalance, u64 seed, bool ultra,
                  u32 *part) {
    require(!hier.empty(), "empty hierarchy");
    HMParams p;
    p.hier = hier;
    p.eps = imbalance;
    p.gW = g.gw;
    p.seed = seed;
    p.ultra = ultra;
    const u32 l = (u32) hier.size();
    p.gk = 1;
    for (u32 x: hier) p.gk *= x;
    p.index_vec = {1};  // hierarchical_multisection.h:262-263
    for (u32 i = 0; i + 1 < l; ++i) p.index_vec.push_back(p.index_vec[i] * hier[i]);
    p.k_rem.resize(l);  // :265-270
    u32 acc = 1;
    for (u32 i = 0; i < l; ++i) {
        p.k_rem[i] = acc * hier[i];
        acc *= hier[i];
    }
    const size_t depth0 = c.arena.front_depth();
    u32 *n_to_o = (u32 *) c.arena.push_front(4ull * (g.n ? g.n : 1));
    launch(c, "hm_init", 4.0 * g.n, dim3(grid_for(c, g.n, 256)), dim3(256), 0, iota_kernel<u32>, n_to_o, (u64) g.n);
    std::vector<u32> ident;
    struct KeyRestore {
        Ctx &c;
        u64 key;
        ~KeyRestore() { c.hkey = key; }
    } kr{c, c.hkey};
    hm_rec(c, g, n_to_o, l - 1, ident, p, part);
    c.sync();
    c.arena.release_front(depth0);
}

} // namespace heipa
