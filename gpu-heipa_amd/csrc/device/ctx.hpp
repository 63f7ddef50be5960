// Device runtime shared by every kernel file: error type, the LIFO HBM arena,
// the per-context stream and the per-kernel event timer.
//
// The arena keeps the reference's two-ended LIFO discipline
// (src/datastructures/kokkos_memory_stack.h:36-187): the front stack holds
// long-lived objects (graphs, mappings, partitions), the back stack holds
// per-phase temporaries; both grow towards each other inside one hipMalloc.
// Exhaustion raises Error(HEIPA_ERR_OOM) instead of abort().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/heipa.h"

namespace heipa {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using s32 = int32_t;
using s64 = int64_t;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                             \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            throw ::heipa::Error(HEIPA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

inline void require(bool ok, const std::string &msg) {
    if (!ok) throw Error(HEIPA_ERR_INVALID, msg);
}

// device view of a CSR graph (heipa_graph without the C struct)
struct DGraph {
    u32 n = 0, m = 0;
    s64 gw = 0;
    const u32 *xadj = nullptr;
    const u32 *adj = nullptr;
    const s32 *ew = nullptr;  // null: uniform
    const s32 *vw = nullptr;  // null: uniform
};

inline DGraph to_dgraph(const heipa_graph *g) {
    DGraph d;
    d.n = g->n;
    d.m = g->m;
    d.gw = g->g_weight;
    d.xadj = g->xadj;
    d.adj = g->adjncy;
    d.ew = (g->flags & HEIPA_UNIFORM_EW) ? nullptr : g->adjwgt;
    d.vw = (g->flags & HEIPA_UNIFORM_VW) ? nullptr : g->vwgt;
    return d;
}

class Arena {
public:
    void init(void *base, size_t cap) {
        base_ = (char *) base;
        cap_ = cap;
        front_ = 0;
        back_ = cap;
        fstack_.clear();
        bstack_.clear();
        peak_ = 0;
    }
    void *push_front(size_t bytes) {
        size_t b = align(bytes);
        if (front_ + b > back_) oom(bytes);
        void *p = base_ + front_;
        fstack_.push_back(front_);
        front_ += b;
        track();
        return p;
    }
    void pop_front() {
        if (fstack_.empty()) throw Error(HEIPA_ERR_INVALID, "arena: pop_front on empty stack");
        front_ = fstack_.back();
        fstack_.pop_back();
    }
    void *push_back(size_t bytes) {
        size_t b = align(bytes);
        if (back_ < b || back_ - b < front_) oom(bytes);
        back_ -= b;
        bstack_.push_back(back_ + b);
        track();
        return base_ + back_;
    }
    void pop_back() {
        if (bstack_.empty()) throw Error(HEIPA_ERR_INVALID, "arena: pop_back on empty stack");
        back_ = bstack_.back();
        bstack_.pop_back();
    }
    size_t front_depth() const { return fstack_.size(); }
    size_t back_depth() const { return bstack_.size(); }
    void release_front(size_t depth) {
        while (fstack_.size() > depth) pop_front();
    }
    void release_back(size_t depth) {
        while (bstack_.size() > depth) pop_back();
    }
    size_t peak() const { return peak_; }
    void reset_peak() { peak_ = in_use(); }
    size_t capacity() const { return cap_; }
    size_t in_use() const { return front_ + (cap_ - back_); }

private:
    static size_t align(size_t b) { return (b + 255) & ~(size_t) 255; }
    void track() {
        size_t u = in_use();
        if (u > peak_) peak_ = u;
    }
    [[noreturn]] void oom(size_t bytes) const {
        throw Error(HEIPA_ERR_OOM, "arena exhausted: requested " + std::to_string(bytes) + " bytes, " +
                                       std::to_string(back_ - front_) + " free of " + std::to_string(cap_));
    }
    char *base_ = nullptr;
    size_t cap_ = 0, front_ = 0, back_ = 0, peak_ = 0;
    std::vector<size_t> fstack_, bstack_;
};

// Pops every back-stack chunk pushed during its lifetime.
struct BackScope {
    Arena &a;
    size_t depth;
    explicit BackScope(Arena &ar) : a(ar), depth(ar.back_depth()) {}
    ~BackScope() { a.release_back(depth); }
    template<typename T>
    T *get(size_t count) { return (T *) a.push_back(sizeof(T) * (count ? count : 1)); }
};

// Coarse levels of graphs the caller names by a key, kept in device memory
// between calls.  The hierarchy does not depend on k (matching always uses
// seed 12345, two_hop_matching.h:590; only the stop level n <= 8k does,
// solver.h:343-347), so a later solve of the same graph with any k reuses
// the levels it needs and appends the ones it builds beyond them.
struct HierLevel {
    DGraph g;           // level i + 1
    u32 *map = nullptr; // level i -> level i + 1, fine_n entries
    u32 fine_n = 0;
};
struct HierEntry {
    std::mutex m;
    u32 n0 = 0, m0 = 0;  // the finest graph this entry was built from
    s64 gw0 = 0;
    std::vector<HierLevel> levels;
    std::vector<void *> mem;
    size_t bytes = 0;
    u64 last_use = 0;
    ~HierEntry() {
        for (void *p: mem) (void) hipFree(p);
    }
};
struct HierCache {
    std::mutex m;
    std::map<u64, std::shared_ptr<HierEntry>> entries;
    size_t budget = 0, bytes = 0;
    u64 clock = 0;
    u64 hits = 0, levels_reused = 0, levels_built = 0;
};

struct KernelStat {
    u64 launches = 0;  // all launches
    u64 timed = 0;     // launches bracketed by events (every prof_sample-th)
    double ms = 0.0;   // event time of the timed launches
    double bytes = 0.0;
};

struct Ctx {
    int device = 0;
    int num_cus = 256;
    size_t max_lds = 65536;  // LDS bytes one workgroup may allocate
    u32 resolve_cap = ~0u;   // HEIPA_OPT_RESOLVE_CAP
    u32 block_jet_max_n = 2048;  // HEIPA_OPT_BLOCK_JET
    u32 group_cap = 64;         // HEIPA_OPT_GROUP_CAP
    u32 block_group_cap = 8;    // HEIPA_OPT_BLOCK_GROUP_CAP
    u32 region_edges = 1024;    // HEIPA_OPT_REGION_EDGES
    bool use_graphs = true;  // HEIPA_OPT_GRAPHS
    u32 bucket_copies = 0;   // HEIPA_OPT_BUCKET_COPIES (0: automatic)
    bool block_lds = true;   // HEIPA_OPT_BLOCK_LDS
    u32 hm_workers = 8;      // HEIPA_OPT_HM_WORKERS
    u64 hm_worker_bytes = 0; // HEIPA_OPT_HM_WORKER_BYTES (0: the estimate)
    bool race = false;       // HEIPA_OPT_RACE: atomic order instead of the deterministic rules
    std::shared_ptr<HierCache> hcache;  // HEIPA_OPT_HIER_CACHE (shared by a pool's workers)
    u64 hkey = 0;                       // HEIPA_OPT_GRAPH_KEY: the input graph's key (0: no reuse)
    hipStream_t stream = nullptr;
    void *arena_mem = nullptr;
    Arena arena;
    // pinned host staging for scalar readbacks
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    std::string last_error;
    // profiling
    bool prof_on = false;
    std::string prof_filter;  // record only kernels whose name starts with this (empty: all)
    u32 prof_sample = 1;      // bracket about one launch in prof_sample of a kernel with events
    bool prof_l0 = false;     // launches of a solve's finest level are recorded as "<name>@L0" as well
    std::vector<std::string> prof_names;
    std::map<std::string, size_t> prof_index;
    std::vector<KernelStat> prof_stats;
    struct Pending {
        size_t idx;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // worker contexts (own stream, arena and staging) for independent
    // subproblems run concurrently, kept between calls
    std::vector<std::unique_ptr<Ctx>> workers;
    // a pool's idle worker contexts lent to this context's HM subtree solves
    // (set by the pool for its workers; null elsewhere): borrow(max) returns
    // up to max of them, give_back returns them to the pool
    std::function<std::vector<Ctx *>(size_t)> borrow;
    std::function<void(const std::vector<Ctx *> &)> give_back;
    void *handle = nullptr;  // the C-ABI object that holds this context

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    size_t prof_slot(const char *name) {
        if (prof_l0) {  // a solve's finest level: its own entry
            std::string key = std::string(name) + "@L0";
            auto it = prof_index.find(key);
            if (it != prof_index.end()) return it->second;
            size_t i = prof_names.size();
            prof_names.push_back(key);
            prof_stats.emplace_back();
            prof_index.emplace(key, i);
            return i;
        }
        auto it = prof_index.find(name);
        if (it != prof_index.end()) return it->second;
        size_t i = prof_names.size();
        prof_names.emplace_back(name);
        prof_stats.emplace_back();
        prof_index.emplace(name, i);
        return i;
    }
    // resolve completed timing events (call after a stream sync)
    void prof_collect() {
        for (auto &p: pending) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
            auto &s = prof_stats[p.idx];
            s.timed++;
            s.ms += ms;
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
    }
    void sync() {
        HIP_CHECK(hipStreamSynchronize(stream));
        if (prof_on && !pending.empty()) prof_collect();
    }
    template<typename T>
    T *pinned_as() { return (T *) pinned; }
};

// Opens a context on `device`: a non-blocking stream, an arena of
// arena_bytes (0: a quarter of the free device memory) and pinned staging.
// On failure everything opened so far is released and the error rethrown.
inline void ctx_close(Ctx &c);
inline void ctx_open(Ctx &c, int device, uint64_t arena_bytes) {
    try {
        c.device = device;
        HIP_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        c.num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
            c.max_lds = (size_t) lds;
        HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        if (arena_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
            arena_bytes = (uint64_t) (free_b / 4);
        }
        arena_bytes = (arena_bytes + 255) & ~(uint64_t) 255;
        HIP_CHECK(hipMalloc(&c.arena_mem, arena_bytes));
        c.arena.init(c.arena_mem, arena_bytes);
        c.pinned_bytes = 1 << 18;
        HIP_CHECK(hipHostMalloc(&c.pinned, c.pinned_bytes, hipHostMallocDefault));
    } catch (...) {
        ctx_close(c);
        throw;
    }
}

// Releases a context and its workers (errors ignored: teardown).
inline void ctx_close(Ctx &c) {
    for (auto &w: c.workers)
        if (w) ctx_close(*w);
    c.workers.clear();
    (void) hipSetDevice(c.device);
    if (c.stream) (void) hipStreamSynchronize(c.stream);
    for (auto &p: c.pending) {
        (void) hipEventDestroy(p.a);
        (void) hipEventDestroy(p.b);
    }
    c.pending.clear();
    for (auto e: c.event_pool) (void) hipEventDestroy(e);
    c.event_pool.clear();
    if (c.pinned) (void) hipHostFree(c.pinned);
    if (c.arena_mem) (void) hipFree(c.arena_mem);
    if (c.stream) (void) hipStreamDestroy(c.stream);
    c.pinned = nullptr;
    c.arena_mem = nullptr;
    c.stream = nullptr;
}

// Closes the cached worker contexts beyond the first `keep` (their arenas go
// back to the device).
inline void ctx_trim_workers(Ctx &c, size_t keep) {
    for (size_t i = keep; i < c.workers.size(); ++i)
        if (c.workers[i]) ctx_close(*c.workers[i]);
    if (c.workers.size() > keep) c.workers.resize(keep);
    (void) hipSetDevice(c.device);
}

// The tuning options of `from` (results never depend on them) and its
// profiling settings, for a worker that runs part of its call.
inline void ctx_copy_settings(Ctx &to, const Ctx &from) {
    to.resolve_cap = from.resolve_cap;
    to.block_jet_max_n = from.block_jet_max_n;
    to.use_graphs = from.use_graphs;
    to.bucket_copies = from.bucket_copies;
    to.block_lds = from.block_lds;
    to.group_cap = from.group_cap;
    to.block_group_cap = from.block_group_cap;
    to.region_edges = from.region_edges;
    to.hm_workers = from.hm_workers;
    to.hm_worker_bytes = from.hm_worker_bytes;
    to.race = from.race;
    to.hcache = from.hcache;
    to.prof_on = from.prof_on;
    to.prof_filter = from.prof_filter;
    to.prof_sample = from.prof_sample;
}

// Adds a worker's kernel statistics to `to` and clears them in the worker
// (after the worker's stream was synchronised).
inline void ctx_merge_prof(Ctx &to, Ctx &w) {
    for (size_t i = 0; i < w.prof_names.size(); ++i) {
        const KernelStat &s = w.prof_stats[i];
        KernelStat &t = to.prof_stats[to.prof_slot(w.prof_names[i].c_str())];
        t.launches += s.launches;
        t.timed += s.timed;
        t.ms += s.ms;
        t.bytes += s.bytes;
    }
    w.prof_names.clear();
    w.prof_index.clear();
    w.prof_stats.clear();
}

// Adds data-dependent algorithmic bytes to a timed kernel's total (for kernels
// whose compulsory traffic is only known after they ran).
inline void prof_bytes(Ctx &c, const char *name, double bytes) {
    if (!c.prof_on) return;
    if (!c.prof_filter.empty() && std::strncmp(name, c.prof_filter.c_str(), c.prof_filter.size()) != 0) return;
    c.prof_stats[c.prof_slot(name)].bytes += bytes;
}

// Kernel launch on the context stream with optional event timing.
// `bytes` is the launch's algorithmic (compulsory) HBM traffic.
template<typename... KArgs, typename... Args>
inline void launch(Ctx &c, const char *name, double bytes, dim3 grid, dim3 block, size_t shmem,
                   void (*kernel)(KArgs...), Args... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    hipEvent_t ea = nullptr, eb = nullptr;
    bool timed = false;
    size_t slot = 0;
    if (c.prof_on && (c.prof_filter.empty() || std::strncmp(name, c.prof_filter.c_str(), c.prof_filter.size()) == 0)) {
        slot = c.prof_slot(name);
        KernelStat &st = c.prof_stats[slot];
        st.launches++;
        st.bytes += bytes;
        // about one launch in prof_sample, chosen by a hash of the launch count so
        // that periodic launch patterns (LP / rebalance alternation) do not bias it
        u64 z = st.launches * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ull;
        timed = c.prof_sample <= 1 || (z ^ (z >> 29)) % c.prof_sample == 0;
    }
    if (timed) {
        ea = c.get_event();
        eb = c.get_event();
        HIP_CHECK(hipEventRecord(ea, c.stream));
    }
    kernel<<<grid, block, shmem, c.stream>>>(static_cast<KArgs>(args)...);
    HIP_CHECK(hipGetLastError());
    if (timed) {
        HIP_CHECK(hipEventRecord(eb, c.stream));
        c.pending.push_back({slot, ea, eb});
        if (c.pending.size() > 4096) c.sync();
    }
}

inline u32 div_up(u64 a, u64 b) { return (u32) ((a + b - 1) / b); }

// grid size for a grid-stride loop over `items` with `per_block` items per block step
inline u32 grid_for(const Ctx &c, u64 items, u32 per_block, u32 max_per_cu = 16) {
    u64 g = (items + per_block - 1) / per_block;
    u64 cap = (u64) c.num_cus * max_per_cu;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (u32) g;
}

inline void memset_async(Ctx &c, void *p, int v, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemsetAsync(p, v, bytes, c.stream));
}
inline void copy_d2h(Ctx &c, void *dst, const void *src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
}
inline void copy_h2d(Ctx &c, void *dst, const void *src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
}
inline void copy_d2d(Ctx &c, void *dst, const void *src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}

// read `count` values of T from device memory (synchronous)
template<typename T>
inline void readback(Ctx &c, T *host, const T *dev, size_t count) {
    size_t bytes = sizeof(T) * count;
    if (bytes <= c.pinned_bytes) {
        copy_d2h(c, c.pinned, dev, bytes);
        c.sync();
        std::memcpy(host, c.pinned, bytes);
    } else {
        copy_d2h(c, host, dev, bytes);
        c.sync();
    }
}

} // namespace heipa
