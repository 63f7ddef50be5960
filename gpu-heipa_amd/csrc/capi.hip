// extern "C" entry points (include/heipa.h).  Every call runs on the context's
// stream, is synchronous at return, and converts internal errors to status
// codes (the reference exits or aborts instead).
#include "device/eval.hpp"
#include "host/host_graph.hpp"
#include "host/lend.h"

#include <cstdio>
#include <cstring>
#include <new>

using namespace heipa;

struct heipa_ctx {
    Ctx c;
};

struct heipa_host_graph {
    HostGraph g;
};

namespace {

template<typename F>
int guarded(heipa_ctx *ctx, F &&f) {
    if (!ctx) return HEIPA_ERR_INVALID;
    try {
        HIP_CHECK(hipSetDevice(ctx->c.device));
        const size_t fd = ctx->c.arena.front_depth(), bd = ctx->c.arena.back_depth();
        try {
            f(ctx->c);
        } catch (...) {
            // leave the arena as it was on entry
            ctx->c.arena.release_back(bd);
            ctx->c.arena.release_front(fd);
            throw;
        }
        ctx->c.last_error.clear();
        return HEIPA_OK;
    } catch (const Error &e) {
        ctx->c.last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->c.last_error = "host allocation failed";
        return HEIPA_ERR_OOM;
    } catch (const std::exception &e) {
        ctx->c.last_error = e.what();
        return HEIPA_ERR_INVALID;
    }
}

template<typename F>
int guarded_host(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return HEIPA_ERR_OOM;
    } catch (...) {
        return HEIPA_ERR_INVALID;
    }
}

void check_graph(const heipa_graph *g) {
    require(g != nullptr, "null graph");
    require(g->n == 0 || (g->xadj && (g->m == 0 || g->adjncy)), "graph arrays missing");
    require((g->flags & HEIPA_UNIFORM_EW) || g->m == 0 || g->adjwgt, "edge weights missing");
    require((g->flags & HEIPA_UNIFORM_VW) || g->n == 0 || g->vwgt, "vertex weights missing");
}

} // namespace

extern "C" {

int heipa_ctx_create(int device, uint64_t arena_bytes, heipa_ctx **out) {
    if (!out) return HEIPA_ERR_INVALID;
    *out = nullptr;
    heipa_ctx *ctx = new (std::nothrow) heipa_ctx();
    if (!ctx) return HEIPA_ERR_OOM;
    ctx->c.handle = ctx;
    try {
        ctx_open(ctx->c, device, arena_bytes);
    } catch (const Error &e) {
        std::fprintf(stderr, "heipa_ctx_create: %s\n", e.what());
        delete ctx;
        return e.code == HEIPA_ERR_DEVICE ? HEIPA_ERR_DEVICE : e.code;
    }
    *out = ctx;
    return HEIPA_OK;
}

int heipa_ctx_destroy(heipa_ctx *ctx) {
    if (!ctx) return HEIPA_ERR_INVALID;
    ctx_close(ctx->c);
    delete ctx;
    return HEIPA_OK;
}

const char *heipa_ctx_last_error(const heipa_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

void *heipa_ctx_stream(heipa_ctx *ctx) { return ctx ? (void *) ctx->c.stream : nullptr; }

uint64_t heipa_arena_mark(const heipa_ctx *ctx) { return ctx ? ctx->c.arena.front_depth() : 0; }

int heipa_arena_release(heipa_ctx *ctx, uint64_t mark) {
    return guarded(ctx, [&](Ctx &c) {
        require(mark <= c.arena.front_depth(), "arena mark is above the current depth");
        c.sync();
        c.arena.release_front((size_t) mark);
    });
}

uint64_t heipa_arena_peak(const heipa_ctx *ctx) { return ctx ? ctx->c.arena.peak() : 0; }

int heipa_arena_reset_peak(heipa_ctx *ctx) {
    if (!ctx) return HEIPA_ERR_INVALID;
    ctx->c.arena.reset_peak();
    return HEIPA_OK;
}

int heipa_download(heipa_ctx *ctx, void *host_dst, const void *dev_src, uint64_t bytes) {
    return guarded(ctx, [&](Ctx &c) {
        require(bytes == 0 || (host_dst && dev_src), "null argument");
        copy_d2h(c, host_dst, dev_src, bytes);
        c.sync();
    });
}

int heipa_upload(heipa_ctx *ctx, void *dev_dst, const void *host_src, uint64_t bytes) {
    return guarded(ctx, [&](Ctx &c) {
        require(bytes == 0 || (dev_dst && host_src), "null argument");
        copy_h2d(c, dev_dst, host_src, bytes);
        c.sync();
    });
}

int heipa_edge_cut(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, int64_t *cut) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(cut && (part || g->n == 0), "null argument");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        BackScope bs(c.arena);
        GraphMeta meta = make_meta(c, d, true);
        *cut = edge_cut(c, d, meta, part);
    });
}

int heipa_comm_cost(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, const int32_t *D, uint32_t k,
                    int64_t *J) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(J && D && k > 0 && (part || g->n == 0), "null argument");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        validate_part(c, part, d.n, k);
        BackScope bs(c.arena);
        GraphMeta meta = make_meta(c, d, true);
        *J = comm_cost(c, d, meta, part, D, k);
    });
}

int heipa_distance_matrix(const uint32_t *hier, const int32_t *dist, uint32_t l, int32_t *D_host) {
    if (!hier || !dist || !D_host || l == 0) return HEIPA_ERR_INVALID;
    u64 k = 1;
    for (u32 i = 0; i < l; ++i) {
        if (hier[i] == 0) return HEIPA_ERR_INVALID;
        k *= hier[i];
    }
    // distance_oracle_matrix.h:64-82
    for (u64 i = 0; i < k; ++i) {
        for (u64 j = 0; j < k; ++j) {
            if (i == j) {
                D_host[i * k + j] = 0;
                continue;
            }
            u32 level = 0;
            u64 group = 1;
            for (; level < l; ++level) {
                group *= hier[level];
                if (i / group == j / group) break;
            }
            D_host[i * k + j] = dist[level < l ? level : l - 1];
        }
    }
    return HEIPA_OK;
}

int heipa_block_weights(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *part, uint32_t k, int64_t *bw_host,
                        int64_t *maxw) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(k > 0 && bw_host, "bad argument");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        validate_part(c, part, d.n, k);
        BackScope bs(c.arena);
        s64 *bw = bs.get<s64>(k);
        block_weights_dev(c, d, part, k, bw);
        readback(c, (s64 *) bw_host, bw, k);
        if (maxw) {
            s64 m = 0;
            for (u32 i = 0; i < k; ++i) m = std::max<s64>(m, bw_host[i]);
            *maxw = m;
        }
    });
}

int heipa_match(heipa_ctx *ctx, const heipa_graph *g, uint32_t seed, uint32_t *mapping, uint32_t *nc,
                heipa_match_stats *stats) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(nc && (mapping || g->n == 0), "null argument");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        BackScope bs(c.arena);
        GraphMeta meta = make_meta(c, d, true);
        *nc = match(c, d, meta, seed, mapping, stats);
    });
}

int heipa_contract(heipa_ctx *ctx, const heipa_graph *fine, const uint32_t *mapping, uint32_t nc,
                   heipa_graph *coarse) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(fine);
        require(coarse && (mapping || fine->n == 0), "null argument");
        require(nc <= fine->n, "nc exceeds n");
        DGraph d = to_dgraph(fine);
        validate_graph(c, d);
        validate_part(c, mapping, d.n, nc);
        GraphMeta meta;  // unused by contraction
        CoarseGraph cg = contract(c, d, meta, mapping, nc);
        c.sync();
        coarse->n = cg.g.n;
        coarse->m = cg.g.m;
        coarse->g_weight = cg.g.gw;
        coarse->flags = 0;
        coarse->reserved = 0;
        coarse->xadj = cg.g.xadj;
        coarse->adjncy = cg.g.adj;
        coarse->adjwgt = cg.g.ew;
        coarse->vwgt = cg.g.vw;
    });
}

int heipa_project(heipa_ctx *ctx, const uint32_t *mapping, uint32_t fine_n, const uint32_t *coarse_part,
                  uint32_t *fine_part) {
    return guarded(ctx, [&](Ctx &c) {
        require(fine_n == 0 || (mapping && coarse_part && fine_part), "null argument");
        project(c, mapping, fine_n, coarse_part, fine_part);
        c.sync();
    });
}

static int jet_common(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, uint32_t k,
                      int64_t lmax, int ultra, uint32_t level, int64_t *cut, int64_t *maxw, heipa_jet_trace *trace,
                      uint32_t trace_cap, uint32_t *trace_len, const int32_t *D = nullptr) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(cut && maxw && k > 0 && (part || g->n == 0), "null argument");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        validate_part(c, part, d.n, k);
        BackScope bs(c.arena);
        GraphMeta meta = make_meta(c, d, true);
        std::vector<heipa_jet_trace> tr;
        JetResult r =
            jet_refine(c, d, meta, part, k, lmax, ultra != 0, level, *cut, *maxw, trace ? &tr : nullptr, D);
        // the reference's copy_into carries the best partition's block weights along with its map
        // (partition.h:192-200), so partition.bweights is current on return
        if (bweights) block_weights_dev(c, d, part, k, bweights);
        c.sync();
        *cut = r.cut;
        *maxw = r.maxw;
        if (trace) {
            const u32 cnt = (u32) std::min<size_t>(tr.size(), trace_cap);
            for (u32 i = 0; i < cnt; ++i) trace[i] = tr[i];
            if (trace_len) *trace_len = (u32) tr.size();
        }
    });
}

int heipa_jet_refine(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, uint32_t k,
                     int64_t lmax, int ultra, uint32_t level, int64_t *cut, int64_t *maxw) {
    return jet_common(ctx, g, part, bweights, k, lmax, ultra, level, cut, maxw, nullptr, 0, nullptr);
}

int heipa_jet_refine_traced(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, uint32_t k,
                            int64_t lmax, int ultra, uint32_t level, int64_t *cut, int64_t *maxw,
                            heipa_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len) {
    return jet_common(ctx, g, part, bweights, k, lmax, ultra, level, cut, maxw, trace, trace_cap, trace_len);
}

int heipa_map_refine(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights, const int32_t *D,
                     uint32_t k, int64_t lmax, uint32_t level, int64_t *J, int64_t *maxw) {
    if (!D) return guarded(ctx, [](Ctx &) { require(false, "null distance matrix"); });
    return jet_common(ctx, g, part, bweights, k, lmax, 0, level, J, maxw, nullptr, 0, nullptr, D);
}

int heipa_map_refine_traced(heipa_ctx *ctx, const heipa_graph *g, uint32_t *part, int64_t *bweights,
                            const int32_t *D, uint32_t k, int64_t lmax, uint32_t level, int64_t *J, int64_t *maxw,
                            heipa_jet_trace *trace, uint32_t trace_cap, uint32_t *trace_len) {
    if (!D) return guarded(ctx, [](Ctx &) { require(false, "null distance matrix"); });
    return jet_common(ctx, g, part, bweights, k, lmax, 0, level, J, maxw, trace, trace_cap, trace_len, D);
}

static void fill_stats(heipa_solve_stats *stats, const SolveResult &r) {
    if (!stats) return;
    stats->cut = r.cut;
    stats->maxw = r.maxw;
    stats->lmax = r.lmax;
    stats->initial_cut = r.initial_cut;
    stats->levels = r.levels;
    stats->jet_iterations = r.jet_iterations;
    stats->ms_coarsening = r.ms[0];
    stats->ms_contraction = r.ms[1];
    stats->ms_initial = r.ms[2];
    stats->ms_uncontraction = r.ms[3];
    stats->ms_refinement = r.ms[4];
}

int heipa_im_partition(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *hier, const int32_t *dist, uint32_t l,
                       double imbalance, uint64_t seed, uint32_t *part, heipa_solve_stats *stats,
                       heipa_level_trace *levels, uint32_t level_cap) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(hier && dist && l > 0 && (part || g->n == 0), "bad argument");
        require(imbalance >= 0.0, "imbalance must be non-negative");
        MapSpec ms;
        ms.hier.assign(hier, hier + l);
        u64 k = 1;
        for (u32 x: ms.hier) {
            require(x > 0, "hierarchy entries must be positive");
            k *= x;
            require(k <= (1u << 16), "hierarchy product too large");
        }
        // the mapping Jet runs on every level above the coarsest, which exists once n > 8k
        require(g->n <= 8 * k || k <= JET_MAX_K, "hierarchy product too large for the mapping refinement");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        std::vector<s32> Dh(k * k);
        require(heipa_distance_matrix(hier, dist, l, Dh.data()) == HEIPA_OK, "bad hierarchy");
        const size_t depth0 = c.arena.front_depth();
        s32 *Dd = (s32 *) c.arena.push_front(4ull * k * k);
        copy_h2d(c, Dd, Dh.data(), 4ull * k * k);
        ms.D = Dd;
        std::vector<heipa_level_trace> lv;
        SolveResult r = solve(c, d, (u32) k, imbalance, seed, false, part, levels ? &lv : nullptr, &ms);
        c.sync();
        c.arena.release_front(depth0);
        fill_stats(stats, r);
        if (levels) {
            for (u32 i = 0; i < lv.size() && i < level_cap; ++i) levels[i] = lv[i];
        }
    });
}

int heipa_partition(heipa_ctx *ctx, const heipa_graph *g, uint32_t k, double imbalance, uint64_t seed, int ultra,
                    uint32_t *part, heipa_solve_stats *stats, heipa_level_trace *levels, uint32_t level_cap) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(k > 0 && (part || g->n == 0), "bad argument");
        require(imbalance >= 0.0, "imbalance must be non-negative");
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        std::vector<heipa_level_trace> lv;
        SolveResult r = solve(c, d, k, imbalance, seed, ultra != 0, part, levels ? &lv : nullptr);
        if (stats) {
            stats->cut = r.cut;
            stats->maxw = r.maxw;
            stats->lmax = r.lmax;
            stats->initial_cut = r.initial_cut;
            stats->levels = r.levels;
            stats->jet_iterations = r.jet_iterations;
            stats->ms_coarsening = r.ms[0];
            stats->ms_contraction = r.ms[1];
            stats->ms_initial = r.ms[2];
            stats->ms_uncontraction = r.ms[3];
            stats->ms_refinement = r.ms[4];
        }
        if (levels) {
            for (u32 i = 0; i < lv.size() && i < level_cap; ++i) levels[i] = lv[i];
        }
    });
}

int heipa_hm_partition(heipa_ctx *ctx, const heipa_graph *g, const uint32_t *hier, uint32_t l, double imbalance,
                       uint64_t seed, int ultra, uint32_t *part) {
    return guarded(ctx, [&](Ctx &c) {
        check_graph(g);
        require(hier && l > 0 && (part || g->n == 0), "bad argument");
        require(imbalance >= 0.0, "imbalance must be non-negative");
        std::vector<u32> h(hier, hier + l);
        u64 prod = 1;
        for (u32 x: h) {
            require(x > 0, "hierarchy entries must be positive");
            prod *= x;
            require(prod <= (1u << 16), "hierarchy product too large");
        }
        DGraph d = to_dgraph(g);
        validate_graph(c, d);
        hm_partition(c, d, h, imbalance, seed, ultra != 0, part);
    });
}

double heipa_adaptive_imbalance(double global_imbalance, int64_t global_g_weight, uint32_t global_k,
                                int64_t local_g_weight, uint32_t local_k_rem, uint64_t depth) {
    return adaptive_imbalance(global_imbalance, global_g_weight, global_k, local_g_weight, local_k_rem, depth);
}

int heipa_prof_enable(heipa_ctx *ctx, int on) {
    return guarded(ctx, [&](Ctx &c) {
        c.sync();
        c.prof_on = on != 0;
    });
}

int heipa_ctx_set_option(heipa_ctx *ctx, int option, uint64_t value) {
    return guarded(ctx, [&](Ctx &c) {
        c.sync();
        if (option == HEIPA_OPT_RESOLVE_CAP) c.resolve_cap = value > 0xFFFFFFFFull ? ~0u : (u32) value;
        else if (option == HEIPA_OPT_BLOCK_JET) c.block_jet_max_n = value > 0xFFFFFFFFull ? ~0u : (u32) value;
        else if (option == HEIPA_OPT_GRAPHS) c.use_graphs = value != 0;
        else if (option == HEIPA_OPT_PROF_SAMPLE) c.prof_sample = value < 1 ? 1u : (u32) std::min<uint64_t>(value, 1u << 30);
        else if (option == HEIPA_OPT_BUCKET_COPIES) c.bucket_copies = (u32) std::min<uint64_t>(value, 32);
        else if (option == HEIPA_OPT_BLOCK_LDS) c.block_lds = value != 0;
        else if (option == HEIPA_OPT_GROUP_CAP) c.group_cap = (u32) std::min<uint64_t>(value, 64);
        else if (option == HEIPA_OPT_BLOCK_GROUP_CAP) c.block_group_cap = (u32) std::min<uint64_t>(value, 64);
        else if (option == HEIPA_OPT_REGION_EDGES) c.region_edges = (u32) std::min<uint64_t>(value, 0xFFFFFFFFull);
        else if (option == HEIPA_OPT_HM_WORKERS) {
            c.hm_workers = (u32) std::min<uint64_t>(value, 64);
            ctx_trim_workers(c, c.hm_workers > 1 ? c.hm_workers : 0);  // frees the arenas of the ones dropped
        } else if (option == HEIPA_OPT_RACE) {
            c.race = value != 0;
        } else if (option == HEIPA_OPT_HIER_CACHE) {
            if (value) {
                c.hcache = std::make_shared<HierCache>();
                c.hcache->budget = (size_t) value;
            } else {
                c.hcache.reset();
            }
        } else if (option == HEIPA_OPT_GRAPH_KEY) {
            c.hkey = value;
        } else if (option == HEIPA_OPT_HM_WORKER_BYTES) {
            c.hm_worker_bytes = value;
            ctx_trim_workers(c, 0);  // reopened at the next split with the new size
        }
        else require(false, "unknown option");
    });
}

int heipa_ctx_set_lender_internal(heipa_ctx *ctx, heipa_borrow_fn borrow, heipa_give_back_fn give_back, void *user) {
    if (!ctx) return HEIPA_ERR_INVALID;
    if (!borrow || !give_back) {
        ctx->c.borrow = nullptr;
        ctx->c.give_back = nullptr;
        return HEIPA_OK;
    }
    ctx->c.borrow = [borrow, user](size_t max) {
        std::vector<heipa_ctx *> got(max);
        got.resize(max ? borrow(user, got.data(), max) : 0);
        std::vector<Ctx *> out;
        for (heipa_ctx *h: got) out.push_back(&h->c);
        return out;
    };
    ctx->c.give_back = [give_back, user](const std::vector<Ctx *> &lent) {
        std::vector<heipa_ctx *> hs;
        for (Ctx *x: lent) hs.push_back(static_cast<heipa_ctx *>(x->handle));
        give_back(user, hs.data(), hs.size());
    };
    return HEIPA_OK;
}

int heipa_ctx_share_hier_cache(heipa_ctx *dst, const heipa_ctx *src) {
    if (!src) return HEIPA_ERR_INVALID;
    return guarded(dst, [&](Ctx &c) {
        c.sync();
        c.hcache = src->c.hcache;
    });
}

int heipa_hier_cache_stats(const heipa_ctx *ctx, uint64_t *entries, uint64_t *bytes, uint64_t *levels_reused,
                           uint64_t *levels_built) {
    if (!ctx) return HEIPA_ERR_INVALID;
    const std::shared_ptr<HierCache> hc = ctx->c.hcache;
    uint64_t e = 0, b = 0, r = 0, bl = 0;
    if (hc) {
        std::lock_guard<std::mutex> lk(hc->m);
        e = hc->entries.size();
        b = hc->bytes;
        r = hc->levels_reused;
        bl = hc->levels_built;
    }
    if (entries) *entries = e;
    if (bytes) *bytes = b;
    if (levels_reused) *levels_reused = r;
    if (levels_built) *levels_built = bl;
    return HEIPA_OK;
}

int heipa_prof_filter(heipa_ctx *ctx, const char *prefix) {
    return guarded(ctx, [&](Ctx &c) {
        c.sync();
        c.prof_filter = prefix ? prefix : "";
    });
}

int heipa_prof_reset(heipa_ctx *ctx) {
    return guarded(ctx, [&](Ctx &c) {
        c.sync();
        for (auto &s: c.prof_stats) s = KernelStat{};
    });
}

uint32_t heipa_prof_count(heipa_ctx *ctx) {
    if (!ctx) return 0;
    try {
        ctx->c.sync();
    } catch (...) {
    }
    return (uint32_t) ctx->c.prof_names.size();
}

int heipa_prof_get(heipa_ctx *ctx, uint32_t idx, char *name, uint32_t name_cap, uint64_t *launches, double *total_ms,
                   double *alg_bytes) {
    return guarded(ctx, [&](Ctx &c) {
        c.sync();
        require(idx < c.prof_names.size(), "profile index out of range");
        if (name && name_cap) {
            std::strncpy(name, c.prof_names[idx].c_str(), name_cap - 1);
            name[name_cap - 1] = 0;
        }
        const KernelStat &s = c.prof_stats[idx];
        if (launches) *launches = s.launches;
        // the timed launches' mean, extrapolated to every launch
        if (total_ms) *total_ms = s.timed ? s.ms * (double) s.launches / (double) s.timed : 0.0;
        if (alg_bytes) *alg_bytes = s.bytes;
    });
}

// ---- host graph tools
int heipa_read_metis(const char *path, heipa_host_graph **out) {
    if (!path || !out) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        auto *h = new heipa_host_graph();
        std::string err = read_metis(path, h->g);
        if (!err.empty()) {
            std::fprintf(stderr, "%s\n", err.c_str());
            delete h;
            return err.find("does not exist") != std::string::npos ? HEIPA_ERR_IO : HEIPA_ERR_INVALID;
        }
        *out = h;
        return HEIPA_OK;
    });
}

int heipa_write_metis(const char *path, const heipa_host_graph *g) {
    if (!path || !g) return HEIPA_ERR_INVALID;
    return write_metis(path, g->g).empty() ? HEIPA_OK : HEIPA_ERR_IO;
}

int heipa_write_partition(const char *path, const uint32_t *part_host, uint32_t n) {
    if (!path || (!part_host && n)) return HEIPA_ERR_INVALID;
    return write_partition(path, part_host, n).empty() ? HEIPA_OK : HEIPA_ERR_IO;
}

int heipa_gen_grid(uint32_t rows, uint32_t cols, heipa_host_graph **out) {
    if (!out) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        auto *h = new heipa_host_graph();
        h->g = gen_grid(rows, cols);
        *out = h;
        return HEIPA_OK;
    });
}

int heipa_gen_rgg(uint32_t n, double avg_deg, uint64_t seed, uint32_t wmod, heipa_host_graph **out) {
    if (!out || n == 0) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        auto *h = new heipa_host_graph();
        h->g = gen_rgg(n, avg_deg, seed, wmod);
        *out = h;
        return HEIPA_OK;
    });
}

int heipa_gen_rmat(uint32_t scale, uint32_t n, double edge_factor, double a, double b, double c, uint64_t seed,
                   uint32_t wmod, int explicit_vw, heipa_host_graph **out) {
    if (!out || scale == 0 || scale > 31) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        auto *h = new heipa_host_graph();
        h->g = gen_rmat(scale, n, edge_factor, a, b, c, seed, wmod, explicit_vw != 0);
        *out = h;
        return HEIPA_OK;
    });
}

int heipa_host_graph_from_arrays(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                                 const int32_t *vwgt, heipa_host_graph **out) {
    if (!out || !xadj || (!adjncy && xadj[n] > 0)) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        auto *h = new heipa_host_graph();
        HostGraph &g = h->g;
        g.n = n;
        g.m = xadj[n];
        g.xadj.assign(xadj, xadj + n + 1);
        g.adjncy.assign(adjncy, adjncy + g.m);
        g.uniform_edge_weights = adjwgt == nullptr;
        g.uniform_vertex_weights = vwgt == nullptr;
        if (adjwgt) g.adjwgt.assign(adjwgt, adjwgt + g.m);
        if (vwgt) g.vwgt.assign(vwgt, vwgt + n);
        s64 w = 0;
        for (u32 u = 0; u < n; ++u) w += vwgt ? vwgt[u] : 1;
        g.g_weight = w;
        *out = h;
        return HEIPA_OK;
    });
}

int heipa_host_graph_view(const heipa_host_graph *h, uint32_t *n, uint32_t *m, int64_t *g_weight, uint32_t *flags,
                          const uint32_t **xadj, const uint32_t **adjncy, const int32_t **adjwgt,
                          const int32_t **vwgt) {
    if (!h) return HEIPA_ERR_INVALID;
    const HostGraph &g = h->g;
    if (n) *n = g.n;
    if (m) *m = g.m;
    if (g_weight) *g_weight = g.g_weight;
    if (flags) *flags = (g.uniform_vertex_weights ? HEIPA_UNIFORM_VW : 0u) | (g.uniform_edge_weights ? HEIPA_UNIFORM_EW : 0u);
    if (xadj) *xadj = g.xadj.data();
    if (adjncy) *adjncy = g.adjncy.data();
    if (adjwgt) *adjwgt = g.uniform_edge_weights ? nullptr : g.adjwgt.data();
    if (vwgt) *vwgt = g.uniform_vertex_weights ? nullptr : g.vwgt.data();
    return HEIPA_OK;
}

void heipa_host_graph_free(heipa_host_graph *g) { delete g; }

int heipa_initial_partition(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                            const int32_t *vwgt, uint32_t k, double imbalance, uint64_t seed, uint32_t *part) {
    if (!xadj || !part || k == 0) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        initial_partition(n, xadj, adjncy, adjwgt, vwgt, k, imbalance, seed, part);
        return HEIPA_OK;
    });
}

int heipa_global_multisection(uint32_t n, const uint32_t *xadj, const uint32_t *adjncy, const int32_t *adjwgt,
                              const int32_t *vwgt, const uint32_t *hier, uint32_t l, double imbalance, uint64_t seed,
                              uint32_t *part) {
    if (!xadj || (!part && n) || !hier || l == 0) return HEIPA_ERR_INVALID;
    for (u32 i = 0; i < l; ++i)
        if (hier[i] == 0) return HEIPA_ERR_INVALID;
    return guarded_host([&]() {
        global_multisection(n, xadj, adjncy, adjwgt, vwgt, hier, l, imbalance, seed, part);
        return HEIPA_OK;
    });
}

} // extern "C"
