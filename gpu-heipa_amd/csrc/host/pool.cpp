// Batched and served forms: a pool of workers on one device.
//
// Each worker owns a heipa_ctx (its own HIP stream, arena and pinned staging)
// and a host thread.  Jobs wait in one queue in arrival order; a free worker
// takes the first queued job whose device-memory estimate fits beside the
// running jobs' within the pool's budget (admission by estimate, so requests
// are admitted as workers and memory free up).  A job that does not fit yet
// may be passed by later, smaller ones, at most MAX_PASSED times: after that
// nothing behind it starts before it, so it cannot starve.  A job runs exactly the
// single-graph entry point (heipa_partition / heipa_hm_partition) on its
// worker's context: its result is the same alone and in the mix, and the
// kernels of jobs on different workers overlap on the device.
#include "../../../include/heipa.h"
#include "lend.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

struct heipa_pool {
    int device = 0;
    uint64_t arena_bytes = 0, budget = 0;
    std::vector<heipa_ctx *> ctxs;
    std::vector<std::thread> threads;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    struct Queued {
        heipa_job *job;
        uint32_t passed;  // later jobs admitted before it
    };
    std::deque<Queued> queue;
    std::vector<uint8_t> idle, lent;  // per worker: waiting for a job / lent to another job's HM subtrees
    uint64_t inflight = 0;
    bool stop = false;
    // the launch-shape options heipa_pool_autotune searches, as last set
    uint32_t group_cap = 64, block_group_cap = 8, block_jet = 2048;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// An envelope of the arena peaks measured on MI355X over 1144 jobs of the
// batched and served mixes (RGG, grid and R-MAT graphs of 10^3 to 10^6
// vertices, k from 2 to 256, HM 4:8:6 and 2:4:8): about 72 bytes per
// directed edge (R-MAT levels sort their coarse edges through device-wide
// radix passes; RGG and grids stay near 25), 80 per vertex, the minibucket
// arrays of 50 k sections entries, and an HM job's children.  Every measured
// peak is below it (the smallest margin 1.17x, the median 4x: a small RGG
// needs a third of what an R-MAT graph of its size does).
uint64_t estimate(const heipa_job &j) {
    const uint64_t n = j.graph.n, m = j.graph.m;
    uint64_t k = j.k;
    if (j.kind == HEIPA_JOB_HM) {
        k = 1;
        for (uint32_t i = 0; i < j.l && i < 8; ++i) k = std::max<uint64_t>(k, j.hier[i]);
    }
    const uint64_t sections = std::max<uint64_t>(1, std::min<uint64_t>(128, n / (4096 * std::max<uint64_t>(k, 1)) + 1));
    const uint64_t bucket_bytes = 50ull * k * sections * 32;
    uint64_t bytes = 80 * n + 72 * m + bucket_bytes + (2ull << 20);
    if (j.kind == HEIPA_JOB_HM) bytes += 16 * m;  // child graphs of the recursion
    return bytes;
}

void run_job(heipa_ctx *ctx, heipa_job &j) {
    if (j.kind == HEIPA_JOB_PARTITION) {
        heipa_solve_stats st{};
        j.status = heipa_partition(ctx, &j.graph, j.k, j.imbalance, j.seed, j.ultra, j.part, &st, nullptr, 0);
        if (j.status == HEIPA_OK) {
            j.cut = st.cut;
            j.maxw = st.maxw;
        }
        return;
    }
    if (j.kind != HEIPA_JOB_HM || j.l == 0 || j.l > 8) {
        j.status = HEIPA_ERR_INVALID;
        return;
    }
    j.status = heipa_hm_partition(ctx, &j.graph, j.hier, j.l, j.imbalance, j.seed, j.ultra, j.part);
    if (j.status != HEIPA_OK) return;
    uint32_t k = 1;
    for (uint32_t i = 0; i < j.l; ++i) k *= j.hier[i];
    j.status = heipa_edge_cut(ctx, &j.graph, j.part, &j.cut);
    if (j.status != HEIPA_OK) return;
    std::vector<int64_t> bw(k);
    j.status = heipa_block_weights(ctx, &j.graph, j.part, k, bw.data(), &j.maxw);
}

constexpr uint32_t MAX_PASSED = 16;

// An HM job's subtrees below its first split may run on workers that are
// idle while no queued job can start: they are lent to the job (no queued
// job starts on them) and come back when its subtrees are done.
size_t borrow_cb(void *user, heipa_ctx **out, size_t max);
void give_back_cb(void *user, heipa_ctx *const *ctxs, size_t n);

// index of the job a free worker takes now, or -1: the first queued job that
// fits the budget, unless an earlier one has been passed MAX_PASSED times
long admissible(const heipa_pool *p) {
    for (size_t i = 0; i < p->queue.size(); ++i) {
        const heipa_pool::Queued &q = p->queue[i];
        if (p->inflight + q.job->estimate <= p->budget) return (long) i;
        if (q.passed >= MAX_PASSED) return -1;
    }
    return -1;
}

void worker(heipa_pool *p, uint32_t slot) {
    for (;;) {
        heipa_job *job = nullptr;
        {
            std::unique_lock<std::mutex> lk(p->m);
            long pick = -1;
            p->idle[slot] = 1;
            p->cv_work.wait(lk, [&] {
                if (p->lent[slot]) return false;
                if (p->stop && p->queue.empty()) return true;
                pick = admissible(p);
                return pick >= 0;
            });
            p->idle[slot] = 0;
            if (pick < 0) return;  // stopping
            job = p->queue[(size_t) pick].job;
            for (long i = 0; i < pick; ++i) p->queue[(size_t) i].passed++;
            p->queue.erase(p->queue.begin() + pick);
            p->inflight += job->estimate;
            job->state = 2;
            job->slot = (int32_t) slot;
            job->t_start = p->now();
        }
        heipa_arena_reset_peak(p->ctxs[slot]);
        if (job->graph_key) heipa_ctx_set_option(p->ctxs[slot], HEIPA_OPT_GRAPH_KEY, job->graph_key);
        run_job(p->ctxs[slot], *job);
        if (job->graph_key) heipa_ctx_set_option(p->ctxs[slot], HEIPA_OPT_GRAPH_KEY, 0);
        job->peak = heipa_arena_peak(p->ctxs[slot]);
        {
            std::lock_guard<std::mutex> lk(p->m);
            p->inflight -= job->estimate;
            job->t_done = p->now();
            job->state = 3;
        }
        p->cv_work.notify_all();
        p->cv_done.notify_all();
    }
}

size_t borrow_cb(void *user, heipa_ctx **out, size_t max) {
    heipa_pool *p = static_cast<heipa_pool *>(user);
    std::lock_guard<std::mutex> lk(p->m);
    if (admissible(p) >= 0) return 0;  // a queued job can start: the idle workers are its
    size_t n = 0;
    for (size_t s = 0; s < p->ctxs.size() && n < max; ++s)
        if (p->idle[s] && !p->lent[s]) {
            p->lent[s] = 1;
            out[n++] = p->ctxs[s];
        }
    return n;
}

void give_back_cb(void *user, heipa_ctx *const *ctxs, size_t n) {
    heipa_pool *p = static_cast<heipa_pool *>(user);
    {
        std::lock_guard<std::mutex> lk(p->m);
        for (size_t i = 0; i < n; ++i)
            for (size_t s = 0; s < p->ctxs.size(); ++s)
                if (p->ctxs[s] == ctxs[i]) p->lent[s] = 0;
    }
    p->cv_work.notify_all();
}

} // namespace

extern "C" {

uint64_t heipa_job_estimate(const heipa_job *job) { return job ? estimate(*job) : 0; }

int heipa_pool_create(int device, uint32_t slots, uint64_t arena_bytes, uint64_t memory_budget, heipa_pool **out) {
    if (!out || slots == 0 || arena_bytes == 0) return HEIPA_ERR_INVALID;
    *out = nullptr;
    heipa_pool *p = new (std::nothrow) heipa_pool();
    if (!p) return HEIPA_ERR_OOM;
    p->device = device;
    p->arena_bytes = arena_bytes;
    p->budget = memory_budget ? memory_budget : arena_bytes * slots;
    for (uint32_t s = 0; s < slots; ++s) {
        heipa_ctx *c = nullptr;
        const int rc = heipa_ctx_create(device, arena_bytes, &c);
        if (rc != HEIPA_OK) {
            for (heipa_ctx *x: p->ctxs) heipa_ctx_destroy(x);
            delete p;
            return rc;
        }
        // the pool's slots already overlap whole jobs: an HM job solves its
        // subtrees on its own slot and on the slots idle at that moment
        heipa_ctx_set_option(c, HEIPA_OPT_HM_WORKERS, 1);
        heipa_ctx_set_lender_internal(c, borrow_cb, give_back_cb, p);
        p->ctxs.push_back(c);
    }
    p->idle.assign(slots, 0);
    p->lent.assign(slots, 0);
    try {
        for (uint32_t s = 0; s < slots; ++s) p->threads.emplace_back(worker, p, s);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(p->m);
            p->stop = true;
        }
        p->cv_work.notify_all();
        for (auto &t: p->threads) t.join();
        for (heipa_ctx *x: p->ctxs) heipa_ctx_destroy(x);
        delete p;
        return HEIPA_ERR_OOM;
    }
    *out = p;
    return HEIPA_OK;
}

int heipa_pool_destroy(heipa_pool *p) {
    if (!p) return HEIPA_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(p->m);
        p->stop = true;
    }
    p->cv_work.notify_all();
    for (auto &t: p->threads) t.join();
    for (heipa_ctx *x: p->ctxs) heipa_ctx_destroy(x);
    delete p;
    return HEIPA_OK;
}

int heipa_pool_submit(heipa_pool *p, heipa_job *job) {
    if (!p || !job) return HEIPA_ERR_INVALID;
    std::lock_guard<std::mutex> lk(p->m);
    if (p->stop) return HEIPA_ERR_INVALID;
    job->slot = -1;
    job->cut = job->maxw = 0;
    job->peak = 0;
    job->t_submit = p->now();
    job->t_start = job->t_done = 0.0;
    job->estimate = estimate(*job);
    if (job->estimate > p->arena_bytes || job->estimate > p->budget) {
        job->status = HEIPA_ERR_OOM;
        job->state = 3;
        job->t_done = job->t_submit;
        p->cv_done.notify_all();
        return HEIPA_OK;
    }
    job->status = HEIPA_OK;
    job->state = 1;
    p->queue.push_back({job, 0});
    p->cv_work.notify_all();
    return HEIPA_OK;
}

int heipa_pool_wait(heipa_pool *p, heipa_job *job, double timeout_ms) {
    if (!p || !job) return 0;
    std::unique_lock<std::mutex> lk(p->m);
    auto done = [&] { return job->state == 3; };
    if (timeout_ms < 0) {
        p->cv_done.wait(lk, done);
        return 1;
    }
    return p->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done) ? 1 : 0;
}

int heipa_pool_run(heipa_pool *p, heipa_job *jobs, uint32_t count) {
    if (!p || (count && !jobs)) return HEIPA_ERR_INVALID;
    for (uint32_t i = 0; i < count; ++i) {
        const int rc = heipa_pool_submit(p, &jobs[i]);
        if (rc != HEIPA_OK) return rc;
    }
    for (uint32_t i = 0; i < count; ++i) heipa_pool_wait(p, &jobs[i], -1.0);
    return HEIPA_OK;
}

double heipa_pool_clock(const heipa_pool *p) { return p ? p->now() : 0.0; }

int heipa_pool_set_option(heipa_pool *p, int option, uint64_t value) {
    if (!p) return HEIPA_ERR_INVALID;
    std::lock_guard<std::mutex> lk(p->m);
    if (option == HEIPA_OPT_HIER_CACHE) {  // one cache, shared by every worker
        int rc = heipa_ctx_set_option(p->ctxs[0], option, value);
        for (size_t i = 1; rc == HEIPA_OK && i < p->ctxs.size(); ++i)
            rc = heipa_ctx_share_hier_cache(p->ctxs[i], p->ctxs[0]);
        return rc;
    }
    for (heipa_ctx *c: p->ctxs) {
        const int rc = heipa_ctx_set_option(c, option, value);
        if (rc != HEIPA_OK) return rc;
    }
    const uint32_t v = (uint32_t) std::min<uint64_t>(value, 0xFFFFFFFFull);
    if (option == HEIPA_OPT_GROUP_CAP) p->group_cap = v;
    else if (option == HEIPA_OPT_BLOCK_GROUP_CAP) p->block_group_cap = v;
    else if (option == HEIPA_OPT_BLOCK_JET) p->block_jet = v;
    return HEIPA_OK;
}

int heipa_pool_autotune(heipa_pool *p, heipa_job *sample, uint32_t count, heipa_tune *out) {
    if (!p || !sample || count == 0 || !out) return HEIPA_ERR_INVALID;
    *out = heipa_tune{};
    int status = HEIPA_OK;
    auto run = [&]() -> double {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = heipa_pool_run(p, sample, count);
        for (uint32_t i = 0; rc == HEIPA_OK && i < count; ++i) rc = sample[i].status;
        if (rc != HEIPA_OK && status == HEIPA_OK) status = rc;
        out->runs++;
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    struct Knob {
        int option;
        uint32_t cand[3];
        uint32_t *cur;
    };
    Knob knobs[3] = {{HEIPA_OPT_GROUP_CAP, {64, 16, 8}, &p->group_cap},
                     {HEIPA_OPT_BLOCK_GROUP_CAP, {16, 8, 4}, &p->block_group_cap},
                     {HEIPA_OPT_BLOCK_JET, {1024, 2048, 4096}, &p->block_jet}};
    run();  // warm-up: first-touch costs do not belong to any setting
    double best = run();
    out->ms_start = best;
    for (Knob &kb: knobs) {
        const uint32_t start = *kb.cur;  // timed already: the best run so far
        uint32_t chosen = start;
        for (uint32_t v: kb.cand) {
            if (status != HEIPA_OK) break;
            if (v == start) continue;
            const int rc = heipa_pool_set_option(p, kb.option, v);
            if (rc != HEIPA_OK) return rc;
            const double t = run();
            if (t < 0.98 * best) {
                best = t;
                chosen = v;
            }
        }
        const int rc = heipa_pool_set_option(p, kb.option, chosen);
        if (rc != HEIPA_OK) return rc;
        if (status != HEIPA_OK) return status;
    }
    out->group_cap = p->group_cap;
    out->block_group_cap = p->block_group_cap;
    out->block_jet = p->block_jet;
    out->ms_best = best;
    return HEIPA_OK;
}

int heipa_pool_hier_cache_stats(const heipa_pool *p, uint64_t *entries, uint64_t *bytes, uint64_t *levels_reused,
                                uint64_t *levels_built) {
    if (!p || p->ctxs.empty()) return HEIPA_ERR_INVALID;
    return heipa_hier_cache_stats(p->ctxs[0], entries, bytes, levels_reused, levels_built);
}

} // extern "C"
