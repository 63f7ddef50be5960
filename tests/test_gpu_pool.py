"""Batched and served forms (heipa_pool): every job's result equals the same
request run alone on a single context (bit-exact partition arrays), under
concurrency, arrival gaps and memory-budget admission."""
import time

import numpy as np
import pytest

import heipa
from graphs import star_plus_grid
from heipa import workloads
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu


def _requests():
    g = [heipa.gen_grid(40, 50), heipa.gen_rgg(3000, 8.0, 3, 10), heipa.gen_rmat(11, 1500, 8.0, 0.57, 0.19, 0.19, 4, 10, False),
         heipa.gen_rgg(5000, 12.0, 9, 100), star_plus_grid(), heipa.gen_grid(17, 90),
         heipa.gen_rmat(12, 0, 16.0, 0.57, 0.19, 0.19, 8, 100, True), heipa.gen_rgg(2500, 8.0, 11, 0)]
    specs = [dict(k=8), dict(k=4, ultra=True), dict(k=16), dict(hier=[2, 4, 3]), dict(k=3),
             dict(k=32), dict(hier=[4, 8, 6]), dict(k=2)]
    return [(heipa.DeviceGraph(x, "cuda:0"), s) for x, s in zip(g, specs)]


def _alone(ctx, dg, s):
    if "hier" in s:
        return heipa.to_numpy_u32(ctx.hm_partition(dg, s["hier"], 0.03, 1, s.get("ultra", False)))
    part, _, _ = ctx.partition(dg, s["k"], 0.03, 1, s.get("ultra", False))
    return heipa.to_numpy_u32(part)


def test_pool_batched_equals_alone(ctx):
    reqs = _requests()
    pool = heipa.Pool(3, 512 << 20)
    try:
        jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs]
        pool.run(jobs)
    finally:
        pool.close()
    slots = set()
    for (dg, s), j in zip(reqs, jobs):
        assert j.status == heipa.OK and j.state == 3, (s, j.status)
        assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s)), s
        assert j.cut == ctx.edge_cut(dg, j._part)
        slots.add(j.slot)
    assert len(slots) > 1  # the work was spread over workers


def test_pool_served_equals_alone(ctx):
    reqs = _requests()
    pool = heipa.Pool(4, 512 << 20)
    try:
        jobs = []
        for i, (dg, s) in enumerate(reqs):
            j = heipa.Pool.job(dg, imbalance=0.03, seed=1, **s)
            pool.submit(j)
            jobs.append(j)
            time.sleep(0.002 * (i % 3))
        for j in jobs:
            assert pool.wait(j, 60000.0)
    finally:
        pool.close()
    for (dg, s), j in zip(reqs, jobs):
        assert j.status == heipa.OK
        assert j.t_submit <= j.t_start <= j.t_done
        assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s)), s


def test_pool_admission_respects_budget(ctx):
    """With a budget of the largest estimate, jobs run together only while
    their estimates fit it (a smaller job may start before a larger one that
    waits), the largest job runs alone, and every result equals its alone run."""
    reqs = _requests()
    probe = heipa.Pool(1, 512 << 20)
    jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs]
    est = [probe.estimate(j) for j in jobs]
    budget = max(est)
    probe.close()
    pool = heipa.Pool(4, 512 << 20, budget=budget)
    try:
        pool.run(jobs)
    finally:
        pool.close()
    for i, ji in enumerate(jobs):
        running = sum(est[j] for j, jj in enumerate(jobs)
                      if j != i and jj.t_start <= ji.t_start < jj.t_done)
        assert running + est[i] <= budget, (i, running, est[i], budget)
    big = jobs[est.index(budget)]
    assert not any(j is not big and j.t_start < big.t_done and big.t_start < j.t_done for j in jobs)
    for (dg, s), j in zip(reqs, jobs):
        assert j.status == heipa.OK
        assert j.estimate >= j.peak > 0
        assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s))


def test_pool_rejects_job_larger_than_arena():
    g = heipa.DeviceGraph(heipa.gen_grid(64, 64), "cuda:0")
    pool = heipa.Pool(1, 1 << 20)
    try:
        j = heipa.Pool.job(g, k=4)
        pool.submit(j)
        assert pool.wait(j, 1000.0)
        assert j.status == heipa.ERR_OOM and j.slot == -1
    finally:
        pool.close()


def test_pool_failed_job_does_not_affect_the_others(ctx):
    """A malformed graph (a neighbour id >= n, rejected by the validation
    launches) fails alone with HEIPA_ERR_INVALID; the jobs around it, on the
    same workers, still equal their alone runs, and the pool accepts more
    work afterwards."""
    reqs = _requests()[:4]
    bad = heipa.HostGraph(np.array([0, 1, 2], np.uint32), np.array([1, 7], np.uint32))
    bad_dg = heipa.DeviceGraph(bad, "cuda:0")
    pool = heipa.Pool(2, 512 << 20)
    try:
        jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs]
        jobs.insert(2, heipa.Pool.job(bad_dg, k=2, imbalance=0.03, seed=1))
        pool.run(jobs)
        assert jobs[2].status == heipa.ERR_INVALID
        rest = [j for i, j in enumerate(jobs) if i != 2]
        for (dg, s), j in zip(reqs, rest):
            assert j.status == heipa.OK, (s, j.status)
            assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s)), s
        again = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs[:2]]
        pool.run(again)
        for (dg, s), j in zip(reqs[:2], again):
            assert j.status == heipa.OK
            assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s)), s
    finally:
        pool.close()


def test_pool_autotune_keeps_results(ctx):
    """The autotuner times a sample under each candidate launch shape and
    leaves the fastest on the workers; the sample's partitions (from its last
    run) and a later run under the chosen setting equal the alone runs."""
    reqs = workloads.batched(12)
    dgs = [_dev(g) for g, _, _ in reqs]
    want = [heipa.to_numpy_u32(ctx.partition(dg, spec["k"], 0.03, 1, False)[0]).copy()
            for dg, (_, spec, _) in zip(dgs, reqs)]
    pool = heipa.Pool(4, 256 << 20)
    try:
        sample = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
        t = pool.autotune(sample)
        assert t["group_cap"] in (64, 16, 8) and t["block_group_cap"] in (16, 8, 4)
        assert t["block_jet"] in (1024, 2048, 4096)
        assert t["runs"] == 8 and 0 < t["ms_best"] <= t["ms_start"]
        for j, w, dg in zip(sample, want, dgs):
            assert j.status == heipa.OK and np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], w)
        jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
        pool.run(jobs)
        for j, w, dg in zip(jobs, want, dgs):
            assert j.status == heipa.OK and np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], w)
    finally:
        pool.close()


def test_pool_hm_job_on_lent_workers(ctx):
    """An HM job alone in a pool solves its subtrees below the first split on
    the pool's idle workers (lent to it while no queued job can start); a
    second run beside partition jobs lends fewer.  Both equal the alone run."""
    g = heipa.gen_rmat(14, 0, 16.0, 0.57, 0.19, 0.19, 7, 100, True)
    dg = _dev(g)
    want = _alone(ctx, dg, dict(hier=[4, 8, 6]))
    small = [(heipa.DeviceGraph(heipa.gen_rgg(4000 + 500 * i, 8.0, 20 + i, 10), "cuda:0"), dict(k=8)) for i in range(5)]
    pool = heipa.Pool(6, 512 << 20)
    try:
        j = heipa.Pool.job(dg, hier=[4, 8, 6], imbalance=0.03, seed=1)
        pool.run([j])
        assert j.status == heipa.OK
        assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], want)
        jobs = [heipa.Pool.job(dg, hier=[4, 8, 6], imbalance=0.03, seed=1)] + \
               [heipa.Pool.job(d, imbalance=0.03, seed=1, **s) for d, s in small]
        pool.run(jobs)
        assert all(x.status == heipa.OK for x in jobs)
        assert np.array_equal(heipa.to_numpy_u32(jobs[0]._part)[:dg.n], want)
        for (d, s), x in zip(small, jobs[1:]):
            assert np.array_equal(heipa.to_numpy_u32(x._part)[:d.n], _alone(ctx, d, s))
    finally:
        pool.close()


def test_pool_autotune_reports_a_failed_sample_job(ctx):
    """A malformed graph in the autotuner's sample ends the tuning with that
    job's status; the pool keeps working afterwards."""
    reqs = _requests()[:3]
    bad = heipa.HostGraph(np.array([0, 1, 2], np.uint32), np.array([1, 7], np.uint32))
    bad_dg = heipa.DeviceGraph(bad, "cuda:0")
    pool = heipa.Pool(2, 512 << 20)
    try:
        sample = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs]
        sample.append(heipa.Pool.job(bad_dg, k=2, imbalance=0.03, seed=1))
        with pytest.raises(heipa.HeipaError) as e:
            pool.autotune(sample)
        assert e.value.code == heipa.ERR_INVALID
        jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **s) for dg, s in reqs]
        pool.run(jobs)
        for (dg, s), j in zip(reqs, jobs):
            assert j.status == heipa.OK
            assert np.array_equal(heipa.to_numpy_u32(j._part)[:dg.n], _alone(ctx, dg, s)), s
    finally:
        pool.close()

