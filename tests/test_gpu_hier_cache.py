"""Resident hierarchies (HEIPA_OPT_HIER_CACHE / HEIPA_OPT_GRAPH_KEY): a graph's
coarse levels stay in device memory between calls and serve later solves of
the same graph with any k; every result equals a fresh solve's, bit for bit."""
import numpy as np
import pytest

import heipa
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu


def _fresh(ctx, dg, spec):
    if "hier" in spec:
        return heipa.to_numpy_u32(ctx.hm_partition(dg, spec["hier"], 0.03, 1, False)).copy()
    part, st, _ = ctx.partition(dg, spec["k"], 0.03, 1, False)
    return heipa.to_numpy_u32(part).copy()


SPECS = [dict(k=8), dict(k=64), dict(k=2), dict(hier=[4, 8, 6]), dict(k=8), dict(k=300)]


def test_resident_hierarchy_serves_any_k(ctx):
    """k = 8 builds the levels down to n <= 64; k = 64 stops earlier (a
    prefix); k = 2 goes deeper (appends); HM's first split (k = 6) and a
    repeat of k = 8 reuse them; every partition equals a fresh solve."""
    g = heipa.gen_rmat(15, 0, 16.0, 0.57, 0.19, 0.19, 3, 100, True)
    dg = _dev(g)
    want = [_fresh(ctx, dg, s) for s in SPECS]
    c = heipa.Context(0, 2 << 30)
    try:
        c.set_option(heipa.OPT_HIER_CACHE, 1 << 30)
        c.set_option(heipa.OPT_GRAPH_KEY, 17)
        for s, w in zip(SPECS, want):
            assert np.array_equal(_fresh(c, dg, s), w), s
        entries, nbytes, reused, built = c.hier_cache_stats()
        assert entries == 1 and nbytes > 0 and built > 0 and reused > built
        # the key now names another graph: its entry is replaced, results stay exact
        g2 = heipa.gen_rgg(1 << 15, 10.0, 4, 10)
        dg2 = _dev(g2)
        assert np.array_equal(_fresh(c, dg2, dict(k=16)), _fresh(ctx, dg2, dict(k=16)))
        assert c.hier_cache_stats()[0] == 1
    finally:
        c.close()


def test_pool_shares_resident_hierarchies(ctx):
    """Partition and HM jobs on three graphs, several k each, through a pool
    whose workers share one cache: every job equals its fresh solve, and
    later jobs reuse levels built by jobs on other workers."""
    graphs = [heipa.gen_grid(180, 200), heipa.gen_rgg(1 << 15, 12.0, 6, 100),
              heipa.gen_rmat(14, 0, 16.0, 0.57, 0.19, 0.19, 5, 100, False)]
    dgs = [_dev(g) for g in graphs]
    specs = [dict(k=16), dict(k=4), dict(hier=[2, 4, 8]), dict(k=128), dict(k=16)]
    reqs = [(gi, s) for s in specs for gi in range(len(graphs))]
    want = {(gi, i): _fresh(ctx, dgs[gi], s) for i, (gi, s) in enumerate(reqs)}
    pool = heipa.Pool(4, 1 << 30)
    try:
        pool.set_option(heipa.OPT_HIER_CACHE, 2 << 30)
        jobs = [heipa.Pool.job(dgs[gi], imbalance=0.03, seed=1, graph_key=100 + gi, **s) for gi, s in reqs]
        pool.run(jobs)
        entries, nbytes, reused, built = pool.hier_cache_stats()
    finally:
        pool.close()
    for i, ((gi, s), j) in enumerate(zip(reqs, jobs)):
        assert j.status == heipa.OK, (i, j.status)
        assert np.array_equal(heipa.to_numpy_u32(j._part)[:dgs[gi].n], want[(gi, i)]), (gi, s)
    assert entries == 3 and reused > 0 and built > 0


def test_hierarchy_cache_over_budget_evicts(ctx):
    """A budget smaller than one graph's levels: least recently used entries
    go, and results stay those of fresh solves."""
    graphs = [heipa.gen_rgg(1 << 14, 10.0, s, 10) for s in (1, 2, 3)]
    dgs = [_dev(g) for g in graphs]
    want = [_fresh(ctx, dg, dict(k=8)) for dg in dgs]
    c = heipa.Context(0, 1 << 30)
    try:
        c.set_option(heipa.OPT_HIER_CACHE, 256 << 10)
        for rep in range(2):
            for i, dg in enumerate(dgs):
                c.set_option(heipa.OPT_GRAPH_KEY, 1 + i)
                assert np.array_equal(_fresh(c, dg, dict(k=8)), want[i]), (rep, i)
        entries, nbytes, _, built = c.hier_cache_stats()
        assert built > 0 and entries <= 3
    finally:
        c.close()
