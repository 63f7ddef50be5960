"""Racing mode (HEIPA_OPT_RACE): the reference's races resolved by atomic
order on the device (matching commit by the first proposer's CAS, minibucket
offsets and strong overflow positions from fetch-adds).  Its results vary
from run to run, so the bar is statistical: over seeds 1..10 the cut (or J)
is at most 2% above the deterministic mode's in geometric mean (a lower one
passes: single seeds scatter by several percent either way on these graphs,
and racing mode has come out up to 3% lower), and every partition meets lmax
wherever the deterministic one does.  Each seed's racing side is the
geometric mean of RUNS racing runs: one racing draw per seed scatters by up
to 14% on the 512x512 grid, which puts the standard error of a 10-draw mean
at the 2% bar itself (one such draw came out at 1.0207)."""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu

SEEDS = range(1, 11)
RUNS = 3


def _geomean(xs):
    return math.exp(sum(math.log(x) for x in xs) / len(xs))


def _balanced(g, part, k):
    vw = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    return np.bincount(part, weights=vw, minlength=k).max() <= math.ceil(1.03 * int(vw.sum()) / k)


def _race_vs_det(ctx, solve):
    ratios = []
    try:
        for s in SEEDS:
            ctx.set_option(heipa.OPT_RACE, 0)
            qd, okd = solve(s)
            ctx.set_option(heipa.OPT_RACE, 1)
            qs = []
            for _ in range(RUNS):
                qr, okr = solve(s)
                assert okr or not okd, (s, qd, qr)
                qs.append(qr)
            ratios.append(_geomean(qs) / qd)
    finally:
        ctx.set_option(heipa.OPT_RACE, 0)
    gm = _geomean(ratios)
    print(f"race/det geometric mean {gm:.4f} ratios {[round(r, 4) for r in ratios]}", flush=True)
    assert gm <= 1.02, gm
    return ratios


@pytest.mark.parametrize("name,make,k,ultra", [
    ("rgg16-ultra-k64", lambda: heipa.gen_rgg(1 << 16, 12.0, 1, 10), 64, True),
    ("grid512-default-k32", lambda: heipa.gen_grid(512, 512), 32, False),
    ("rmat15-default-k16", lambda: heipa.gen_rmat(15, 0, 16.0, 0.57, 0.19, 0.19, 1, 100, True), 16, False),
])
def test_race_partition_quality(ctx, name, make, k, ultra):
    g = make()
    dg = _dev(g)

    def solve(seed):
        p, st, _ = ctx.partition(dg, k, 0.03, seed, ultra)
        p = heipa.to_numpy_u32(p)
        assert orc.edge_cut(g, p) == st.cut
        return st.cut, _balanced(g, p, k)

    _race_vs_det(ctx, solve)


def test_race_hm_quality(ctx):
    g = heipa.gen_rmat(14, 0, 16.0, 0.57, 0.19, 0.19, 1, 100, True)
    dg = _dev(g)
    hier = [4, 8, 6]
    D = orc.distance_matrix(hier, [1, 10, 100])

    def solve(seed):
        p = heipa.to_numpy_u32(ctx.hm_partition(dg, hier, 0.03, seed, False))
        return orc.comm_cost(g, p, D), _balanced(g, p, 192)

    _race_vs_det(ctx, solve)


@pytest.mark.parametrize("case", ["heavy1", "heavy2", "leak1", "leak2"])
def test_race_jet_strong_paths(ctx, case):
    """Racing mode through whole refine calls whose first iteration is a strong
    rebalance (blocks of heavy vertices; a cluster in bucket 49): the result's
    cut and largest block weight equal a recount, and the trace's first
    iteration is strong."""
    from brute import heavy_block_case, two_heavy_blocks_case
    make = heavy_block_case if case.startswith("heavy") else two_heavy_blocks_case
    g, part0, k, lmax, bw = make(int(case[-1]))
    cut0, maxw0 = orc.edge_cut(g, part0), max(bw)
    ctx.set_option(heipa.OPT_RACE, 1)
    try:
        pd = heipa.from_numpy_u32(part0)
        cut, maxw, tr = ctx.jet_refine(_dev(g), pd, k, lmax, False, 0, cut0, maxw0, trace_cap=4096)
    finally:
        ctx.set_option(heipa.OPT_RACE, 0)
    p = heipa.to_numpy_u32(pd)
    assert tr and tr[0].mode == 2
    assert orc.edge_cut(g, p) == cut
    assert np.bincount(p, weights=g.vwgt.astype(np.int64), minlength=k).max() == maxw
