"""GPU parity: every stage of the HIP path against the CPU oracle, bit-exact
(integer work), through the C-ABI on cuda:0."""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from graphs import block_weights, many_hubs, random_partition, star_plus_grid

pytestmark = pytest.mark.gpu


def _vertex_weighted(g, hi, seed):
    """Same structure with vertex weights uniform in [1, hi] (coarse levels of
    the solver are vertex-weighted; the generators are not)."""
    vw = np.random.default_rng(seed).integers(1, hi + 1, g.n).astype(np.int32)
    return heipa.HostGraph(g.xadj, g.adjncy, g.adjwgt, vw)


def _graphs():
    return {
        "grid64": heipa.gen_grid(64, 64),
        "rgg12vw": _vertex_weighted(heipa.gen_rgg(1 << 12, 10.0, 6, 10), 20, 3),
        "grid_rect": heipa.gen_grid(37, 91),
        "rgg14w": heipa.gen_rgg(1 << 14, 12.0, 1, 10),
        "rgg12u": heipa.gen_rgg(1 << 12, 8.0, 5, 0),
        "rmat12w": heipa.gen_rmat(12, 0, 16.0, 0.57, 0.19, 0.19, 2, 100, True),
        "rmat11u": heipa.gen_rmat(11, 0, 8.0, 0.57, 0.19, 0.19, 3, 0, False),
        "star": star_plus_grid(),
        "hubs100": many_hubs(),
    }


GRAPHS = _graphs()


@pytest.fixture(params=sorted(GRAPHS))
def graph(request):
    return request.param, GRAPHS[request.param]


def _dev(g):
    return heipa.DeviceGraph(g, "cuda:0")


def test_edge_cut_comm_cost(ctx, graph):
    name, g = graph
    dg = _dev(g)
    for k, seed in ((2, 1), (7, 2), (48, 3)):
        part = random_partition(g.n, k, seed)
        pd = heipa.from_numpy_u32(part)
        assert ctx.edge_cut(dg, pd) == orc.edge_cut(g, part), (name, k)
    hier, dist = [4, 8, 6], [1, 10, 100]
    D = orc.distance_matrix(hier, dist)
    part = random_partition(g.n, 192, 9)
    Dd = heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))
    assert ctx.comm_cost(dg, heipa.from_numpy_u32(part), Dd, 192) == orc.comm_cost(g, part, D), name
    bw, mx = ctx.block_weights(dg, heipa.from_numpy_u32(part), 192)
    assert np.array_equal(bw, block_weights(g, part, 192))


def test_match(ctx, graph):
    name, g = graph
    dg = _dev(g)
    mp, nc, st = ctx.match(dg, 12345)
    omp, onc, ost = orc.match(g, 12345)
    assert nc == onc, (name, nc, onc)
    assert np.array_equal(heipa.to_numpy_u32(mp), omp), name
    for f in ("rounds", "after_hem", "after_leaf", "after_twin", "after_relative", "ran_leaf", "ran_twin",
              "ran_relative"):
        assert getattr(st, f) == getattr(ost, f), (name, f, getattr(st, f), getattr(ost, f))


def test_contract_project(ctx, graph):
    name, g = graph
    dg = _dev(g)
    omp, onc, _ = orc.match(g, 12345)
    cg = ctx.contract(dg, heipa.from_numpy_u32(omp), onc)
    cx, ca, cw, cv, rc = orc.contract(g, omp, onc)
    assert rc == 0
    assert np.array_equal(cg.xadj, cx), name
    assert np.array_equal(cg.adjncy, ca), name
    assert np.array_equal(cg.adjwgt, cw), name
    assert np.array_equal(cg.vwgt, cv), name
    cpart = random_partition(onc, 5, 4)
    fine = ctx.project(heipa.from_numpy_u32(omp), heipa.from_numpy_u32(cpart), g.n)
    assert np.array_equal(heipa.to_numpy_u32(fine), orc.project(omp, cpart))


def _coarsened(g, levels):
    for _ in range(levels):
        mp, nc, _ = orc.match(g, 12345)
        cx, ca, cw, cv, rc = orc.contract(g, mp, nc)
        assert rc == 0
        g = heipa.HostGraph(cx, ca, cw, cv)
    return g


COARSE = {
    "rmat8u_l2": lambda: _coarsened(heipa.gen_rmat(8, 0, 6.0, 0.57, 0.19, 0.19, 6, 0, False), 2),
    "rmat11u_l1": lambda: _coarsened(GRAPHS["rmat11u"], 1),
    "rmat11u_l3": lambda: _coarsened(GRAPHS["rmat11u"], 3),
    "rmat12w_l2": lambda: _coarsened(GRAPHS["rmat12w"], 2),
    "star_l1": lambda: _coarsened(GRAPHS["star"], 1),
    "hubs100_l2": lambda: _coarsened(GRAPHS["hubs100"], 2),
    "rgg14w_l3": lambda: _coarsened(GRAPHS["rgg14w"], 3),
}


@pytest.mark.parametrize("name", sorted(COARSE))
def test_match_contract_coarse_levels(ctx, name):
    """Matching and contraction on coarse levels (vertex weights, merged edge
    weights; the relative pass runs on the R-MAT levels)."""
    g = COARSE[name]()
    print(f"coarse level {name} n={g.n}", flush=True)
    test_match(ctx, (name, g))
    test_contract_project(ctx, (name, g))


@pytest.mark.parametrize("region_edges", [0, 16])
@pytest.mark.parametrize("name", sorted(COARSE))
def test_match_repick_pieces(ctx, name, region_edges):
    """Repick pieces of a full workgroup (0) or of 8 list entries (16) give
    the same matching on the coarse levels."""
    g = COARSE[name]()
    ctx.set_option(heipa.OPT_REGION_EDGES, region_edges)
    try:
        test_match(ctx, (name, g))
    finally:
        ctx.set_option(heipa.OPT_REGION_EDGES, 1024)


def _trace_tuple(t):
    return (t.ratio_idx, t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash)


@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.0, 8), (False, 1, 0.0, 8), (True, 0, 0.0, 8),
                                                (False, 0, 0.6, 8), (True, 2, 0.9, 8), (False, 0, 0.3, 3),
                                                (True, 1, 0.0, 5), (False, 1, 0.7, 7)])
def test_jet(ctx, graph, ultra, level, skew, k):
    name, g = graph
    dg = _dev(g)
    # start from a coarse-ish partition: blocks of consecutive ids, perturbed
    part0 = (np.arange(g.n, dtype=np.int64) * k // max(g.n, 1)).astype(np.uint32)
    rnd = random_partition(g.n, k, 11, skew)
    flip = np.random.default_rng(5).random(g.n) < 0.15
    part0 = np.where(flip, rnd, part0).astype(np.uint32)
    vw = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    lmax = math.ceil(1.03 * vw.sum() / k)
    cut0 = orc.edge_cut(g, part0)
    maxw0 = int(block_weights(g, part0, k).max())
    op, ocut, omaxw, otr = orc.jet_refine(g, part0, k, lmax, ultra, level, cut0, maxw0, trace_cap=4096)
    pd = heipa.from_numpy_u32(part0)
    import torch
    bwd = torch.full((k,), -1, dtype=torch.int64, device="cuda:0")
    gcut, gmaxw, gtr = ctx.jet_refine(dg, pd, k, lmax, ultra, level, cut0, maxw0, trace_cap=4096, bweights=bwd)
    assert len(gtr) == len(otr), (name, len(gtr), len(otr))
    for i, (a, b) in enumerate(zip(gtr, otr)):
        assert _trace_tuple(a) == _trace_tuple(b), (name, i, _trace_tuple(a), _trace_tuple(b))
    assert (gcut, gmaxw) == (ocut, omaxw)
    gp = heipa.to_numpy_u32(pd)
    assert np.array_equal(gp, op), name
    assert orc.edge_cut(g, gp) == gcut
    # the returned partition's block weights (partition.bweights after copy_into)
    bw = bwd.cpu().numpy()
    assert np.array_equal(bw, block_weights(g, op, k)) and bw.max() == gmaxw, name


@pytest.mark.parametrize("graphs", [1, 0])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.0, 8), (True, 0, 0.0, 8), (False, 0, 0.6, 8),
                                                (True, 2, 0.9, 8), (False, 0, 0.3, 3), (False, 1, 0.7, 7)])
def test_jet_multi_launch(ctx, graph, graphs, ultra, level, skew, k):
    """Small graphs run Jet in one workgroup by default; this forces the
    multi-launch form on them, with batches replayed from a captured HIP graph
    (every replay is checked through the per-iteration trace) or launched
    kernel by kernel."""
    ctx.set_option(heipa.OPT_BLOCK_JET, 0)
    ctx.set_option(heipa.OPT_GRAPHS, graphs)
    try:
        test_jet(ctx, graph, ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)
        ctx.set_option(heipa.OPT_GRAPHS, 1)


@pytest.mark.parametrize("block", [1 << 20, 0])
@pytest.mark.parametrize("cap", [0, 16])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.6, 8), (True, 2, 0.9, 8), (False, 1, 0.7, 7)])
def test_jet_global_resolve(ctx, graph, cap, block, ultra, level, skew, k):
    """Rebalance iterations whose resolve list exceeds the LDS capacity take the
    device-wide radix-sort path (from either Jet form); a capacity of 0 sends
    every rebalance there, 16 mixes both paths within one refinement."""
    ctx.set_option(heipa.OPT_RESOLVE_CAP, cap)
    ctx.set_option(heipa.OPT_BLOCK_JET, block)
    try:
        test_jet(ctx, graph, ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_RESOLVE_CAP, 1 << 32)
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)


SMALL = {"grid40": lambda: heipa.gen_grid(40, 40), "rgg2kw": lambda: heipa.gen_rgg(2000, 10.0, 4, 10),
         "rmat11u": lambda: GRAPHS["rmat11u"], "rgg12vw": lambda: GRAPHS["rgg12vw"]}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("lds", [1, 0])
@pytest.mark.parametrize("cap", [0, 1 << 32])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.6, 8), (True, 2, 0.9, 8), (False, 1, 0.7, 7),
                                                (True, 0, 0.3, 64)])
def test_jet_block_state(ctx, name, lds, cap, ultra, level, skew, k):
    """The single-workgroup form with its per-vertex state in LDS or in device
    memory; a resolve capacity of 0 stalls every rebalance (the state goes
    back to device memory for the host's global path, then the next launch
    copies it in again)."""
    ctx.set_option(heipa.OPT_BLOCK_JET, 1 << 20)
    ctx.set_option(heipa.OPT_BLOCK_LDS, lds)
    ctx.set_option(heipa.OPT_RESOLVE_CAP, cap)
    try:
        test_jet(ctx, (name, SMALL[name]()), ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)
        ctx.set_option(heipa.OPT_BLOCK_LDS, 1)
        ctx.set_option(heipa.OPT_RESOLVE_CAP, 1 << 32)


@pytest.mark.parametrize("opts", [{heipa.OPT_BLOCK_JET: 0, heipa.OPT_GROUP_CAP: 4},
                                  {heipa.OPT_BLOCK_JET: 1 << 20, heipa.OPT_BLOCK_GROUP_CAP: 64},
                                  {heipa.OPT_BLOCK_JET: 1 << 20, heipa.OPT_BLOCK_GROUP_CAP: 4}],
                         ids=["multi-cap4", "block-cap64", "block-cap4"])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.6, 8), (True, 2, 0.9, 8), (True, 0, 0.3, 64)])
def test_jet_group_caps(ctx, graph, opts, ultra, level, skew, k):
    """Lane groups narrower than the level's average degree (afterburner,
    apply and rows of the multi-launch form; the single-workgroup form) or as
    wide as it give the same iterations."""
    for o, v in opts.items():
        ctx.set_option(o, v)
    try:
        test_jet(ctx, graph, ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)
        ctx.set_option(heipa.OPT_GROUP_CAP, 64)
        ctx.set_option(heipa.OPT_BLOCK_GROUP_CAP, 8)


@pytest.mark.parametrize("copies", [2, 32])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.6, 8), (True, 2, 0.9, 8), (False, 1, 0.7, 7)])
def test_jet_bucket_copies(ctx, graph, copies, ultra, level, skew, k):
    """Rebalance minibucket totals spread over several copies (automatic only on
    large levels) give the same iterations."""
    ctx.set_option(heipa.OPT_BUCKET_COPIES, copies)
    ctx.set_option(heipa.OPT_BLOCK_JET, 0)
    try:
        test_jet(ctx, graph, ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_BUCKET_COPIES, 0)
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)


@pytest.mark.parametrize("region_edges", [0, 16])
@pytest.mark.parametrize("ultra,level,skew,k", [(False, 0, 0.6, 8), (True, 2, 0.9, 8), (True, 0, 0.3, 64)])
def test_jet_region_sizes(ctx, graph, region_edges, ultra, level, skew, k):
    """List regions of one vertex per sweep thread (0) or of 8 vertices (16
    adjacency entries per region forces the floor on every level) in the
    multi-launch form give the same iterations."""
    ctx.set_option(heipa.OPT_REGION_EDGES, region_edges)
    ctx.set_option(heipa.OPT_BLOCK_JET, 0)
    try:
        test_jet(ctx, graph, ultra, level, skew, k)
    finally:
        ctx.set_option(heipa.OPT_REGION_EDGES, 1024)
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)


@pytest.mark.parametrize("k,ultra", [(4, False), (16, True), (32, False), (3, False), (7, True)])
def test_partition(ctx, graph, k, ultra):
    name, g = graph
    dg = _dev(g)
    part, st, lv = ctx.partition(dg, k, 0.03, 1, ultra, level_cap=64)
    op, ocut, omaxw, olv = orc.partition(g, k, 0.03, 1, ultra)
    assert len(lv) == len(olv), name
    for a, b in zip(lv, olv):
        assert (a.n, a.m, a.nc, a.mapping_hash, a.cut, a.maxw) == (b.n, b.m, b.nc, b.mapping_hash, b.cut, b.maxw), name
    assert np.array_equal(heipa.to_numpy_u32(part), op), name
    assert (st.cut, st.maxw) == (ocut, omaxw)
    assert st.lmax == math.ceil((1.03) * g.g_weight / k)


@pytest.mark.parametrize("hier", [[4, 8, 6], [2, 4, 8], [3, 1, 2]])
def test_hm(ctx, hier):
    g = heipa.gen_rmat(12, 0, 16.0, 0.57, 0.19, 0.19, 7, 100, True)
    dg = _dev(g)
    gp = heipa.to_numpy_u32(ctx.hm_partition(dg, hier, 0.03, 1, False))
    op = orc.hm_partition(g, hier, 0.03, 1, False)
    assert np.array_equal(gp, op)
    k = int(np.prod(hier))
    D = orc.distance_matrix(hier, [1, 10, 100])
    Dd = heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))
    assert ctx.comm_cost(dg, heipa.from_numpy_u32(gp), Dd, k) == orc.comm_cost(g, op, D)


def test_hm_worker_out_of_memory_falls_back(ctx):
    """Worker arenas of 1 MiB cannot hold a child's solve: each worker hands
    its unfinished subtrees back to the calling context, and the mapping is the
    oracle's."""
    ctx.set_option(heipa.OPT_HM_WORKER_BYTES, 1 << 20)
    try:
        test_hm(ctx, [4, 8, 6])
    finally:
        ctx.set_option(heipa.OPT_HM_WORKER_BYTES, 0)


@pytest.mark.parametrize("workers", [0, 2])
def test_hm_workers(ctx, workers):
    """The subtrees below the first split solved one after another (0) or on
    two worker contexts for six children (in turns) give the same mapping."""
    ctx.set_option(heipa.OPT_HM_WORKERS, workers)
    try:
        test_hm(ctx, [4, 8, 6])
        test_hm(ctx, [3, 1, 2])
    finally:
        ctx.set_option(heipa.OPT_HM_WORKERS, 8)
