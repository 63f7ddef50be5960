"""GPU parity of the IM process-mapping path against the CPU oracle,
bit-exact, through the C-ABI on cuda:0: the mapping form of Jet
(ProMap_jet_refine) iteration by iteration, and the whole IM solver."""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from graphs import block_weights, random_partition
from test_gpu_parity import GRAPHS, _dev, _trace_tuple

pytestmark = pytest.mark.gpu

HIERS = {"4:8:6": ([4, 8, 6], [1, 10, 100]), "2:3": ([2, 3], [1, 7]), "2:2:2": ([2, 2, 2], [3, 5, 40]),
         "5": ([5], [1])}


@pytest.fixture(params=sorted(GRAPHS))
def graph(request):
    return request.param, GRAPHS[request.param]


def _D(hier, dist):
    D = orc.distance_matrix(hier, dist)
    return D, heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))


def _start(g, k, skew, seed=11):
    part0 = (np.arange(g.n, dtype=np.int64) * k // max(g.n, 1)).astype(np.uint32)
    rnd = random_partition(g.n, k, seed, skew)
    flip = np.random.default_rng(5).random(g.n) < 0.15
    return np.where(flip, rnd, part0).astype(np.uint32)


@pytest.mark.parametrize("hier,skew,slack", [("4:8:6", 0.0, 1.03), ("2:3", 0.6, 1.03), ("2:2:2", 0.9, 1.03),
                                             ("5", 0.3, 1.10), ("2:3", 0.0, 1.0)])
def test_map_refine(ctx, graph, hier, skew, slack):
    name, g = graph
    dg = _dev(g)
    D, Dd = _D(*HIERS[hier])
    k = D.shape[0]
    part0 = _start(g, k, skew)
    vw = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    lmax = math.ceil(slack * vw.sum() / k)
    J0 = orc.comm_cost(g, part0, D)
    maxw0 = int(block_weights(g, part0, k).max())
    op, oJ, omaxw, otr = orc.map_refine(g, part0, D, lmax, J0, maxw0, trace_cap=4096)
    pd = heipa.from_numpy_u32(part0)
    gJ, gmaxw, gtr = ctx.map_refine(dg, pd, Dd, k, lmax, 0, J0, maxw0, trace_cap=4096)
    assert len(gtr) == len(otr), (name, len(gtr), len(otr))
    for i, (a, b) in enumerate(zip(gtr, otr)):
        assert _trace_tuple(a) == _trace_tuple(b), (name, i, _trace_tuple(a), _trace_tuple(b))
    assert (gJ, gmaxw) == (oJ, omaxw)
    gp = heipa.to_numpy_u32(pd)
    assert np.array_equal(gp, op), name
    assert orc.comm_cost(g, gp, D) == gJ


@pytest.mark.parametrize("graphs", [1, 0])
@pytest.mark.parametrize("hier,skew", [("4:8:6", 0.0), ("2:2:2", 0.9)])
def test_map_refine_multi_launch(ctx, graph, graphs, hier, skew):
    """Forces the multi-launch form (replayed HIP graphs or kernel by kernel) on
    levels small enough for the single-workgroup form."""
    ctx.set_option(heipa.OPT_BLOCK_JET, 0)
    ctx.set_option(heipa.OPT_GRAPHS, graphs)
    try:
        test_map_refine(ctx, graph, hier, skew, 1.03)
    finally:
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)
        ctx.set_option(heipa.OPT_GRAPHS, 1)


@pytest.mark.parametrize("lds", [1, 0])
def test_map_refine_block_state(ctx, lds):
    """The mapping objective in the single-workgroup form, per-vertex state in
    LDS or in device memory."""
    ctx.set_option(heipa.OPT_BLOCK_JET, 1 << 20)
    ctx.set_option(heipa.OPT_BLOCK_LDS, lds)
    try:
        for hier, skew in (("4:8:6", 0.0), ("2:2:2", 0.9)):
            test_map_refine(ctx, ("grid40", heipa.gen_grid(40, 40)), hier, skew, 1.03)
    finally:
        ctx.set_option(heipa.OPT_BLOCK_JET, 2048)
        ctx.set_option(heipa.OPT_BLOCK_LDS, 1)


@pytest.mark.parametrize("cap", [0, 16])
def test_map_refine_global_resolve(ctx, graph, cap):
    ctx.set_option(heipa.OPT_RESOLVE_CAP, cap)
    try:
        test_map_refine(ctx, graph, "2:2:2", 0.9, 1.03)
    finally:
        ctx.set_option(heipa.OPT_RESOLVE_CAP, 1 << 32)


@pytest.mark.parametrize("hier", ["4:8:6", "2:3", "2:2:2"])
def test_im_partition(ctx, graph, hier):
    name, g = graph
    h, dist = HIERS[hier]
    dg = _dev(g)
    part, st, lv = ctx.im_partition(dg, h, dist, 0.03, 1, level_cap=64)
    op, oJ, omaxw, olv = orc.im_partition(g, h, dist, 0.03, 1)
    assert len(lv) == len(olv), name
    for a, b in zip(lv, olv):
        assert (a.n, a.m, a.nc, a.mapping_hash, a.cut, a.maxw) == (b.n, b.m, b.nc, b.mapping_hash, b.cut, b.maxw), name
    assert np.array_equal(heipa.to_numpy_u32(part), op), name
    assert (st.cut, st.maxw) == (oJ, omaxw)
    k = int(np.prod(h))
    assert st.lmax == math.ceil(1.03 * g.g_weight / k)


def test_map_refine_rejects_null_distances(ctx):
    g = heipa.gen_grid(8, 8)
    dg = _dev(g)
    pd = heipa.from_numpy_u32(np.zeros(g.n, np.uint32))
    with pytest.raises(heipa.HeipaError):
        ctx._check("heipa_map_refine", ctx.L.heipa_map_refine(ctx.h, heipa.ctypes.byref(dg.desc), pd.data_ptr(),
                                                               None, None, 4, 100, 0, heipa.ctypes.byref(
                                                                   heipa.ctypes.c_int64(0)),
                                                               heipa.ctypes.byref(heipa.ctypes.c_int64(0))))
