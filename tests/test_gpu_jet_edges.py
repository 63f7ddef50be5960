"""GPU parity of Jet on the inputs the oracle's brute-force pins were built
around (rebalancing corner cases and whole refine calls on small graphs), in
every device form of Jet: one workgroup with its state in LDS or in device
memory, and the multi-launch form replayed from HIP graphs or launched
kernel by kernel.  Every traced iteration and the result must equal the
oracle's, bit for bit."""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from brute import heavy_block_case, slight_overload_case, two_heavy_blocks_case, vertex_weighted
from graphs import block_weights, random_partition
from test_gpu_parity import _dev, _trace_tuple

pytestmark = pytest.mark.gpu

FORMS = {
    "block-lds": {heipa.OPT_BLOCK_JET: 1 << 20, heipa.OPT_BLOCK_LDS: 1},
    "block-mem": {heipa.OPT_BLOCK_JET: 1 << 20, heipa.OPT_BLOCK_LDS: 0},
    "multi-graph": {heipa.OPT_BLOCK_JET: 0, heipa.OPT_GRAPHS: 1},
    "multi-eager": {heipa.OPT_BLOCK_JET: 0, heipa.OPT_GRAPHS: 0},
}
DEFAULTS = {heipa.OPT_BLOCK_JET: 2048, heipa.OPT_BLOCK_LDS: 1, heipa.OPT_GRAPHS: 1}


def _random_case(make, k, skew, seed):
    g = make()
    return g, random_partition(g.n, k, seed, skew), k, math.ceil(1.03 * int(g.g_weight) / k)


def _ranges(g, sizes, lmax):
    """Blocks of consecutive vertex ids with the given sizes."""
    part = np.repeat(np.arange(len(sizes)), sizes).astype(np.uint32)
    assert len(part) == g.n
    return g, part, len(sizes), lmax


def _cases():
    out = {}
    for s in (1, 2, 3):
        out[f"heavy{s}"] = heavy_block_case(s)[:4]  # weak finds nothing, strong runs
    for s in (1, 2):
        out[f"slight{s}"] = slight_overload_case(s)[:4]  # too heavy for weak, buckets per weight
    # a neighbour block at floor(0.99 lmax) takes nothing
    out["fullnb"] = _ranges(heipa.gen_grid(10, 9), [40, 30, 20], 31)
    # no block below max_b_w = 298 for the weak iteration (every vertex stays),
    # but two below the strong iteration's max(opt + 1, 298) = 301
    out["nounder"] = _ranges(heipa.gen_grid(30, 30), [303, 299, 298], 302)
    out["rgg60w"] = _random_case(lambda: heipa.gen_rgg(60, 6.0, 3, 9), 4, 0.8, 1)
    out["grid8k6"] = _random_case(lambda: heipa.gen_grid(8, 8), 6, 1.0, 2)
    out["rgg70vw"] = _random_case(lambda: vertex_weighted(heipa.gen_rgg(70, 6.0, 4, 9), 4, 40), 3, 3.0, 2)
    out["rgg50vwk6"] = _random_case(lambda: vertex_weighted(heipa.gen_rgg(50, 5.0, 7, 9), 5, 60), 6, 0.0, 3)
    # strong candidates of block 0 in bucket 49 leak into block 1's range (A11)
    for s in (1, 2):
        out[f"leak{s}"] = two_heavy_blocks_case(s)[:4]
    # full rows keep entries of weight 0 that strong rebalance counts
    out["zerorow"] = _random_case(lambda: vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80), 4, 1.0, 3)
    # lmax within 0.1-1% of W / k: strong evicts by min(weight, excess)
    for k, eps, skew, seed in ((4, 0.001, 0.5, 1), (2, 0.003, 2.0, 1), (3, 0.01, 2.0, 2)):
        g = vertex_weighted(heipa.gen_grid(8, 8), 3, 20)
        out[f"tight{k}"] = (g, random_partition(g.n, k, seed, skew), k, math.ceil((1 + eps) * int(g.g_weight) / k))
    return out


CASES = _cases()


def _with_form(ctx, form, fn):
    for opt, v in FORMS[form].items():
        ctx.set_option(opt, v)
    try:
        fn()
    finally:
        for opt, v in DEFAULTS.items():
            ctx.set_option(opt, v)


def _same(name, gtr, otr):
    assert len(gtr) == len(otr), (name, len(gtr), len(otr))
    for i, (a, b) in enumerate(zip(gtr, otr)):
        assert _trace_tuple(a) == _trace_tuple(b), (name, i, _trace_tuple(a), _trace_tuple(b))


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("ultra,level", [(False, 0), (False, 1), (True, 0)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_jet_edges(ctx, name, ultra, level, form):
    g, part0, k, lmax = CASES[name]
    cut0, maxw0 = orc.edge_cut(g, part0), int(block_weights(g, part0, k).max())
    op, ocut, omaxw, otr = orc.jet_refine(g, part0, k, lmax, ultra, level, cut0, maxw0, trace_cap=4096)
    print(f"jet edges {name} ultra={ultra} level={level} form={form} n={g.n} k={k}", flush=True)

    def run():
        dg = _dev(g)
        pd = heipa.from_numpy_u32(part0)
        gcut, gmaxw, gtr = ctx.jet_refine(dg, pd, k, lmax, ultra, level, cut0, maxw0, trace_cap=4096)
        _same(name, gtr, otr)
        assert (gcut, gmaxw) == (ocut, omaxw)
        assert np.array_equal(heipa.to_numpy_u32(pd), op), name

    _with_form(ctx, form, run)


MAP_HIERS = {2: ([2], [3]), 3: ([3], [4]), 4: ([2, 2], [1, 10]), 6: ([3, 2], [1, 5]), 8: ([2, 2, 2], [1, 10, 100])}


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_map_refine_edges(ctx, name, form):
    g, part0, k, lmax = CASES[name]
    D = orc.distance_matrix(*MAP_HIERS[k])
    Dd = heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))
    J0, maxw0 = orc.comm_cost(g, part0, D), int(block_weights(g, part0, k).max())
    op, oJ, omaxw, otr = orc.map_refine(g, part0, D, lmax, J0, maxw0, trace_cap=4096)
    print(f"map edges {name} form={form} n={g.n} k={k}", flush=True)

    def run():
        dg = _dev(g)
        pd = heipa.from_numpy_u32(part0)
        gJ, gmaxw, gtr = ctx.map_refine(dg, pd, Dd, k, lmax, 0, J0, maxw0, trace_cap=4096)
        _same(name, gtr, otr)
        assert (gJ, gmaxw) == (oJ, omaxw)
        assert np.array_equal(heipa.to_numpy_u32(pd), op), name

    _with_form(ctx, form, run)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_map_refine_zero_entries(ctx, form):
    """Distances 7:1 (not growing with the level): a block whose connection
    fell to 0 in a full row wins the mapping LP (promap:237-246)."""
    g = heipa.gen_rmat(7, 0, 8.0, 0.57, 0.19, 0.19, 6, 9, True)
    D = orc.distance_matrix([3, 2], [7, 1])
    k = D.shape[0]
    part0 = random_partition(g.n, k, 4, 3.0)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    Dd = heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))
    J0, maxw0 = orc.comm_cost(g, part0, D), int(block_weights(g, part0, k).max())
    op, oJ, omaxw, otr = orc.map_refine(g, part0, D, lmax, J0, maxw0, trace_cap=4096)
    print(f"map zero entries form={form}", flush=True)

    def run():
        pd = heipa.from_numpy_u32(part0)
        gJ, gmaxw, gtr = ctx.map_refine(_dev(g), pd, Dd, k, lmax, 0, J0, maxw0, trace_cap=4096)
        _same("zero", gtr, otr)
        assert (gJ, gmaxw) == (oJ, omaxw)
        assert np.array_equal(heipa.to_numpy_u32(pd), op)

    _with_form(ctx, form, run)


LARGE_K = heipa.gen_rgg(1 << 14, 10.0, 9, 10)


@pytest.mark.parametrize("form", ["multi-graph", "multi-eager"])
@pytest.mark.parametrize("ultra,level", [(False, 0), (False, 1), (True, 0)])
@pytest.mark.parametrize("k", [700, 1500, 3000])
def test_jet_large_k(ctx, k, ultra, level, form):
    """Wide k: the sweep keeps fewer, wider lane groups per workgroup (one
    LDS counter array of 12k bytes per group), and with 5-25 vertices per
    block most iterations rebalance."""
    g = LARGE_K
    part0 = (np.arange(g.n, dtype=np.int64) * k // g.n).astype(np.uint32)
    rnd = random_partition(g.n, k, 11, 0.3)
    flip = np.random.default_rng(5).random(g.n) < 0.15
    part0 = np.where(flip, rnd, part0).astype(np.uint32)
    lmax = math.ceil(1.03 * g.n / k)
    cut0, maxw0 = orc.edge_cut(g, part0), int(block_weights(g, part0, k).max())
    op, ocut, omaxw, otr = orc.jet_refine(g, part0, k, lmax, ultra, level, cut0, maxw0, trace_cap=8192)
    print(f"jet large k={k} ultra={ultra} level={level} form={form}", flush=True)

    def run():
        pd = heipa.from_numpy_u32(part0)
        gcut, gmaxw, gtr = ctx.jet_refine(_dev(g), pd, k, lmax, ultra, level, cut0, maxw0, trace_cap=8192)
        _same(f"k{k}", gtr, otr)
        assert (gcut, gmaxw) == (ocut, omaxw)
        assert np.array_equal(heipa.to_numpy_u32(pd), op)

    _with_form(ctx, form, run)


@pytest.mark.parametrize("n", [2062, 2063, 2200, 2268, 2269, 2282, 2290, 2450])
def test_jet_block_lds_boundary(ctx, n):
    """The single-workgroup form keeps a level's state in LDS only if the
    state, the dynamic work areas and the kernel's own __shared__ variables
    fit the 160 KiB together (the static part once pushed a launch over the
    limit).  With the form allowed for any n and k = 8, the work areas take
    49,312 bytes (the resolve step's), the state 50 n + 64 and the static
    part is counted as at least 1 KiB: 2200 vertices fit with room, 2268 is
    the largest n that keeps its state in LDS, 2269-2282 would still fit the
    kernel's actual static LDS, 2283-2289 would have crossed the limit by the
    static part alone, and from 2290 on the dynamic part alone does not fit.
    The BlockConn row ids and flags (5 n bytes; these graphs have vertices of
    degree >= k, so the rows exist) join the state up to n = 2062 and stay in
    device memory from 2063 on.  Every result equals the oracle's."""
    g = heipa.gen_rgg(n, 8.0, 3, 10)
    k = 8
    part0 = random_partition(g.n, k, 2, 1.0)
    lmax = math.ceil(1.03 * g.n / k)
    cut0, maxw0 = orc.edge_cut(g, part0), int(block_weights(g, part0, k).max())
    op, ocut, omaxw, otr = orc.jet_refine(g, part0, k, lmax, False, 0, cut0, maxw0, trace_cap=4096)
    print(f"jet block boundary n={n}", flush=True)

    def run():
        pd = heipa.from_numpy_u32(part0)
        gcut, gmaxw, gtr = ctx.jet_refine(_dev(g), pd, k, lmax, False, 0, cut0, maxw0, trace_cap=4096)
        _same(f"n{n}", gtr, otr)
        assert (gcut, gmaxw) == (ocut, omaxw)
        assert np.array_equal(heipa.to_numpy_u32(pd), op)

    _with_form(ctx, "block-lds", run)
