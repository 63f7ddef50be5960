"""GPU error paths of the C-ABI: malformed graphs, block ids out of range and
bad arguments are rejected with HEIPA_ERR_INVALID and a message, by the host
checks or the validation launches that run before any computation, and the
context stays usable afterwards (the reference exits the process instead,
host_graph.h:148-151)."""
import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu


def _raw(xadj, adjncy, adjwgt=None, vwgt=None):
    return heipa.HostGraph(np.array(xadj, np.uint32), np.array(adjncy, np.uint32),
                           None if adjwgt is None else np.array(adjwgt, np.int32),
                           None if vwgt is None else np.array(vwgt, np.int32))


def _hub_with_self_loop():
    """Vertex 0 adjacent to itself and 40 leaves: a row longer than one
    thread scans, so the wave-wide self-loop scan finds it."""
    leaves = list(range(1, 41))
    xadj = [0, 41] + [41 + i for i in range(1, 41)]
    return _raw(xadj, [0] + leaves + [0] * 40)


BAD = {
    "xadj_not_monotone": (lambda: _raw([0, 3, 2, 4], [1, 2, 0, 1]), "monotone"),
    "neighbour_out_of_range": (lambda: _raw([0, 1, 2], [1, 7]), "neighbour id >= n"),
    "zero_edge_weight": (lambda: _raw([0, 1, 2], [1, 0], adjwgt=[0, 0]), "edge weights must be positive"),
    "negative_vertex_weight": (lambda: _raw([0, 1, 2], [1, 0], vwgt=[1, -2]), "vertex weights must be non-negative"),
    "self_loop": (lambda: _raw([0, 2, 3], [0, 1, 0]), "self loops"),
    "hub_self_loop": (_hub_with_self_loop, "self loops"),
}


def _good(ctx):
    g = heipa.gen_grid(6, 7)
    part = (np.arange(g.n) % 3).astype(np.uint32)
    cut = ctx.edge_cut(_dev(g), heipa.from_numpy_u32(part))
    assert cut > 0
    return cut


@pytest.mark.parametrize("name", sorted(BAD))
def test_malformed_graph_is_rejected(ctx, name):
    make, msg = BAD[name]
    g = make()
    print(f"malformed graph {name} n={g.n} m={g.m}", flush=True)
    before = _good(ctx)
    dg = _dev(g)
    part = heipa.from_numpy_u32(np.zeros(g.n, np.uint32))
    with pytest.raises(heipa.HeipaError) as e:
        ctx.edge_cut(dg, part)
    assert e.value.code == heipa.ERR_INVALID and msg in str(e.value), str(e.value)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.partition(dg, 2)
    assert e.value.code == heipa.ERR_INVALID and msg in str(e.value), str(e.value)
    assert _good(ctx) == before  # the context is still usable


def test_block_id_out_of_range_is_rejected(ctx):
    g = heipa.gen_grid(6, 7)
    dg = _dev(g)
    part = (np.arange(g.n) % 3).astype(np.uint32)
    part[17] = 4
    pd = heipa.from_numpy_u32(part)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.jet_refine(dg, pd, 4, 100, False, 0, 0, 0)
    assert e.value.code == heipa.ERR_INVALID and "block id >= k" in str(e.value)
    assert np.array_equal(heipa.to_numpy_u32(pd), part)  # untouched
    with pytest.raises(heipa.HeipaError) as e:
        ctx.block_weights(dg, pd, 4)
    assert e.value.code == heipa.ERR_INVALID
    _good(ctx)


def test_bad_arguments_are_rejected(ctx):
    g = heipa.gen_grid(6, 7)
    dg = _dev(g)
    pd = heipa.from_numpy_u32(np.zeros(g.n, np.uint32))
    with pytest.raises(heipa.HeipaError) as e:
        ctx.jet_refine(dg, pd, 0, 100, False, 0, 0, 0)  # k = 0
    assert e.value.code == heipa.ERR_INVALID
    with pytest.raises(heipa.HeipaError) as e:
        ctx.partition(dg, 0)
    assert e.value.code == heipa.ERR_INVALID
    with pytest.raises(heipa.HeipaError) as e:
        ctx.partition(dg, 4, imbalance=-0.1)
    assert e.value.code == heipa.ERR_INVALID and "imbalance" in str(e.value)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.hm_partition(dg, [2, 2], -0.1, 1, False)
    assert e.value.code == heipa.ERR_INVALID and "imbalance" in str(e.value)
    with pytest.raises(heipa.HeipaError) as e:  # product 2^32 wraps u32 arithmetic
        ctx.hm_partition(dg, [65536, 65536], 0.03, 1, False)
    assert e.value.code == heipa.ERR_INVALID and "product" in str(e.value)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.hm_partition(dg, [2, 32769], 0.03, 1, False)
    assert e.value.code == heipa.ERR_INVALID and "product" in str(e.value)
    _good(ctx)


def test_im_rejects_a_product_beyond_the_jet_limit_before_the_matrix(ctx):
    """An IM call whose mapping refinement would exceed the Jet k limit is
    rejected before the k x k distance matrix is built."""
    g = heipa.gen_rgg(40000, 8.0, 3, 10)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.im_partition(_dev(g), [4, 1024], [1, 10])
    assert e.value.code == heipa.ERR_INVALID and "too large" in str(e.value)
    _good(ctx)


def test_block_weights_beyond_the_lds_histogram(ctx):
    """k above 8192 sums block weights with global atomics: heipa_block_weights,
    a whole partition (n <= 8k, so no Jet) and a one-level HM pool job."""
    g = heipa.gen_rgg(30000, 8.0, 5, 10)
    dg = _dev(g)
    rng = np.random.default_rng(3)
    part = rng.integers(0, 9000, g.n).astype(np.uint32)
    bw, mw = ctx.block_weights(dg, heipa.from_numpy_u32(part), 9000)
    ref = np.bincount(part, minlength=9000)
    assert np.array_equal(np.asarray(bw), ref) and mw == ref.max()
    gp, st, _ = ctx.partition(dg, 8193, 0.03, 1, False)
    op, cut, maxw, _ = orc.partition(g, 8193, 0.03, 1, False)
    assert np.array_equal(heipa.to_numpy_u32(gp), op) and st.cut == cut and st.maxw == maxw
    hp = heipa.to_numpy_u32(ctx.hm_partition(dg, [8193], 0.03, 1, False))
    assert np.array_equal(hp, orc.hm_partition(g, [8193], 0.03, 1, False))


def test_arena_exhaustion_is_reported_and_recovered(ctx):
    """A 16 MiB arena cannot hold the coarsening of an RGG with 2^18 vertices
    (its contraction tables alone take about 25 MB): the solve returns
    HEIPA_ERR_OOM, leaves the arena as it was, and the same context then
    solves a small graph with the result a large context gives."""
    small = heipa.gen_grid(12, 12)
    ref, _, _ = ctx.partition(_dev(small), 4)
    ref = heipa.to_numpy_u32(ref).copy()
    big = _dev(heipa.gen_rgg(1 << 18, 12.0, 3, 10))
    c = heipa.Context(0, 16 << 20)
    try:
        for _ in range(2):
            with pytest.raises(heipa.HeipaError) as e:
                c.partition(big, 16)
            assert e.value.code == heipa.ERR_OOM, str(e.value)
            p, _, _ = c.partition(_dev(small), 4)
            assert np.array_equal(heipa.to_numpy_u32(p), ref)
    finally:
        c.close()


def test_coarse_weight_overflow_is_reported(ctx):
    """Coarse weights that do not fit the s32 graph format return
    HEIPA_ERR_OVERFLOW (the reference wraps them): merged edge weights, merged
    vertex weights, and a whole solve whose first contraction overflows."""
    from graphs import csr_from_edges
    big = 2**31 - 1
    g = csr_from_edges(4, [(0, 2), (1, 2), (0, 3)], [big, big, 1])
    mp = heipa.from_numpy_u32(np.array([0, 0, 1, 2], np.uint32))  # (0,2) and (1,2) merge
    with pytest.raises(heipa.HeipaError) as e:
        ctx.contract(_dev(g), mp, 3)
    assert e.value.code == heipa.ERR_OVERFLOW, str(e.value)
    g = csr_from_edges(4, [(0, 1), (1, 2), (2, 3)], None, [big, big, 1, 1])
    with pytest.raises(heipa.HeipaError) as e:
        ctx.contract(_dev(g), mp, 3)  # vertices 0 and 1 merge
    assert e.value.code == heipa.ERR_OVERFLOW, str(e.value)
    path = csr_from_edges(40, [(i, i + 1) for i in range(39)], None, [1 << 30] * 40)
    with pytest.raises(heipa.HeipaError) as e:
        ctx.partition(_dev(path), 2)  # n > 8k: any matched pair weighs 2^31
    assert e.value.code == heipa.ERR_OVERFLOW, str(e.value)
    _good(ctx)


def test_jet_k_limit_is_reported(ctx):
    """The sweep keeps a 12k-byte LDS counter array per lane group (at most
    48 KiB) and the hub sweep one array plus 4k bytes of block weights (at
    most 64 KiB), so one Jet call takes k <= 4093; k = 4094 is rejected
    before any Jet launch, and k = 4093 runs."""
    g = heipa.gen_grid(6, 7)
    dg = _dev(g)
    pd = heipa.from_numpy_u32(np.zeros(g.n, np.uint32))
    with pytest.raises(heipa.HeipaError) as e:
        ctx.jet_refine(dg, pd, 4094, 100, False, 0, 0, g.n)
    assert e.value.code == heipa.ERR_INVALID and "k too large" in str(e.value), str(e.value)
    cut, maxw, _ = ctx.jet_refine(dg, pd, 4093, 1, False, 0, 0, g.n)
    assert maxw <= g.n and cut == ctx.edge_cut(dg, pd)
    _good(ctx)
