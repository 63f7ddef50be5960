"""heipa_edge_cut_noalloc: the edge cut with no device allocation (a static
device array holds its partial sums), through the C-ABI, against
heipa_edge_cut on the same inputs.  Two calls in a row with the second on a
smaller input show that nothing the first call left in the shared array leaks
into the second."""
import ctypes

import numpy as np
import pytest
import torch

import heipa
from graphs import many_hubs
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu


def _noalloc(dg, part):
    fn = heipa.lib().heipa_edge_cut_noalloc
    out = torch.full((1,), -7, dtype=torch.int64, device="cuda:0")
    stream = torch.cuda.current_stream()
    rc = fn(ctypes.byref(dg.desc), part.data_ptr() if part is not None else None, out.data_ptr(),
            ctypes.c_void_p(stream.cuda_stream))
    assert rc == heipa.OK, rc
    return int(out.item())  # .item() waits for the torch stream the kernels ran on


def _case(ctx, g, k, seed):
    dg = _dev(g)
    part = np.random.default_rng(seed).integers(0, k, max(g.n, 1)).astype(np.uint32)[:g.n]
    pd = heipa.from_numpy_u32(part) if g.n else None
    ref = ctx.edge_cut(dg, pd) if g.n else 0
    return ref, _noalloc(dg, pd)


@pytest.mark.parametrize("big,small", [
    (lambda: heipa.gen_rgg(1 << 19, 12.0, 3, 10), lambda: heipa.gen_grid(9, 11)),       # 2048 groups, then 1
    (lambda: heipa.gen_rmat(15, 0, 16.0, 0.57, 0.19, 0.19, 4, 100, True), lambda: many_hubs()),  # hubs
    (lambda: heipa.gen_rgg(1 << 14, 8.0, 5, 0), lambda: heipa.gen_rgg(700, 8.0, 6, 0)),  # uniform weights
])
def test_noalloc_edge_cut_matches_edge_cut(ctx, big, small):
    for i, make in enumerate((big, small)):
        g = make()
        ref, got = _case(ctx, g, 13, 100 + i)
        assert got == ref, (i, g.n, got, ref)


def test_noalloc_edge_cut_empty_and_malformed(ctx):
    # an empty graph after a large one: the cut is 0 (no partial-sum launch)
    _case(ctx, heipa.gen_grid(300, 300), 4, 1)
    e = heipa.HostGraph(np.zeros(1, np.uint32), np.zeros(0, np.uint32), None, None)
    assert _noalloc(_dev(e), None) == 0
    # a neighbour id >= n is never dereferenced; the result is -1 where
    # heipa_edge_cut reports HEIPA_ERR_INVALID
    bad = heipa.HostGraph(np.array([0, 1, 2, 3], np.uint32), np.array([1, 0, 7], np.uint32), None, None)
    dg = _dev(bad)
    pd = heipa.from_numpy_u32(np.zeros(3, np.uint32))
    with pytest.raises(heipa.HeipaError):
        ctx.edge_cut(dg, pd)
    assert _noalloc(dg, pd) == -1
    # and a valid call afterwards is unaffected
    ref, got = _case(ctx, heipa.gen_grid(40, 40), 5, 2)
    assert got == ref
