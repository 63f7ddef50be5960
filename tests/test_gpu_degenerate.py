"""GPU parity on degenerate inputs: the empty graph, graphs without edges,
isolated vertices beside a component, k = 1, k larger than n, and a star.
The whole solve, its cut and maximum block weight must equal the oracle's."""
import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from graphs import csr_from_edges, star_plus_grid
from test_gpu_parity import _dev

pytestmark = pytest.mark.gpu


def _grid_edges(rows, cols):
    return [(r * cols + c, r * cols + c + 1) for r in range(rows) for c in range(cols - 1)] + \
           [(r * cols + c, (r + 1) * cols + c) for r in range(rows - 1) for c in range(cols)]


CASES = {
    "empty": (lambda: csr_from_edges(0, []), 4),
    "edgeless100": (lambda: csr_from_edges(100, []), 4),
    "edgeless_vw": (lambda: csr_from_edges(90, [], None, list(range(1, 91))), 3),
    "one_edge": (lambda: csr_from_edges(2, [(0, 1)]), 2),
    "isolated_and_grid": (lambda: csr_from_edges(60, _grid_edges(4, 5)), 3),
    "path_k_above_n": (lambda: csr_from_edges(10, [(i, i + 1) for i in range(9)]), 16),
    "grid_k1": (lambda: heipa.gen_grid(10, 10), 1),
    "star_k2": (lambda: star_plus_grid(200, 2, 3), 2),
    "star_k8": (lambda: star_plus_grid(200, 2, 3), 8),
}


@pytest.mark.parametrize("ultra", [False, True])
@pytest.mark.parametrize("name", sorted(CASES))
def test_degenerate_partition(ctx, name, ultra):
    make, k = CASES[name]
    g = make()
    print(f"degenerate {name} n={g.n} m={g.m} k={k} ultra={ultra}", flush=True)
    op, ocut, omaxw, olv = orc.partition(g, k, 0.03, 1, ultra)
    dg = _dev(g)
    p, st, _ = ctx.partition(dg, k, 0.03, 1, ultra)
    assert (st.cut, st.maxw) == (ocut, omaxw), name
    assert np.array_equal(heipa.to_numpy_u32(p)[:g.n], op), name
    if g.n:
        assert ctx.edge_cut(dg, heipa.from_numpy_u32(op)) == ocut


MAPPING = ["edgeless100", "edgeless_vw", "empty", "isolated_and_grid", "one_edge", "path_k_above_n", "star_k8"]


@pytest.mark.parametrize("hier,dist", [([2, 2], [1, 10]), ([3, 4], [1, 5]), ([1, 2], [2, 3])])
@pytest.mark.parametrize("name", MAPPING)
def test_degenerate_im_and_hm(ctx, name, hier, dist):
    """The IM solver and the HM recursion on the same inputs (the hierarchy
    product may exceed n; a level of 1 splits nothing)."""
    g = CASES[name][0]()
    print(f"degenerate mapping {name} n={g.n} hier={hier}", flush=True)
    dg = _dev(g)
    op, oJ, omaxw, _ = orc.im_partition(g, hier, dist, 0.03, 1)
    p, st, _ = ctx.im_partition(dg, hier, dist, 0.03, 1)
    assert (st.cut, st.maxw) == (oJ, omaxw), name
    assert np.array_equal(heipa.to_numpy_u32(p)[:g.n], op), name
    oh = orc.hm_partition(g, hier, 0.03, 1, False)
    ph = ctx.hm_partition(dg, hier, 0.03, 1, False)
    assert np.array_equal(heipa.to_numpy_u32(ph)[:g.n], oh), name
