"""CPU pins of the oracle's IM process-mapping path (ProMap_jet_refine, global
multisection, ProMapSolver::solve).

The reference ships no tests or fixtures for this path, so the oracle is
pinned by what the mathematics fixes:
  * the LP gain of a move (promap_jet_label_propagation.h:180-204) is half the
    drop of J = sum over directed edges w * D(block(u), block(v)) when the
    vertex moves alone, and the afterburner's change (:284-322) is the change
    of that gain once the neighbours ordered before it have moved: both are
    recomputed here from whole-graph J evaluations, never from the formula;
  * the incrementally tracked J equals a recount (the reference's own
    after-refinement assert, promap_solver.h via assert_state_after_partition);
  * with one hierarchy level and distance 1, J is twice the edge cut;
  * global multisection is restated a second time here in Python over the
    shared host partitioner and must agree entry for entry.
"""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from brute import Rows, heavy_block_case, strong_iteration, vertex_weighted, weak_iteration
from graphs import block_weights, random_partition

GAIN_MIN = -(1 << 63)
U64 = (1 << 64) - 1


def _J(g, part, D):
    """Brute-force comm cost: every directed CSR entry."""
    src = np.repeat(np.arange(g.n), np.diff(g.xadj.astype(np.int64)))
    w = np.ones(len(g.adjncy), np.int64) if g.adjwgt is None else g.adjwgt.astype(np.int64)
    return int((w * D[part[src], part[g.adjncy.astype(np.int64)]].astype(np.int64)).sum())


def _moved(part, moves):
    p = part.copy()
    for u, b in moves.items():
        p[u] = b
    return p


def _first_lp_iteration(g, part, D):
    """One ProMap LP iteration from scratch (no cache, no locks) by brute force."""
    J0 = _J(g, part, D)
    gain, dest = {}, {}
    for u in range(g.n):
        blocks = sorted({int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]} - {int(part[u])})
        best, best_b = -1, None
        for b in blocks:
            d2 = J0 - _J(g, _moved(part, {u: b}), D)
            assert d2 % 2 == 0
            if d2 // 2 > best:
                best, best_b = d2 // 2, b
        if best_b is not None:
            gain[u], dest[u] = best, best_b
    moves = []
    for u in sorted(gain):
        first = {int(v): dest[int(v)] for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]
                 if int(v) in gain and (gain[int(v)] > gain[u] or (gain[int(v)] == gain[u] and v < u))}
        p1 = _moved(part, first)
        after = (_J(g, p1, D) - _J(g, _moved(p1, {u: dest[u]}), D)) // 2
        before = (J0 - _J(g, _moved(part, {u: dest[u]}), D)) // 2
        if gain[u] + (after - before) >= 0:
            moves.append(u)
    gh = sum(orc.mix64((orc.mix64(u) + (gain[u] & U64)) & U64) for u in gain) & U64
    mh = sum(orc.mix64((u << 32) | dest[u]) for u in moves) & U64
    return len(gain), moves, gh, mh, _J(g, _moved(part, {u: dest[u] for u in moves}), D)


TINY = [
    ("rgg70w", lambda: heipa.gen_rgg(70, 6.0, 3, 9), [2, 2, 2], [1, 10, 100]),
    ("grid8x9", lambda: heipa.gen_grid(8, 9), [3, 2], [1, 7]),
    ("rmat6w", lambda: heipa.gen_rmat(6, 0, 6.0, 0.57, 0.19, 0.19, 5, 9, True), [2, 4], [2, 5]),
]


@pytest.mark.parametrize("name,make,hier,dist", TINY)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_map_lp_first_iteration_brute_force(name, make, hier, dist, seed):
    g = make()
    D = orc.distance_matrix(hier, dist)
    k = D.shape[0]
    part = random_partition(g.n, k, seed)
    lmax = int(g.g_weight)  # never overloaded: the first iteration is LP
    n_cand, moves, gh, mh, J1 = _first_lp_iteration(g, part, D)
    J0 = orc.comm_cost(g, part, D)
    _, _, _, tr = orc.map_refine(g, part, D, lmax, J0, int(block_weights(g, part, k).max()), trace_cap=64)
    if not moves:
        assert not tr or tr[0].iteration > 1
        return
    t = tr[0]
    assert (t.iteration, t.mode) == (1, 0)
    assert (t.n_cand, t.n_moves, t.gain_hash, t.moves_hash, t.cut) == (n_cand, len(moves), gh, mh, J1)


def _map_lp_iteration(g, part, D, cache, locked, rows=None):
    """One ProMap LP iteration (promap:206-330) with the decision cache and
    locks of the reference (a decision is recomputed only after the vertex or
    a neighbour moved; a vertex moved by the previous LP iteration is no
    candidate), decisions and afterburner from whole-graph J evaluations.
    With `rows`, a row entry of weight 0 is a candidate destination too
    (promap:237-246 walks every valid entry)."""
    J0 = _J(g, part, D)
    for u in range(g.n):
        if u in cache:
            continue
        best, best_b = -1, None
        cands = {int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]}
        if rows:
            cands |= rows.zero_entries(part, u)
        for b in sorted(cands - {int(part[u])}):
            d2 = J0 - _J(g, _moved(part, {u: b}), D)
            if d2 // 2 > best:
                best, best_b = d2 // 2, b
        cache[u] = (best, best_b) if best_b is not None else (0, None)
    gain = {u: cache[u][0] for u in range(g.n) if cache[u][1] is not None and u not in locked}
    dest = {u: cache[u][1] for u in gain}
    moves = []
    for u in sorted(gain):
        first = {int(v): dest[int(v)] for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]
                 if int(v) in gain and (gain[int(v)] > gain[u] or (gain[int(v)] == gain[u] and v < u))}
        p1 = _moved(part, first)
        after = (_J(g, p1, D) - _J(g, _moved(p1, {u: dest[u]}), D)) // 2
        before = (J0 - _J(g, _moved(part, {u: dest[u]}), D)) // 2
        if gain[u] + (after - before) >= 0:
            moves.append(u)
    gh = sum(orc.mix64((orc.mix64(u) + (gain[u] & U64)) & U64) for u in gain) & U64
    return len(gain), moves, dest, gh


def _map_refine(g, part, D, lmax, use_rows=True):
    """ProMap_jet_refine (promap:1219-1311) over brute-force iterations: at
    most 12 iterations; LP while balanced, else the balance counter goes up
    first and its first value runs a weak iteration (the rebalancing is the
    plain one, promap:333-718 differing from jet:356-751 only in the order of
    the candidate list), later ones strong; an iteration without moves is
    skipped; a lower maximum block weight while the best is overloaded, or a
    lower J at a balanced or strictly lighter state, becomes the best."""
    k = D.shape[0]
    w = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    J = _J(g, part, D)
    maxw = int(np.bincount(part, weights=w, minlength=k).max())
    best = (part.copy(), J, maxw)
    cache, locked, trace = {}, set(), []
    rows = Rows(g, part, k) if use_rows else None
    bal = it = 0
    while it < 12:
        it += 1
        gh = 0
        if maxw <= lmax:
            mode = 0
            n_cand, moves, dest, gh = _map_lp_iteration(g, part, D, cache, locked, rows)
            locked = set(moves)
            bal = 0
        else:
            bal += 1
            mode = 1 if bal < 2 else 2
            if mode == 1:
                n_cand, moves, dest = weak_iteration(g, part, k, lmax)
            else:
                n_cand, moves, dest = strong_iteration(g, part, k, lmax, rows)
        if not moves:
            continue
        mh = sum(orc.mix64((u << 32) | dest[u]) for u in moves) & U64
        for u in moves:
            cache.pop(u, None)
            for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]:
                cache.pop(int(v), None)
        part = _moved(part, {u: dest[u] for u in moves})
        if rows:
            rows.update(part, moves)
        J = _J(g, part, D)
        maxw = int(np.bincount(part, weights=w, minlength=k).max())
        trace.append((it, mode, n_cand, len(moves), J, maxw, mh, gh))
        if best[2] > lmax and maxw < best[2]:
            best = (part.copy(), J, maxw)
            it = 0
        elif J < best[1] and (maxw <= lmax or maxw < best[2]):
            if J < 0.999 * best[1]:
                it = 0
            best = (part.copy(), J, maxw)
    return best, trace


MAP_LOOP = [
    ("rgg60w", lambda: heipa.gen_rgg(60, 6.0, 3, 9), [2, 2], [1, 10], 0.8, 1),
    ("grid8x8", lambda: heipa.gen_grid(8, 8), [3, 2], [1, 7], 1.0, 2),
    ("rgg70vw", lambda: vertex_weighted(heipa.gen_rgg(70, 6.0, 4, 9), 4, 40), [2, 3], [2, 5], 3.0, 3),
    ("rmat6w", lambda: heipa.gen_rmat(6, 0, 6.0, 0.57, 0.19, 0.19, 5, 9, True), [2, 2, 2], [1, 10, 100], 0.0, 1),
    # a lower J at an overloaded state as heavy as the best one does not
    # become the best (promap:1290, strict `<`)
    ("grid8x8-k6", lambda: heipa.gen_grid(8, 8), [3, 2], [1, 5], 1.0, 1),
    ("rgg50vw-k6", lambda: vertex_weighted(heipa.gen_rgg(50, 5.0, 7, 9), 5, 60), [2, 3], [2, 9], 0.0, 3),
]


def test_map_lp_takes_the_zero_entries_of_full_rows():
    """The mapping LP walks every valid row entry (promap:237-246), so a block
    whose connection fell to 0 in a full row (block_conn.h:379) is a
    candidate destination.  With distances that grow with the hierarchy
    level D is an ultrametric and a neighbour's block always beats such a
    block; with distances 7:1 (a sibling block is farther than a block of
    the other group) it can win, here in an LP iteration.  The brute force
    with the rows' counts equals the oracle; without them it equals the
    oracle with positive connections only; the two differ."""
    g = heipa.gen_rmat(7, 0, 8.0, 0.57, 0.19, 0.19, 6, 9, True)
    D = orc.distance_matrix([3, 2], [7, 1])
    k = D.shape[0]
    part = random_partition(g.n, k, 4, 3.0)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    J0, maxw0 = orc.comm_cost(g, part, D), int(block_weights(g, part, k).max())
    results = []
    try:
        for mode in (1, 0):
            orc.set_row_mode(mode)
            (bp, bJ, bm), trace = _map_refine(g, part, D, lmax, use_rows=mode == 1)
            p, J, mw, tr = orc.map_refine(g, part, D, lmax, J0, maxw0, trace_cap=4096)
            got = [(t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash) for t in tr]
            assert got == trace
            assert np.array_equal(p, bp) and (J, mw) == (bJ, bm)
            results.append(trace)
    finally:
        orc.set_row_mode(1)
    first = next(i for i, (a, b) in enumerate(zip(*results)) if a != b)
    assert results[0][first][1] == 0  # an LP iteration


@pytest.mark.parametrize("name,make,hier,dist,skew,seed", MAP_LOOP)
def test_map_refine_loop_brute_force(name, make, hier, dist, skew, seed):
    """Every iteration of a whole ProMap_jet_refine call, and its result."""
    g = make()
    D = orc.distance_matrix(hier, dist)
    k = D.shape[0]
    part = random_partition(g.n, k, seed, skew)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    (bp, bJ, bm), trace = _map_refine(g, part, D, lmax)
    J0, maxw0 = orc.comm_cost(g, part, D), int(block_weights(g, part, k).max())
    p, J, mw, tr = orc.map_refine(g, part, D, lmax, J0, maxw0, trace_cap=4096)
    got = [(t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash) for t in tr]
    assert len(trace) > 3
    assert got == trace
    assert np.array_equal(p, bp) and (J, mw) == (bJ, bm)


@pytest.mark.parametrize("seed", [1, 2])
def test_map_refine_loop_skips_an_iteration_without_moves(seed):
    """The weak iteration finds no move in a block only the strong iteration
    may empty (heavy_block_case); the loop skips it and goes on
    (promap:1276)."""
    g, part, k, lmax, bw = heavy_block_case(seed)
    D = orc.distance_matrix([2, 2], [1, 10])
    (bp, bJ, bm), trace = _map_refine(g, part, D, lmax)
    p, J, mw, tr = orc.map_refine(g, part, D, lmax, orc.comm_cost(g, part, D), bw[0], trace_cap=4096)
    got = [(t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash) for t in tr]
    assert trace[0][:2] == (2, 2)  # iteration 1 (weak) had no move
    assert got == trace
    assert np.array_equal(p, bp) and (J, mw) == (bJ, bm)


GRAPHS = [
    ("rgg3k", lambda: heipa.gen_rgg(3000, 10.0, 2, 10)),
    ("grid50", lambda: heipa.gen_grid(50, 50)),
    ("rmat11", lambda: heipa.gen_rmat(11, 0, 8.0, 0.57, 0.19, 0.19, 9, 20, True)),
]


@pytest.mark.parametrize("name,make", GRAPHS)
@pytest.mark.parametrize("hier,dist,skew", [([4, 3], [1, 10], 0.0), ([2, 2, 3], [1, 5, 40], 0.7),
                                            ([12], [3], 0.3)])
def test_map_refine_invariants(name, make, hier, dist, skew):
    g = make()
    D = orc.distance_matrix(hier, dist)
    k = D.shape[0]
    lmax = math.ceil(1.03 * g.g_weight / k)
    part0 = random_partition(g.n, k, 4, skew)
    J0, mw0 = orc.comm_cost(g, part0, D), int(block_weights(g, part0, k).max())
    p, J, mw, tr = orc.map_refine(g, part0, D, lmax, J0, mw0, trace_cap=100000)
    assert J == orc.comm_cost(g, p, D) == _J(g, p, D)  # incremental J == recount
    assert mw == block_weights(g, p, k).max()
    assert p.max() < k
    if mw0 <= lmax:
        assert J <= J0
    for t in tr:
        assert t.n_moves > 0 and t.mode in (0, 1, 2) and 1 <= t.iteration <= 12
    # one weak iteration per rebalancing run, then strong (promap:1265-1271):
    # a weak iteration with moves directly follows an LP iteration or starts
    # the run, never another rebalance iteration
    for a, b in zip(tr, tr[1:]):
        if b.mode == 1:
            assert a.mode == 0
    p2, J2, _, _ = orc.map_refine(g, part0, D, lmax, J0, mw0)  # deterministic
    assert np.array_equal(p, p2) and J == J2


def test_map_refine_uniform_distance_is_twice_cut():
    g = heipa.gen_rgg(2000, 9.0, 5, 10)
    k = 7
    D = orc.distance_matrix([k], [1])
    lmax = math.ceil(1.03 * g.g_weight / k)
    part0 = random_partition(g.n, k, 2, 0.4)
    p, J, _, _ = orc.map_refine(g, part0, D, lmax, orc.comm_cost(g, part0, D),
                                int(block_weights(g, part0, k).max()))
    assert J == 2 * orc.edge_cut(g, p)


def _gms_python(g, hier, eps, seed, split=None):
    """global_multisection_host (global_multisection.h:139-259), restated again;
    `split(subgraph, k, imbalance)` replaces the host partitioner (the
    hierarchical multisection passes the multilevel Solver)."""
    l = len(hier)
    gk = int(np.prod(hier))
    index_vec = [1]
    for i in range(l - 1):
        index_vec.append(index_vec[i] * hier[i])
    k_rem, p = [], 1
    for h in hier:
        k_rem.append(p * h)
        p *= h
    out = np.zeros(g.n, np.uint32)
    W = int(g.g_weight)

    def rec(sg, n_to_o, level, ident):
        k = hier[level]
        lw = int(sg.vwgt.astype(np.int64).sum()) if sg.vwgt is not None else sg.n
        depth = l - len(ident)
        imb = ((1.0 + eps) * (k_rem[l - 1 - len(ident)] * W) / (gk * lw)) ** (1.0 / depth) - 1.0
        tmp = heipa.initial_partition(sg, k, imb, seed) if split is None else split(sg, k, imb)
        if len(ident) == l - 1:
            off = sum(ident[i] * index_vec[len(index_vec) - 1 - i] for i in range(l - 1))
            out[n_to_o] = off + tmp
            return
        for b in range(k):
            verts = np.nonzero(tmp == b)[0]
            if len(verts) == 0:
                continue
            loc = {int(v): i for i, v in enumerate(verts)}
            xadj, adj, ew = [0], [], []
            for v in verts:
                for j in range(sg.xadj[v], sg.xadj[v + 1]):
                    t = int(sg.adjncy[j])
                    if t in loc:
                        adj.append(loc[t])
                        ew.append(1 if sg.adjwgt is None else int(sg.adjwgt[j]))
                xadj.append(len(adj))
            vw = np.ones(len(verts), np.int32) if sg.vwgt is None else sg.vwgt[verts]
            child = heipa.HostGraph(np.array(xadj, np.uint32), np.array(adj, np.uint32), np.array(ew, np.int32),
                                    np.asarray(vw, np.int32))
            rec(child, n_to_o[verts], level - 1, ident + [b])

    rec(g, np.arange(g.n), l - 1, [])
    return out


@pytest.mark.parametrize("hier", [[4, 3], [2, 3, 2], [3, 1, 2], [5]])
def test_global_multisection_python_restatement(hier):
    g = heipa.gen_rgg(1500, 8.0, 11, 10)
    a = orc.global_multisection(g, hier, 0.03, 7)
    b = _gms_python(g, hier, 0.03, 7)
    assert np.array_equal(a, b)
    assert a.max() < np.prod(hier)


@pytest.mark.parametrize("hier,ultra", [([2, 3, 2], False), ([3, 1, 2], True), ([4, 2], False)])
def test_hm_partition_python_restatement(hier, ultra):
    """recursive_multisection_device (hierarchical_multisection.h:56-235): the
    same recursion, child graphs and leaf offsets as the global multisection,
    with every split a full multilevel solve (gpu_heipa_partition :38-54, all
    zeros for k = 1) under the adaptive imbalance.  The Solver itself is
    pinned by its own restatement in test_oracle.py."""
    g = heipa.gen_rmat(11, 0, 8.0, 0.57, 0.19, 0.19, 4, 10, True)

    def solve(sg, k, imb):
        if k == 1:
            return np.zeros(sg.n, np.uint32)
        return orc.partition(sg, k, imb, 7, ultra)[0]

    a = orc.hm_partition(g, hier, 0.03, 7, ultra)
    b = _gms_python(g, hier, 0.03, 7, solve)
    assert np.array_equal(a, b)
    assert a.max() < np.prod(hier)


@pytest.mark.parametrize("kind,hier,dist", [("rgg", [4, 8, 6], [1, 10, 100]), ("grid", [2, 2, 2], [1, 4, 16]),
                                            ("rmat", [3, 4], [1, 9])])
def test_im_partition_invariants(kind, hier, dist):
    g = {"grid": lambda: heipa.gen_grid(60, 70), "rgg": lambda: heipa.gen_rgg(20000, 10.0, 1, 10),
         "rmat": lambda: heipa.gen_rmat(11, 0, 8.0, 0.57, 0.19, 0.19, 9, 20, True)}[kind]()
    k = int(np.prod(hier))
    D = orc.distance_matrix(hier, dist)
    p, J, mw, lv = orc.im_partition(g, hier, dist, 0.03, 1)
    assert p.max() < k
    assert J == orc.comm_cost(g, p, D)
    assert mw == block_weights(g, p, k).max()
    assert lv[-1].n <= 8 * k and all(a.n > 8 * k for a in lv[:-1])  # promap_solver.h:220-221
    assert lv[0].cut == J
    # refinement never worsens J at a balanced level
    for fine, coarse in zip(lv, lv[1:]):
        if coarse.maxw <= math.ceil(1.03 * g.g_weight / k):
            assert fine.cut <= coarse.cut
    p2, J2, _, _ = orc.im_partition(g, hier, dist, 0.03, 1)
    assert np.array_equal(p, p2) and J == J2


@pytest.mark.parametrize("kind,hier,dist,imb", [("rgg", [2, 3, 2], [1, 10, 100], 0.03),
                                                ("rmat", [3, 4], [1, 9], 0.05)])
def test_im_level_loop_restated(kind, hier, dist, imb):
    """ProMapSolver::internal_solve (promap_solver.h:216-246, :299, :336-347)
    restated over the pinned pieces: coarsen while n > 8k, the global
    multisection of the coarsest graph (no refinement there), then per level
    projection and the mapping Jet under lmax = ceil((1 + imbalance) W / k),
    J and the maximum block weight carried from level to level."""
    g = {"rgg": lambda: heipa.gen_rgg(3000, 9.0, 2, 10),
         "rmat": lambda: heipa.gen_rmat(10, 0, 8.0, 0.57, 0.19, 0.19, 6, 20, True)}[kind]()
    seed = 4
    k = int(np.prod(hier))
    D = orc.distance_matrix(hier, dist)
    lmax = math.ceil((1.0 + imb) * (int(g.g_weight) / k))
    graphs, maps = [g], []
    while graphs[-1].n > 8 * k:
        mp, nc, _ = orc.match(graphs[-1], 12345)
        cx, ca, cw, cv, rc = orc.contract(graphs[-1], mp, nc)
        assert rc == 0
        maps.append(mp)
        graphs.append(heipa.HostGraph(cx, ca, cw, cv))
    assert len(maps) >= 2
    part = orc.global_multisection(graphs[-1], hier, imb, seed)
    J, maxw = orc.comm_cost(graphs[-1], part, D), int(block_weights(graphs[-1], part, k).max())
    levels = {len(maps): (J, maxw)}
    for lv in range(len(maps) - 1, -1, -1):
        part = part[maps[lv]]
        part, J, maxw, _ = orc.map_refine(graphs[lv], part, D, lmax, J, maxw)
        levels[lv] = (J, maxw)
    p, J1, mw1, lt = orc.im_partition(g, hier, dist, imb, seed)
    assert np.array_equal(p, part) and (J1, mw1) == (J, maxw)
    assert [(t.n, t.cut, t.maxw) for t in lt] == [(graphs[lv].n, *levels[lv]) for lv in range(len(graphs))]


@pytest.mark.parametrize("hier", [[4, 8, 6], [2, 3, 2], [3, 1, 2]])
def test_product_global_multisection_matches_oracle(hier):
    # the library's host routine (no GPU involved) against the oracle's restatement
    for g in (heipa.gen_rgg(1500, 8.0, 11, 10), heipa.gen_grid(30, 41),
              heipa.gen_rmat(10, 0, 8.0, 0.57, 0.19, 0.19, 4, 10, True)):
        assert np.array_equal(heipa.global_multisection(g, hier, 0.03, 3), orc.global_multisection(g, hier, 0.03, 3))
