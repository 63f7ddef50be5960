"""CPU pin of the oracle's plain Jet label propagation (jet_lp,
src/refinement/jet_label_propagation.h:182-354) by brute force.

The reference ships no tests or fixtures for Jet, so its first LP iteration
is recomputed here from whole-graph edge-cut evaluations alone:
  * a vertex's connectivity to its own block is the cut increase when it
    moves to a block nobody else is in;
  * the gain of moving it to an adjacent block b is the cut drop of that lone
    move, and the destination is the adjacent block with the largest drop
    (the smaller id on ties, jet:219-221);
  * a move with negative gain stays a candidate only while the loss is below
    floor(ratio * own connectivity) (jet:227-228); the ratio is 0.25 at the
    finest level, 0.75 above it and 0.85 first in the ultra schedule
    (jet:953-961);
  * the afterburner (jet:269-328) keeps a candidate when its gain plus the
    change of that gain once the candidates ordered before it (larger gain,
    then smaller id) have moved is not negative; the change is recomputed
    from cut evaluations of the graph with and without those moves.
The oracle's trace of that iteration must agree on the candidate count, the
gain hash, the move count, the move hash and the cut after the moves.

The first weak rebalance iteration (rebalance_weak, jet:592-751) is pinned the
same way: connectivities and gains come from cut evaluations, the gain bucket
from the exact-arithmetic restatement (itself pinned in test_oracle.py), and
the selection is stated as a sort, not as the bucket prefix sums the oracle
and the reference use: in every overloaded block the candidates are ordered by
(gain bucket, id mod sections, id) and a candidate moves while the weight of
the candidates before it is below the block's excess over lmax.

Whole jet_refine calls (jet:926-1033) are then run over these brute-force
iterations with the reference's decision cache and locks, and every traced
iteration and the result must agree with the oracle.
"""
import math

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
from brute import U64, Rows, cut as _cut, heavy_block_case, two_heavy_blocks_case, moved as _moved, slight_overload_case, strong_iteration as _strong_iteration, \
    vertex_weighted as _vertex_weighted, weak_iteration as _weak_iteration
from graphs import block_weights, random_partition


def _first_lp_iteration(g, part, k, ratio):
    c0 = _cut(g, part)
    gain, dest = {}, {}
    for u in range(g.n):
        own_b = int(part[u])
        own = _cut(g, _moved(part, {u: k})) - c0  # block k holds no vertex
        blocks = sorted({int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]} - {own_b})
        best, best_b = None, None
        for b in blocks:
            drop = c0 - _cut(g, _moved(part, {u: b}))
            if best is None or drop > best:
                best, best_b = drop, b
        if best_b is None:
            continue
        if best >= 0 or -best < math.floor(ratio * own):
            gain[u], dest[u] = best, best_b
    moves = []
    for u in sorted(gain):
        first = {int(v): dest[int(v)] for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]
                 if int(v) in gain and (gain[int(v)] > gain[u] or (gain[int(v)] == gain[u] and v < u))}
        p1 = _moved(part, first)
        after = _cut(g, p1) - _cut(g, _moved(p1, {u: dest[u]}))
        before = c0 - _cut(g, _moved(part, {u: dest[u]}))
        if gain[u] + (after - before) >= 0:
            moves.append(u)
    gh = sum(orc.mix64((orc.mix64(u) + (gain[u] & U64)) & U64) for u in gain) & U64
    mh = sum(orc.mix64((u << 32) | dest[u]) for u in moves) & U64
    return len(gain), moves, gh, mh, _cut(g, _moved(part, {u: dest[u] for u in moves}))


TINY = [
    ("rgg80w", lambda: heipa.gen_rgg(80, 7.0, 3, 9), 4),
    ("grid9x8", lambda: heipa.gen_grid(9, 8), 3),
    ("rmat6w", lambda: heipa.gen_rmat(6, 0, 6.0, 0.57, 0.19, 0.19, 5, 9, True), 5),
]


@pytest.mark.parametrize("name,make,k", TINY)
@pytest.mark.parametrize("ultra,level,ratio", [(False, 0, 0.25), (False, 2, 0.75), (True, 0, 0.85)])
@pytest.mark.parametrize("seed", [1, 2])
def test_jet_lp_first_iteration_brute_force(name, make, k, ultra, level, ratio, seed):
    g = make()
    part = random_partition(g.n, k, seed)
    lmax = int(g.g_weight)  # never overloaded: the first iteration is LP
    n_cand, moves, gh, mh, cut1 = _first_lp_iteration(g, part, k, ratio)
    cut0 = orc.edge_cut(g, part)
    _, _, _, tr = orc.jet_refine(g, part, k, lmax, ultra, level, cut0, int(block_weights(g, part, k).max()),
                                 trace_cap=64)
    assert moves, "a random partition of these graphs always has improving moves"
    t = tr[0]
    assert (t.ratio_idx, t.iteration, t.mode) == (0, 1, 0)
    assert (t.n_cand, t.n_moves, t.gain_hash, t.moves_hash, t.cut) == (n_cand, len(moves), gh, mh, cut1)


def test_jet_lp_ratio_changes_the_candidates():
    """The three ratios give different candidate sets on one input, so a
    wrong ratio for a level fails the test above."""
    g = heipa.gen_rgg(80, 7.0, 3, 9)
    part = random_partition(g.n, 4, 1)
    counts = {r: _first_lp_iteration(g, part, 4, r)[0] for r in (0.25, 0.75, 0.85)}
    assert len(set(counts.values())) == 3, counts


WEAK = [
    ("rgg120w", lambda: heipa.gen_rgg(120, 7.0, 3, 9), 4, 0.8),
    ("grid10x9", lambda: heipa.gen_grid(10, 9), 3, 1.0),
    ("rmat7w", lambda: heipa.gen_rmat(7, 0, 6.0, 0.57, 0.19, 0.19, 5, 9, True), 6, 0.6),
    ("rgg150vw", lambda: _vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80), 4, 0.8),
]


@pytest.mark.parametrize("name,make,k,skew", WEAK)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jet_weak_first_iteration_brute_force(name, make, k, skew, seed):
    g = make()
    part = random_partition(g.n, k, seed, skew)
    _check_weak(g, part, k)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_jet_weak_slight_overload_weighted(seed):
    """Block 0 just over lmax with vertex weights 1..80 (slight_overload_case):
    its heaviest vertices are no candidates, and which candidate moves first
    depends on the gain per unit of weight."""
    g, part, k, lmax, bw = slight_overload_case(seed)
    opt = (int(g.g_weight) + k - 1) // k
    assert (g.vwgt[part == 0] >= 1.5 * (bw[0] - opt)).any()
    _check_weak(g, part, k)


def test_jet_weak_full_neighbour_block_is_skipped():
    """Row-major grid ids split 40 / 30 / 20 over k = 3: lmax = 31, and block 1
    weighs 30 = floor(0.99 lmax), so it takes no vertex although it is below
    lmax; block 0's vertices all go to block 2, the only one with room."""
    g = heipa.gen_grid(10, 9)
    part = np.array([0] * 40 + [1] * 30 + [2] * 20, np.uint32)
    _check_weak(g, part, 3)


def _check_weak(g, part, k):
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    bw = block_weights(g, part, k)
    assert bw.max() > lmax  # the first iteration is a weak rebalance
    _, _, _, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part), int(bw.max()), trace_cap=64)
    assert (tr[0].ratio_idx, tr[0].iteration, tr[0].mode) == (0, 1, 1)
    _check_iteration(g, tr[0], part, *_weak_iteration(g, part, k, lmax))


def _check_iteration(g, t, part, n_cand, moves, dest):
    """Compares one traced iteration with a brute-force move set; returns the
    partition after the moves."""
    assert moves
    after = _moved(part, {u: dest[u] for u in moves})
    mh = sum(orc.mix64((u << 32) | dest[u]) for u in moves) & U64
    assert (t.n_cand, t.n_moves, t.moves_hash, t.cut) == (n_cand, len(moves), mh, _cut(g, after))
    return after


STRONG = [
    ("rgg120w", lambda: heipa.gen_rgg(120, 7.0, 3, 9), 8, 1.0, 1),
    ("grid10x9", lambda: heipa.gen_grid(10, 9), 6, 1.0, 2),
    ("rmat7w", lambda: heipa.gen_rmat(7, 0, 6.0, 0.57, 0.19, 0.19, 5, 9, True), 4, 1.0, 1),
    ("rgg150vw", lambda: _vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80), 6, 3.0, 2),
    ("rgg150vw", lambda: _vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80), 8, 3.0, 1),
    ("rgg150vw", lambda: _vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80), 4, 1.0, 3),
    ("rgg200vw20", lambda: _vertex_weighted(heipa.gen_rgg(200, 7.0, 5, 9), 6, 20), 3, 1.0, 3),
]


@pytest.mark.parametrize("name,make,k,skew,seed", STRONG)
def test_jet_weak_weak_strong_chain_brute_force(name, make, k, skew, seed):
    """Two weak iterations, then a strong one (jet:976-993), each recomputed
    by brute force from the partition the previous one left."""
    g = make()
    part = random_partition(g.n, k, seed, skew)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    bw = block_weights(g, part, k)
    _, _, _, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part), int(bw.max()), trace_cap=64)
    assert [t.mode for t in tr[:3]] == [1, 1, 2]
    rows = Rows(g, part, k)
    for t in tr[:2]:
        n_cand, moves, dest = _weak_iteration(g, part, k, lmax)
        part = _check_iteration(g, t, part, n_cand, moves, dest)
        rows.update(part, moves)
    _check_iteration(g, tr[2], part, *_strong_iteration(g, part, k, lmax, rows))


def test_strong_counts_the_zero_entries_of_full_rows():
    """A vertex of degree >= k whose row reached all k blocks keeps an entry
    of weight 0 when its last neighbour in a block leaves (block_conn.h:379);
    rebalance_strong counts it in the mean connectivity (jet:418-428).  In
    this chain (two weak iterations, then a strong one) such entries change
    the strong iteration's moves: the brute force with the rows' counts
    equals the oracle, the brute force without them equals the oracle with
    positive connections only, and the two differ."""
    g = _vertex_weighted(heipa.gen_rgg(150, 7.0, 4, 9), 4, 80)
    k = 4
    part = random_partition(g.n, k, 3, 1.0)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    cut0, maxw0 = orc.edge_cut(g, part), int(block_weights(g, part, k).max())
    rows = Rows(g, part, k)
    for _ in range(2):
        n_cand, moves, dest = _weak_iteration(g, part, k, lmax)
        part = _moved(part, {u: dest[u] for u in moves})
        rows.update(part, moves)
    assert any(rows.zero_entries(part, u) for u in range(g.n))
    with_rows = _strong_iteration(g, part, k, lmax, rows)
    without = _strong_iteration(g, part, k, lmax)
    assert with_rows[1] != without[1] or with_rows[2] != without[2]
    p0 = random_partition(g.n, k, 3, 1.0)
    try:
        for mode, expect in ((1, with_rows), (0, without)):
            orc.set_row_mode(mode)
            _, _, _, tr = orc.jet_refine(g, p0, k, lmax, False, 0, cut0, maxw0, trace_cap=64)
            assert [t.mode for t in tr[:3]] == [1, 1, 2]
            _check_iteration(g, tr[2], part, *expect)
    finally:
        orc.set_row_mode(1)


@pytest.mark.parametrize("seed", [1, 2, 5])
def test_jet_strong_minibucket_leak(seed):
    """Appendix A11 (jet:437 against :494-495, :527-530): strong candidates in
    bucket 49 and the last section sit on the first minibucket of the next
    block's range.  two_heavy_blocks_case puts block 0's cluster there; the
    brute force that orders them after their own block and counts their
    weight ahead of every candidate of block 1 equals the oracle, and the
    brute force without the leak does not."""
    g, part, k, lmax, bw = two_heavy_blocks_case(seed)
    _, _, _, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part), max(bw), trace_cap=64)
    assert tr[0].mode == 2
    rows = Rows(g, part, k)
    with_leak = _strong_iteration(g, part, k, lmax, rows)
    without = _strong_iteration(g, part, k, lmax, rows, a11=False)
    assert with_leak[1] != without[1] or with_leak[2] != without[2]
    _check_iteration(g, tr[0], part, *with_leak)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_jet_strong_takes_vertices_too_heavy_for_weak(seed):
    """A block whose vertices only the strong iteration may move (built by
    heavy_block_case): the weak iteration finds no move and the strong one
    runs first."""
    g, part, k, lmax, bw = heavy_block_case(seed)
    _, _, _, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part), bw[0], trace_cap=64)
    assert tr[0].mode == 2
    _check_iteration(g, tr[0], part, *_strong_iteration(g, part, k, lmax, Rows(g, part, k)))


def _lp_iteration(g, part, k, ratio, cache, locked):
    """One LP iteration with the reference's decision cache (jet:193-238: a
    vertex's destination and gain, or its lack of one, are recomputed only
    after it or a neighbour moved, otherwise reused even under a later
    ratio) and locks (jet:247-258: a vertex moved by the previous LP
    iteration is no candidate in this one).  Decisions and the afterburner
    come from cut evaluations as in the first-iteration test."""
    c0 = _cut(g, part)
    for u in range(g.n):
        if u in cache:
            continue
        own = _cut(g, _moved(part, {u: k})) - c0
        best, best_b = None, None
        for b in sorted({int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]} - {int(part[u])}):
            drop = c0 - _cut(g, _moved(part, {u: b}))
            if best is None or drop > best:
                best, best_b = drop, b
        keep = best_b is not None and (best >= 0 or -best < math.floor(ratio * own))
        cache[u] = (best, best_b) if keep else (0, None)
    gain = {u: cache[u][0] for u in range(g.n) if cache[u][1] is not None and u not in locked}
    dest = {u: cache[u][1] for u in gain}
    moves = []
    for u in sorted(gain):
        first = {int(v): dest[int(v)] for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]
                 if int(v) in gain and (gain[int(v)] > gain[u] or (gain[int(v)] == gain[u] and v < u))}
        p1 = _moved(part, first)
        after = _cut(g, p1) - _cut(g, _moved(p1, {u: dest[u]}))
        before = c0 - _cut(g, _moved(part, {u: dest[u]}))
        if gain[u] + (after - before) >= 0:
            moves.append(u)
    gh = sum(orc.mix64((orc.mix64(u) + (gain[u] & U64)) & U64) for u in gain) & U64
    return len(gain), moves, dest, gh


ULTRA = [0.85, 0.80, 0.75, 0.70, 0.65, 0.60, 0.55, 0.50, 0.45, 0.40, 0.35, 0.30, 0.25, 0.20, 0.15, 0.10, 0.05]


def _jet_refine(g, part, k, lmax, ultra, level):
    """jet_refine (jet:926-1033) over the brute-force iterations: per filter
    ratio at most 12 iterations; LP while balanced (no moves ends the
    ratio), else up to 2 weak iterations (no moves: go to strong) then strong
    (no moves ends the ratio); after each move set, a lower maximum block
    weight while the best is overloaded, or a lower cut at a balanced or no
    heavier state, becomes the best, and the first case or a cut below
    0.999 of the best restarts the iteration count."""
    w = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    cut = _cut(g, part)
    maxw = int(np.bincount(part, weights=w, minlength=k).max())
    best = (part.copy(), cut, maxw)
    cache, locked, trace = {}, set(), []
    rows = Rows(g, part, k)
    for ri, ratio in enumerate(ULTRA if ultra else [0.25 if level == 0 else 0.75]):
        bal = it = 0
        while it < 12:
            it += 1
            gh = 0
            if maxw <= lmax:
                mode = 0
                n_cand, moves, dest, gh = _lp_iteration(g, part, k, ratio, cache, locked)
                locked = set(moves)
                bal = 0
                if not moves:
                    break
            else:
                if bal < 2:
                    mode = 1
                    n_cand, moves, dest = _weak_iteration(g, part, k, lmax)
                    if not moves:
                        bal = 2
                        continue
                else:
                    mode = 2
                    n_cand, moves, dest = _strong_iteration(g, part, k, lmax, rows)
                    if not moves:
                        break
                bal += 1
            mh = sum(orc.mix64((u << 32) | dest[u]) for u in moves) & U64
            for u in moves:
                cache.pop(u, None)
                for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]:
                    cache.pop(int(v), None)
            part = _moved(part, {u: dest[u] for u in moves})
            rows.update(part, moves)
            cut = _cut(g, part)
            maxw = int(np.bincount(part, weights=w, minlength=k).max())
            trace.append((ri, it, mode, n_cand, len(moves), cut, maxw, mh, gh))
            if best[2] > lmax and maxw < best[2]:
                best = (part.copy(), cut, maxw)
                it = 0
            elif cut < best[1] and (maxw <= lmax or maxw <= best[2]):
                if cut < 0.999 * best[1]:
                    it = 0
                best = (part.copy(), cut, maxw)
    return best, trace


LOOP = [
    ("rgg60w", lambda: heipa.gen_rgg(60, 6.0, 3, 9), 4, False, 0, 0.8, 1),
    ("rgg60w-coarse", lambda: heipa.gen_rgg(60, 6.0, 3, 9), 3, False, 1, 0.0, 2),
    ("grid8x8-ultra", lambda: heipa.gen_grid(8, 8), 4, True, 0, 1.0, 3),
    ("rgg70vw-ultra", lambda: _vertex_weighted(heipa.gen_rgg(70, 6.0, 4, 9), 4, 40), 3, True, 0, 3.0, 2),
    # a lower cut at an overloaded state as heavy as the best one becomes the
    # best (jet:1010, `<=`)
    ("grid8x8-k6", lambda: heipa.gen_grid(8, 8), 6, False, 0, 1.0, 2),
    ("rgg50vw-k6", lambda: _vertex_weighted(heipa.gen_rgg(50, 5.0, 7, 9), 5, 60), 6, False, 0, 0.0, 3),
]


@pytest.mark.parametrize("name,make,k,ultra,level,skew,seed", LOOP)
def test_jet_refine_loop_brute_force(name, make, k, ultra, level, skew, seed):
    """Every iteration of a whole jet_refine call, and its result."""
    g = make()
    part = random_partition(g.n, k, seed, skew)
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    (bp, bc, bm), trace = _jet_refine(g, part, k, lmax, ultra, level)
    cut0, maxw0 = orc.edge_cut(g, part), int(block_weights(g, part, k).max())
    p, c, mw, tr = orc.jet_refine(g, part, k, lmax, ultra, level, cut0, maxw0, trace_cap=4096)
    got = [(t.ratio_idx, t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash)
           for t in tr]
    assert len(trace) > 3
    assert got == trace
    assert np.array_equal(p, bp) and (c, mw) == (bc, bm)


def test_jet_refine_loop_without_underloaded_block():
    """Blocks of 303 / 299 / 298 consecutive ids of a 30x30 grid, lmax = 302:
    no block is below the weak iteration's max_b_w = 298, so every vertex
    stays (the reference divides by zero there, jet:689) and the weak
    iteration finds no move; the strong iteration, whose max_b_w is
    max(opt + 1, 298) = 301, then runs."""
    g = heipa.gen_grid(30, 30)
    part = np.repeat(np.arange(3), [303, 299, 298]).astype(np.uint32)
    k, lmax = 3, 302
    (bp, bc, bm), trace = _jet_refine(g, part, k, lmax, False, 0)
    p, c, mw, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part), 303, trace_cap=64)
    got = [(t.ratio_idx, t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash)
           for t in tr]
    assert trace[0][1:3] == (2, 2)
    assert got == trace
    assert np.array_equal(p, bp) and (c, mw) == (bc, bm)


@pytest.mark.parametrize("k,eps,skew,seed", [(4, 0.001, 0.5, 1), (2, 0.003, 2.0, 1), (3, 0.01, 2.0, 2)])
def test_jet_refine_loop_tight_lmax(k, eps, skew, seed):
    """An 8x8 grid with vertex weights 1..20 and lmax = ceil((1 + eps) W / k)
    with eps of 0.1-1%: strong iterations meet blocks whose excess over lmax
    is below their candidates' weights, so the bucket weight
    min(weight, excess) (jet:404) decides which candidate is evicted first."""
    g = _vertex_weighted(heipa.gen_grid(8, 8), 3, 20)
    part = random_partition(g.n, k, seed, skew)
    lmax = math.ceil((1 + eps) * int(g.g_weight) / k)
    (bp, bc, bm), trace = _jet_refine(g, part, k, lmax, False, 0)
    p, c, mw, tr = orc.jet_refine(g, part, k, lmax, False, 0, orc.edge_cut(g, part),
                                  int(block_weights(g, part, k).max()), trace_cap=4096)
    got = [(t.ratio_idx, t.iteration, t.mode, t.n_cand, t.n_moves, t.cut, t.maxw, t.moves_hash, t.gain_hash)
           for t in tr]
    assert any(t[2] == 2 for t in trace)
    assert got == trace
    assert np.array_equal(p, bp) and (c, mw) == (bc, bm)
