"""Brute-force Jet iterations for the oracle's CPU pins: connectivities and
gains from whole-graph cut evaluations, rebalancing selections stated as
sorts (rebalance_weak, jet:592-751; rebalance_strong, jet:356-590)."""
import math

import numpy as np

import heipa
import refsim
from graphs import block_weights

U64 = (1 << 64) - 1


def cut(g, part):
    src = np.repeat(np.arange(g.n), np.diff(g.xadj.astype(np.int64)))
    w = np.ones(len(g.adjncy), np.int64) if g.adjwgt is None else g.adjwgt.astype(np.int64)
    tot = int((w * (part[src] != part[g.adjncy.astype(np.int64)])).sum())
    assert tot % 2 == 0
    return tot // 2


def moved(part, moves):
    p = part.copy()
    for u, b in moves.items():
        p[u] = b
    return p


def connectivity(g, part, u, k, c0):
    """Connectivity of u to every block, from cut evaluations: its own block's
    is the cut increase of a move to an empty block, another block's is that
    plus the cut drop of the move there."""
    own = cut(g, moved(part, {u: k})) - c0
    out = {int(part[u]): own}
    for b in {int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]} - {int(part[u])}:
        out[b] = own + c0 - cut(g, moved(part, {u: b}))
    return out


def weak_iteration(g, part, k, lmax):
    vw = (lambda u: 1) if g.vwgt is None else (lambda u: int(g.vwgt[u]))
    W = int(g.g_weight)
    bw = [int(x) for x in block_weights(g, part, k)]
    opt = (W + k - 1) // k
    max_b_w = max(int(lmax * 0.99), lmax - 100)
    sections = 128
    if (g.n + sections * k) // (sections * k) < 4096:
        sections = (g.n + 4096 * k) // (4096 * k)
    under = [b for b in range(k) if bw[b] < max_b_w]
    c0 = cut(g, part)
    dest, key = {}, {}
    for u in range(g.n):
        src = int(part[u])
        if not (bw[src] > lmax and vw(u) < 1.5 * (bw[src] - opt)):
            continue
        conn = connectivity(g, part, u, k, c0)
        own = conn[src]
        room = [(c, b) for b, c in conn.items() if b != src and bw[b] < max_b_w and c > 0]
        if room:
            best_c = max(c for c, _ in room)
            b = min(b for c, b in room if c == best_c)
            gain = best_c - own
        elif under:
            b = under[u % len(under)]
            gain = -own
        else:
            continue  # the reference divides by zero here (jet:689); the vertex stays
        dest[u] = b
        key[u] = (src, refsim.gain_bucket_exact(gain, vw(u)), u % sections, u)
    moves = []
    for src in range(k):
        before = 0
        for u in sorted((u for u in dest if key[u][0] == src), key=lambda u: key[u]):
            if before < bw[src] - lmax:
                moves.append(u)
            before += vw(u)
    moves.sort()
    return len(dest), moves, dest


def strong_iteration(g, part, k, lmax):
    """rebalance_strong (jet:356-590).  Candidates: vertices of an overloaded
    block no heavier than twice the largest room and lighter than twice the
    block's excess over opt; score = mean connectivity to the blocks with room
    minus the own connectivity, bucketed per min(weight, excess over lmax).
    Each overloaded block evicts, in (bucket, id mod sections, id) order,
    while the weight before a candidate is below its excess over lmax.  The
    evicted vertices of all blocks, block by block in that order, are laid on
    one line of weight positions, which the rooms of the blocks below
    max_b_w cut into consecutive chunks; a vertex goes to the chunk that
    holds its position if at least half its weight fits there, otherwise
    (in id order) it is laid again at the end of the line, and it stays when
    its position is past the last chunk."""
    vw = (lambda u: 1) if g.vwgt is None else (lambda u: int(g.vwgt[u]))
    W = int(g.g_weight)
    bw = [int(x) for x in block_weights(g, part, k)]
    opt = (W + k - 1) // k
    max_b_w = max(opt + 1, int(lmax * 0.99))
    sections = 128
    if (g.n + sections * k) // (sections * k) < 4096:
        sections = (g.n + 4096 * k) // (4096 * k)
    rooms = [max_b_w - bw[b] for b in range(k) if bw[b] < max_b_w]
    if not rooms:
        return 0, [], {}
    max_room = max(rooms)
    c0 = cut(g, part)
    key = {}
    for u in range(g.n):
        src = int(part[u])
        if not (bw[src] > lmax and vw(u) <= 2 * max_room and vw(u) < 2 * (bw[src] - opt)):
            continue
        conn = connectivity(g, part, u, k, c0)
        others = [c for b, c in conn.items() if b != src and bw[b] < max_b_w]
        gain = sum(others) // max(len(others), 1) - conn[src]
        key[u] = (src, refsim.gain_bucket_exact(gain, min(vw(u), bw[src] - lmax)), u % sections, u)
    pos, line = {}, 0
    for src in range(k):
        before = 0
        for u in sorted((u for u in key if key[u][0] == src), key=lambda u: key[u]):
            if before < bw[src] - lmax:
                pos[u] = line + before
            before += vw(u)
        line += sum(vw(u) for u in pos if key[u][0] == src)
    ends, acc = [], 0
    for b in range(k):
        acc += max(0, max_b_w - bw[b])
        ends.append(acc)
    dest = {}
    for u in sorted(pos):
        x = pos[u]
        while True:
            chunk = next((b for b in range(k) if x < ends[b]), None)
            if chunk is None:
                dest[u] = int(part[u])
                break
            if vw(u) // 2 <= ends[chunk] - x:
                dest[u] = chunk
                break
            x, line = line, line + vw(u)
    return len(key), sorted(pos), dest


def vertex_weighted(g, seed, hi):
    vw = np.random.default_rng(seed).integers(1, hi + 1, g.n).astype(np.int32)
    return heipa.HostGraph(g.xadj, g.adjncy, g.adjwgt, vw)


def heavy_block_case(seed):
    """Block 0 holds 12 vertices of weight 60..75 and weighs opt + 39, so
    every one of them is at least 1.5 (block weight - opt) (no weak
    candidate: the weak iteration finds no move and the strong one runs
    first) and below 2 (block weight - opt) (all strong candidates); their
    weights exceed the block's excess over lmax, which then weights their
    gain buckets.  Block 1 is light enough to take any of them."""
    g0 = heipa.gen_rgg(200, 7.0, 6, 9)
    rng = np.random.default_rng(seed)
    heavy = rng.choice(g0.n, 12, replace=False)
    vw = np.zeros(g0.n, np.int64)
    vw[heavy] = rng.integers(60, 76, 12)
    b0 = int(vw.sum())
    opt = b0 - 39
    rest = [u for u in range(g0.n) if vw[u] == 0]
    vw[rest] = rng.integers(1, 21, len(rest))
    vw[rest[-1]] += 4 * opt - int(vw.sum())  # total weight 4 opt
    assert vw[rest[-1]] >= 1
    g = heipa.HostGraph(g0.xadj, g0.adjncy, g0.adjwgt, vw.astype(np.int32))
    k = 4
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    part = np.zeros(g.n, np.uint32)
    bw = [b0, 0, 0, 0]
    for u in rest:  # blocks 2 and 3 up to opt + 15, block 1 the rest
        b = next((b for b in (2, 3) if bw[b] + vw[u] <= opt + 15), 1)
        part[u] = b
        bw[b] += int(vw[u])
    assert (int(g.g_weight) + k - 1) // k == opt and bw[0] > lmax
    assert all(1.5 * (bw[0] - opt) <= vw[u] < 2 * (bw[0] - opt) for u in heavy)
    assert all(vw[u] > bw[0] - lmax for u in heavy) and max(bw[1:]) <= lmax
    return g, part, k, lmax, bw


def slight_overload_case(seed):
    """Vertex weights 1..80; block 0 takes, in a random vertex order, every
    vertex that still fits under lmax, then the lightest other vertex that
    takes it over lmax; every other vertex goes to the lightest other block.
    The heaviest vertices of block 0 are then no candidates (weight at least
    1.5 (block weight - opt)), and which candidate moves first depends on the
    gain per unit of weight."""
    g = vertex_weighted(heipa.gen_rgg(200, 7.0, 5, 9), 6, 80)
    k = 8
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    w = [int(x) for x in g.vwgt]
    order = [int(u) for u in np.random.default_rng(seed).permutation(g.n)]
    part = np.full(g.n, k, np.uint32)
    bw = np.zeros(k, np.int64)
    for u in order:
        if bw[0] + w[u] <= lmax:
            part[u] = 0
            bw[0] += w[u]
    last = min((u for u in order if part[u] == k and bw[0] + w[u] > lmax), key=lambda u: (w[u], u))
    part[last] = 0
    bw[0] += w[last]
    for u in order:
        if part[u] == k:
            b = 1 + int(np.argmin(bw[1:]))
            part[u] = b
            bw[b] += w[u]
    return g, part, k, lmax, bw
