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


class Rows:
    """What the reference's BlockConn rows (block_conn.h) hold beyond the
    positive connections, stated as counts rather than slots: only a row of
    capacity k (a vertex of degree >= k) can keep an entry whose weight fell
    to 0, and only while its size is k.  Per such vertex: the row's size and
    the set of its valid entries.
      * init_BlockConn (:48-196): size = number of neighbour blocks P,
        valid = P.
      * After a move list (jet:853): every vertex with a neighbour in the
        list is updated.  A list longer than n / 10 rebuilds the row
        (update_large, :208-345): size = |P| + max(3, |P| // 4) if that is
        below k, else k; valid = P.  Otherwise (update_small, :347-466) an
        entry that falls to 0 becomes a tombstone unless the size is k, and
        a new block reuses a tombstone or an empty slot before the size
        grows: size = max(size, |P|); valid = P, or the old valid set plus P
        in a row of size k."""

    def __init__(self, g, part, k):
        self.g, self.k = g, k
        self.size, self.valid = {}, {}
        for u in range(g.n):
            if g.xadj[u + 1] - g.xadj[u] >= k:
                P = self._blocks(part, u)
                self.size[u], self.valid[u] = len(P), P

    def _blocks(self, part, u):
        g = self.g
        return {int(part[v]) for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]}

    def update(self, part, moves):
        """part: after the moves; moves: the whole list, blocks kept included"""
        g, k = self.g, self.k
        large = len(moves) > g.n // 10
        touched = {int(v) for u in moves for v in g.adjncy[g.xadj[u]:g.xadj[u + 1]]}
        for v in touched:
            if v not in self.size:
                continue
            P = self._blocks(part, v)
            if large:
                ns = len(P) + max(3, len(P) // 4)
                self.size[v], self.valid[v] = (ns if ns < k else k), P
            else:
                full = self.size[v] == k
                self.size[v] = max(self.size[v], len(P))
                self.valid[v] = (self.valid[v] | P) if full else P

    def zero_entries(self, part, u):
        """valid entries of u's row whose connection is 0"""
        if self.size.get(u) != self.k:
            return set()
        return self.valid[u] - self._blocks(part, u)


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


def strong_iteration(g, part, k, lmax, rows=None, a11=True):
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
    its position is past the last chunk.  With `rows`, the mean also counts
    the row entries of weight 0 of blocks with room (jet:418-428)."""
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
        zeros = [b for b in (rows.zero_entries(part, u) if rows else ()) if b != src and bw[b] < max_b_w]
        gain = sum(others) // max(len(others) + len(zeros), 1) - conn[src]
        key[u] = (src, refsim.gain_bucket_exact(gain, min(vw(u), bw[src] - lmax)), u % sections, u)
    # The minibucket of a candidate is (50 block + bucket) sections + id mod
    # sections, plus 1 (jet:437), while each block's range starts without it
    # (jet:494-495, :527-530): a candidate in bucket 49 and the last section
    # lands on the first minibucket of the next block's range (appendix A11).
    # It is ordered after every other candidate of its own block, and its
    # weight is ahead of every candidate of the next block.
    def leaked(u):
        return a11 and key[u][1] == 49 and key[u][2] == sections - 1
    pos, line = {}, 0
    for src in range(k):
        leak_in = sum(vw(u) for u in key if key[u][0] == src - 1 and leaked(u)) if src else 0
        mine = [u for u in key if key[u][0] == src]
        order = sorted((u for u in mine if not leaked(u)), key=lambda u: key[u]) + sorted(u for u in mine if leaked(u))
        before, end = leak_in, None
        for u in order:
            if before < bw[src] - lmax:
                pos[u] = line + before
                if before + vw(u) >= bw[src] - lmax:
                    end = before + vw(u)
            before += vw(u)
        # the block's share of the line: up to the end of the candidate that
        # crosses the excess, else the whole range's weight (jet:508-518)
        line += end if end is not None else leak_in + sum(vw(u) for u in mine if not leaked(u))
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


def two_heavy_blocks_case(seed):
    """Blocks 0 and 1 each hold a connected cluster of 12 vertices of weight
    60..75 (the same weights) and weigh opt + 39: no weak candidate (every
    weight is at least 1.5 (block weight - opt)), so the strong iteration runs
    first.  The edges inside block 0's cluster weigh 100000, so its vertices'
    gains fall in bucket 49; with one section (n < 4096 k) they sit in the
    last section, and the strong iteration's +1 minibucket offset (jet:437)
    puts them on the first minibucket of block 1's range (appendix A11):
    their weight comes before every candidate of block 1.  Blocks 2 and 3
    share the other vertices (weights 1..20) with room for any cluster
    vertex."""
    g0 = heipa.gen_rgg(240, 8.0, seed, 9)
    rng = np.random.default_rng(seed)

    def cluster(start, taken):
        out, queue, seen = [], [start], {start}
        while queue and len(out) < 12:
            u = queue.pop(0)
            if u in taken:
                continue
            out.append(u)
            for v in g0.adjncy[g0.xadj[u]:g0.xadj[u + 1]]:
                v = int(v)
                if v not in seen:
                    seen.add(v)
                    queue.append(v)
        return out

    A = cluster(int(rng.integers(g0.n)), set())
    B = cluster(next(int(u) for u in rng.permutation(g0.n) if int(u) not in A), set(A))
    assert len(A) == 12 and len(B) == 12
    heavy_w = rng.integers(60, 76, 12)
    vw = np.zeros(g0.n, np.int64)
    vw[A] = heavy_w
    vw[B] = heavy_w
    opt = int(heavy_w.sum()) - 39
    rest = [u for u in range(g0.n) if vw[u] == 0]
    vw[rest] = 1
    extra = 4 * opt - int(vw.sum())
    while extra > 0:  # raise the other vertices' weights, at most 20 each
        u = rest[int(rng.integers(len(rest)))]
        step = min(extra, 20 - int(vw[u]))
        vw[u] += step
        extra -= step
    assert int(vw.sum()) == 4 * opt
    inA = np.zeros(g0.n, bool)
    inA[A] = True
    src = np.repeat(np.arange(g0.n), np.diff(g0.xadj.astype(np.int64)))
    ew = g0.adjwgt.astype(np.int64).copy()
    ew[inA[src] & inA[g0.adjncy.astype(np.int64)]] = 100000
    g = heipa.HostGraph(g0.xadj, g0.adjncy, ew.astype(np.int32), vw.astype(np.int32))
    k = 4
    lmax = math.ceil(1.03 * int(g.g_weight) / k)
    part = np.zeros(g.n, np.uint32)
    part[B] = 1
    bw = [int(heavy_w.sum())] * 2 + [0, 0]
    for u in rest:
        b = 2 if bw[2] <= bw[3] else 3
        part[u] = b
        bw[b] += int(vw[u])
    max_b_w = max(opt + 1, int(lmax * 0.99))
    room = max(max_b_w - bw[2], max_b_w - bw[3])
    assert bw[0] > lmax and bw[1] > lmax and max(bw[2:]) < max_b_w
    assert all(1.5 * (bw[0] - opt) <= w < 2 * (bw[0] - opt) and w <= 2 * room for w in heavy_w)
    return g, part, k, lmax, bw
