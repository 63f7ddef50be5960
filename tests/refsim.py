"""Pure-Python serial simulation of the reference's coarsening and evaluators
for tiny graphs (brute force, independent of the C++ oracle).

Serial execution order equals the deterministic tie-break rules wherever the
reference resolves a race by atomic order: proposers are visited in list
(= id) order, so the first CAS on a receiver is by the smallest proposer id;
hash pairing visits vertices in id order.  Citations are into the reference.
"""
from fractions import Fraction

M32 = 0xFFFFFFFF
SENT = M32


def rotl32(x, r):
    return ((x << r) | (x >> (32 - r))) & M32


def fmix32(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    h ^= h >> 16
    return h


def murmur3_32(data: bytes, seed: int) -> int:
    """MurmurHash3_x86_32 (SMHasher, public domain algorithm)."""
    h = seed & M32
    n = len(data) // 4
    for i in range(n):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * 0xCC9E2D51) & M32
        k = rotl32(k, 15)
        k = (k * 0x1B873593) & M32
        h ^= k
        h = rotl32(h, 13)
        h = (h * 5 + 0xE6546B64) & M32
    tail = data[4 * n:]
    k = 0
    if len(tail) >= 3:
        k ^= tail[2] << 16
    if len(tail) >= 2:
        k ^= tail[1] << 8
    if len(tail) >= 1:
        k ^= tail[0]
        k = (k * 0xCC9E2D51) & M32
        k = rotl32(k, 15)
        k = (k * 0x1B873593) & M32
        h ^= k
    h ^= len(data)
    return fmix32(h)


def pod_hash(x):
    return murmur3_32((x & M32).to_bytes(4, "little"), 0)


def xorshift(x):
    x &= M32
    x ^= (x << 13) & M32
    x ^= x >> 17
    x ^= (x << 5) & M32
    return x


def edge_cut(g, part):
    s = 0
    for u in range(g.n):
        for j in range(int(g.xadj[u]), int(g.xadj[u + 1])):
            if part[u] != part[int(g.adjncy[j])]:
                s += 1 if g.adjwgt is None else int(g.adjwgt[j])
    return s // 2


def gain_bucket_exact(gx, vw):
    """jet_label_propagation.h:147-180 with exact arithmetic on the float quotient."""
    import numpy as np
    if gx > 0:
        return 0
    if gx == 0:
        return 1
    g = Fraction(float(np.float32(-gx) / np.float32(vw)))
    if g < 1:
        k = 1
        while k <= 23 and g * Fraction(3, 2) ** k < 1:
            k += 1
        return max(25 - k, 2)
    k = 0
    while k <= 24 and Fraction(3, 2) ** k < g:
        k += 1
    return min(25 + k, 49)


def match(g, seed=12345):
    """two_hop_matcher_get_mapping (two_hop_matching.h:573-650), serial."""
    n = g.n
    xadj, adj = [int(x) for x in g.xadj], [int(x) for x in g.adjncy]
    ew = None if g.adjwgt is None else [int(x) for x in g.adjwgt]
    vcmap = [SENT] * n
    hn = [SENT] * n

    def pick(u, s, initial):  # pick_neighbor_flat :102-141
        r = xorshift(u ^ s)
        h, max_ewt, tb = SENT, 0, 0
        for j in range(xadj[u], xadj[u + 1]):
            v = adj[j]
            if not initial and vcmap[v] != SENT:
                continue
            if ew is not None:
                if max_ewt < ew[j]:
                    max_ewt, h, tb = ew[j], v, xorshift((v + r) & M32)
                    continue
                if max_ewt != ew[j]:
                    continue
            t = xorshift((v + r) & M32)
            if t >= tb:
                h, tb = v, t
        return h

    for u in range(n):  # :278-297
        if ew is None:
            d = xadj[u + 1] - xadj[u]
            if d:
                hn[u] = adj[xadj[u] + xorshift(u ^ seed) % d]
        else:
            hn[u] = pick(u, seed, True)
    lst = list(range(n))
    rnd = 0
    while lst:
        rs = seed ^ ((rnd * 0x9E3779B1) & M32)
        for r in range(4):
            for u in lst:  # p1
                v = hn[u]
                if v == SENT or vcmap[u] != SENT:
                    continue
                cond = pod_hash(u + r) < pod_hash(v + r) if r > 0 else u < v
                if not cond:
                    vcmap[u] = SENT - 1
            for u in lst:  # p2, serial CAS
                v = hn[u]
                if v == SENT or vcmap[u] != SENT:
                    continue
                if vcmap[v] == SENT - 1:
                    vcmap[v] = vcmap[u] = min(u, v)
            for u in lst:  # p3
                if vcmap[u] == SENT - 1:
                    vcmap[u] = SENT
        for u in lst:  # repick
            if vcmap[u] != SENT or hn[u] == SENT or vcmap[hn[u]] == SENT:
                continue
            hn[u] = pick(u, rs, False)
        lst = [u for u in lst if vcmap[u] == SENT and hn[u] != SENT]
        rnd += 1

    def pair(items):  # matchHash, serial: twin slot holds the last unpaired vertex
        slot = {}
        for key, u in items:
            if key in slot:
                t = slot.pop(key)
                vcmap[t] = vcmap[u] = min(t, u)
            else:
                slot[key] = u

    def unmapped():
        return sum(1 for x in vcmap if x == SENT)

    if unmapped() / n > 0.25:  # leaf
        pair([(adj[xadj[u]], u) for u in range(n) if vcmap[u] == SENT and xadj[u + 1] - xadj[u] == 1])
    if unmapped() / n > 0.25:  # twin
        items = []
        for u in range(n):
            if vcmap[u] != SENT:
                continue
            h = 0
            for j in range(xadj[u], xadj[u + 1]):
                y = pod_hash(adj[j])
                h = (h + y * y + y) % (1 << 64)
            items.append((h, u))
        pair(items)
    if unmapped() / n > 0.25:  # relative
        items = []
        for u in range(n):
            if vcmap[u] != SENT:
                continue
            h, max_wgt, min_deg = SENT, 0, SENT
            for j in range(xadj[u], xadj[u + 1]):
                v = adj[j]
                vd = xadj[v + 1] - xadj[v]
                w = 1 if ew is None else ew[j]
                if min_deg > vd:
                    min_deg, max_wgt, h = vd, w, v
                elif min_deg == vd and max_wgt < w:
                    h, max_wgt = v, w
            items.append((h, u))
        pair(items)
    for u in range(n):
        if vcmap[u] == SENT:
            vcmap[u] = u
    cid, nc = {}, 0
    for u in range(n):
        if vcmap[u] == u:
            cid[u] = nc
            nc += 1
    return [cid[vcmap[u]] for u in range(n)], nc


def contract(g, mapping, nc):
    """from_Graph_Mapping (graph.h:127-390) as a multiset aggregation, rows sorted."""
    rows = [dict() for _ in range(nc)]
    cw = [0] * nc
    for u in range(g.n):
        c = mapping[u]
        cw[c] += 1 if g.vwgt is None else int(g.vwgt[u])
        for j in range(int(g.xadj[u]), int(g.xadj[u + 1])):
            x = mapping[int(g.adjncy[j])]
            if x != c:
                rows[c][x] = rows[c].get(x, 0) + (1 if g.adjwgt is None else int(g.adjwgt[j]))
    xadj, adj, w = [0], [], []
    for c in range(nc):
        for x in sorted(rows[c]):
            adj.append(x)
            w.append(rows[c][x])
        xadj.append(len(adj))
    return xadj, adj, w, cw
