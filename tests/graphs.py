"""Test graph builders shared by the CPU and GPU suites."""
import numpy as np

import heipa


def csr_from_edges(n, edges, weights=None, vwgt=None):
    """Symmetric CSR (sorted rows, deduplicated) from undirected (u, v) pairs."""
    adj = [dict() for _ in range(n)]
    for i, (u, v) in enumerate(edges):
        if u == v:
            continue
        w = 1 if weights is None else int(weights[i])
        adj[u][v] = w
        adj[v][u] = w
    xadj = np.zeros(n + 1, np.uint32)
    cols, ws = [], []
    for u in range(n):
        items = sorted(adj[u].items())
        cols += [v for v, _ in items]
        ws += [w for _, w in items]
        xadj[u + 1] = len(cols)
    return heipa.HostGraph(xadj, np.array(cols, np.uint32),
                           None if weights is None else np.array(ws, np.int32),
                           None if vwgt is None else np.asarray(vwgt, np.int32))


def star_plus_grid(leaves=3000, rows=20, cols=20, seed=3):
    """A hub with `leaves` degree-1 neighbours attached to a grid; exercises hub
    kernels and the leaf pass."""
    rng = np.random.default_rng(seed)
    n = 1 + leaves + rows * cols
    edges, ws = [], []
    for i in range(leaves):
        edges.append((0, 1 + i))
        ws.append(int(rng.integers(1, 10)))
    base = 1 + leaves
    for r in range(rows):
        for c in range(cols):
            v = base + r * cols + c
            if c + 1 < cols:
                edges.append((v, v + 1))
                ws.append(int(rng.integers(1, 10)))
            if r + 1 < rows:
                edges.append((v, v + cols))
                ws.append(int(rng.integers(1, 10)))
    edges.append((0, base))
    ws.append(5)
    return csr_from_edges(n, edges, ws)


def random_partition(n, k, seed, skew=0.0):
    rng = np.random.default_rng(seed)
    if skew <= 0:
        return rng.integers(0, k, n).astype(np.uint32)
    p = np.ones(k)
    p[0] += skew * k
    p /= p.sum()
    return rng.choice(k, n, p=p).astype(np.uint32)


def block_weights(g, part, k):
    vw = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    return np.bincount(part, weights=vw, minlength=k).astype(np.int64)


def csr_from_arrays(n, us, vs, ws=None):
    """Symmetric CSR (sorted rows, no self loops) from undirected edge arrays;
    a pair given more than once keeps its first weight, in both directions."""
    us, vs = np.asarray(us, np.int64), np.asarray(vs, np.int64)
    keep = us != vs
    us, vs = us[keep], vs[keep]
    lo, hi = np.minimum(us, vs), np.maximum(us, vs)
    _, first = np.unique(lo * n + hi, return_index=True)
    lo, hi = lo[first], hi[first]
    w = None if ws is None else np.asarray(ws, np.int64)[keep][first]
    a, b = np.concatenate([lo, hi]), np.concatenate([hi, lo])
    order = np.argsort(a * n + b, kind="stable")
    a, b = a[order], b[order]
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, a + 1, 1)
    xadj = np.cumsum(xadj).astype(np.uint32)
    ew = None if w is None else np.concatenate([w, w])[order].astype(np.int32)
    return heipa.HostGraph(xadj, b.astype(np.uint32), ew, None)


def many_hubs(n=60000, hubs=100, hub_degree=3000, seed=5):
    """A ring of n vertices plus `hubs` vertices of degree about hub_degree to
    random vertices: many workgroup-per-hub vertices at once."""
    rng = np.random.default_rng(seed)
    ring = np.arange(n)
    us = [ring, np.repeat(np.arange(hubs), hub_degree)]
    vs = [(ring + 1) % n, rng.integers(0, n, hubs * hub_degree)]
    u, v = np.concatenate(us), np.concatenate(vs)
    return csr_from_arrays(n, u, v, rng.integers(1, 10, len(u)))
