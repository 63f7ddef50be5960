"""Seeded generators of the benchmark workloads (the synthetic inputs the
reference's headline configurations name): the batched form's 4096 small
graphs and the served form's request mix.  Each returns a list of
(HostGraph, spec, kind) with spec = {"k": k} for a partition request or
{"hier": hierarchy} for a process-mapping (HM) request."""
import math

import numpy as np

from . import gen_grid, gen_rgg, gen_rmat


def _generate(params, threads):
    """Calls gen(*args) for every (gen, args) in params, on `threads` host
    threads (the generators are C++ and release the interpreter lock)."""
    if threads <= 1:
        return [gen(*a) for gen, a in params]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda p: p[0](*p[1]), params))


def batched(count, first=0, threads=1):
    """Config 4: graph i uses seed 1000+i; RGG (avg degree 8) or R-MAT (edge
    factor 8, scale ceil(log2 n), ids >= n rejected) with probability 0.5
    each; n log-uniform in [1024, 65536]; edge weights 1 + hash mod 10; k
    from {2,4,8,16,32}; eps 0.03, default config.  `first` skips that many
    graphs of the sequence (the draws stay those of the full sequence)."""
    rng = np.random.default_rng(1)
    params, meta = [], []
    for i in range(first + count):
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(65536)))))
        rgg = rng.random() < 0.5
        k = int(rng.choice([2, 4, 8, 16, 32]))
        if i < first:
            continue
        if rgg:
            params.append((gen_rgg, (n, 8.0, 1000 + i, 10)))
        else:
            params.append((gen_rmat, (max(1, math.ceil(math.log2(n))), n, 8.0, 0.57, 0.19, 0.19, 1000 + i, 10,
                                      False)))
        meta.append((dict(k=k), "rgg" if rgg else "rmat"))
    graphs = _generate(params, threads)
    return [(g, spec, kind) for g, (spec, kind) in zip(graphs, meta)]


def served(count, nmax, threads=1):
    """Config 5 (request mix): 70% partition (k uniform in 8..256), 30%
    mapping (hierarchy 4:8:6 or 2:4:8); grid / RGG (avg degree 12) / R-MAT
    (edge factor 16) with probability 1/3 each; n log-uniform in
    [1e4, nmax]; edge weights 1 + hash mod 100."""
    rng = np.random.default_rng(1)
    params, meta = [], []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(1e4), math.log(nmax)))))
        t = int(rng.integers(3))
        if t == 0:
            side = max(2, int(round(math.sqrt(n))))
            params.append((gen_grid, (side, side)))
            kind = "grid"
        elif t == 1:
            params.append((gen_rgg, (n, 12.0, 2000 + i, 100)))
            kind = "rgg"
        else:
            params.append((gen_rmat, (max(1, math.ceil(math.log2(n))), n, 16.0, 0.57, 0.19, 0.19, 2000 + i, 100,
                                      False)))
            kind = "rmat"
        if rng.random() < 0.7:
            spec = dict(k=int(rng.integers(8, 257)))
        else:
            spec = dict(hier=[4, 8, 6] if rng.random() < 0.5 else [2, 4, 8])
        meta.append((spec, kind))
    graphs = _generate(params, threads)
    return [(g, spec, kind) for g, (spec, kind) in zip(graphs, meta)]


def poisson_arrivals(count, rate, seed=1):
    """Arrival times (s) of a Poisson stream with `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))
