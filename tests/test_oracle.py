"""CPU checks that pin the oracle: published hash values, closed forms,
exact-arithmetic references, brute-force serial simulation on tiny graphs,
and the invariants the reference asserts (src/utility/asserts.h:131-168,
solver.h:522-527)."""
import math
import os
import re
from fractions import Fraction

import numpy as np
import pytest

import heipa
import heipa_oracle as orc
import refsim
from graphs import block_weights, csr_from_edges, random_partition, star_plus_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- hashing
def test_murmur3_smhasher_verification_value():
    # SMHasher's VerificationTest for MurmurHash3_x86_32: hash keys 0..255 of
    # length i with seed 256-i, then hash the concatenated hashes with seed 0.
    key = bytes(range(256))
    for impl in (orc.murmur3, refsim.murmur3_32):
        hashes = b"".join(impl(key[:i], 256 - i).to_bytes(4, "little") for i in range(256))
        assert impl(hashes, 0) == 0xB0F57EE3


def test_murmur3_known_vectors():
    assert orc.murmur3(b"", 0) == 0
    assert orc.murmur3(b"", 1) == 0x514E28B7
    assert orc.murmur3(b"", 0xFFFFFFFF) == 0x81F16F39
    assert orc.murmur3(b"\x00\x00\x00\x00", 0) == 0x2362F9DE


def test_pod_hash_is_murmur_of_little_endian_u32():
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 2**32, 200, dtype=np.uint64):
        x = int(x)
        assert orc.pod_hash(x) == refsim.murmur3_32(x.to_bytes(4, "little"), 0) == refsim.pod_hash(x)


def test_xorshift32_marsaglia():
    # Marsaglia (2003), "Xorshift RNGs": y = 2463534242, shifts (13, 17, 5)
    assert orc.xorshift(2463534242) == 723471715
    assert refsim.xorshift(2463534242) == 723471715
    for x in (1, 0xDEADBEEF, 12345 ^ 7):
        assert orc.xorshift(x) == refsim.xorshift(x)


# ---------------------------------------------------------------- gain bucket
def _pairs():
    out = [(5, 1), (0, 1), (-1, 1), (-2, 3), (-3, 2), (-9, 4), (-27, 8), (-1, 1000), (-1000, 1),
           (-1, 3**20), (-(3**15), 2**15), (-(3**15) - 1, 2**15), (-(2**10), 3**10), (-(2**10) + 1, 3**10)]
    rng = np.random.default_rng(3)
    for _ in range(3000):
        out.append((-int(rng.integers(1, 10**6)), int(rng.integers(1, 10**4))))
    return out


def test_gain_bucket_exact_arithmetic():
    for gx, vw in _pairs():
        assert orc.gain_bucket(gx, vw) == refsim.gain_bucket_exact(gx, vw), (gx, vw)


def test_gain_bucket_matches_float_log_formula_away_from_boundaries():
    # jet_label_propagation.h:160-177 evaluated in float32 with a correctly
    # rounded log2; only quotients within 1e-4 (relative) of a power of 1.5
    # may differ, and none of the sampled ones are that close.
    inv = np.float32(1.0) / np.float32(0.5849625007211562)
    for gx, vw in _pairs():
        if gx >= 0:
            continue
        g = np.float32(-gx) / np.float32(vw)
        x = abs(math.log(float(g), 1.5))
        if abs(x - round(x)) < 1e-4:
            continue
        if g < 1:
            k = int(math.ceil(np.float32(-np.log2(g)) * inv - np.float32(1e-6)))
            b = max(25 - k, 2)
        else:
            k = int(math.ceil(np.float32(np.log2(g)) * inv - np.float32(1e-6)))
            b = min(25 + k, 49)
        assert orc.gain_bucket(gx, vw) == b, (gx, vw)


def test_device_gain_bucket_tables_are_exact():
    src = open(os.path.join(ROOT, "gpu-heipa_amd", "csrc", "device", "jet.hip")).read()
    lo = re.search(r"kLo\[24\] = \{([^}]*)\}", src).group(1)
    hi = re.search(r"kHi\[25\] = \{([^}]*)\}", src).group(1)
    lo = [float.fromhex(t.strip()[:-1]) for t in lo.split(",")]
    hi = [float.fromhex(t.strip()[:-1]) for t in hi.split(",")]
    ulp = lambda f: Fraction(2) ** (math.frexp(f)[1] - 24)  # noqa: E731
    for k, f in enumerate(lo):  # least float >= 1.5^-k
        t = Fraction(2**k, 3**k)
        assert Fraction(f) >= t and Fraction(f) - ulp(f) < t, k
    for k, f in enumerate(hi):  # greatest float <= 1.5^k
        t = Fraction(3**k, 2**k)
        assert Fraction(f) <= t and Fraction(f) + ulp(f) > t, k


# ---------------------------------------------------------------- evaluators
def test_distance_matrix_closed_form():
    for hier, dist in (([4, 8, 6], [1, 10, 100]), ([2, 4, 8], [1, 5, 50])):
        D = orc.distance_matrix(hier, dist)
        D2 = heipa.distance_matrix(hier, dist)
        k = int(np.prod(hier))
        g1, g2 = hier[0], hier[0] * hier[1]
        for i in range(k):
            for j in range(k):
                want = 0 if i == j else dist[0] if i // g1 == j // g1 else dist[1] if i // g2 == j // g2 else dist[2]
                assert D[i, j] == want == D2[i, j]


def test_edge_cut_grid_tiles_closed_form():
    # 512x512 grid cut into 4x8 tiles of 128x64: 3 horizontal + 7 vertical
    # boundaries of 512 edges each -> 5120
    g = heipa.gen_grid(512, 512)
    r, c = np.divmod(np.arange(g.n), 512)
    part = (r // 128) * 8 + (c // 64)
    assert orc.edge_cut(g, part) == 5120


def test_comm_cost_identity_tiles_closed_form():
    # 64x48 grid in 4x8x6 = 192 tiles of 4 rows x 4 cols laid out row-major
    # by tile id; J = 2 * sum over tile boundary edges of D[a][b]
    rows, cols = 64, 48
    g = heipa.gen_grid(rows, cols)
    r, c = np.divmod(np.arange(g.n), cols)
    tr, tc = r // 4, c // 4  # 16 x 12 tiles
    part = (tr * 12 + tc).astype(np.uint32)
    D = orc.distance_matrix([4, 8, 6], [1, 10, 100])
    J = 0
    for a in range(16):
        for b in range(12):
            t = a * 12 + b
            if b + 1 < 12:
                J += 4 * D[t, t + 1]
            if a + 1 < 16:
                J += 4 * D[t, t + 12]
    assert orc.comm_cost(g, part, D) == 2 * J


def test_evaluators_brute_force():
    for seed in range(5):
        g = heipa.gen_rgg(300, 7.0, seed, 10)
        part = random_partition(g.n, 5, seed)
        assert orc.edge_cut(g, part) == refsim.edge_cut(g, part)
        D = orc.distance_matrix([5], [3])
        brute = sum(int(g.adjwgt[j]) * int(D[part[u], part[int(g.adjncy[j])]])
                    for u in range(g.n) for j in range(int(g.xadj[u]), int(g.xadj[u + 1])))
        assert orc.comm_cost(g, part, D) == brute


def test_edge_cut_64bit():
    # weights near 2^31: the reference's s32 sum would overflow (edge_cut.h:75-89)
    g = csr_from_edges(3, [(0, 1), (1, 2)], [2**31 - 1, 2**31 - 1])
    assert orc.edge_cut(g, np.array([0, 1, 0], np.uint32)) == 2 * (2**31 - 1)


# ---------------------------------------------------------------- matching / contraction
TINY = [
    ("path", lambda: csr_from_edges(9, [(i, i + 1) for i in range(8)])),
    ("grid5", lambda: heipa.gen_grid(5, 6)),
    ("rgg", lambda: heipa.gen_rgg(80, 6.0, 3, 10)),
    ("rgg_u", lambda: heipa.gen_rgg(90, 5.0, 4, 0)),
    ("rmat", lambda: heipa.gen_rmat(7, 0, 4.0, 0.57, 0.19, 0.19, 5, 10, False)),
    ("rmat_u", lambda: heipa.gen_rmat(7, 0, 6.0, 0.57, 0.19, 0.19, 6, 0, False)),
    ("star", lambda: star_plus_grid(40, 3, 4)),
    ("isolated", lambda: csr_from_edges(12, [(0, 1), (2, 3)])),
    # coarse levels: merged edge weights and vertex weights
    ("rgg_l1", lambda: _coarsened(heipa.gen_rgg(160, 7.0, 8, 10), 1)),
    ("rmat_u_l2", lambda: _coarsened(heipa.gen_rmat(8, 0, 6.0, 0.57, 0.19, 0.19, 6, 0, False), 2)),
    ("grid_l2", lambda: _coarsened(heipa.gen_grid(12, 13), 2)),
]


def _coarsened(g, levels):
    for _ in range(levels):
        mp, nc, _ = orc.match(g)
        cx, ca, cw, cv, rc = orc.contract(g, mp, nc)
        assert rc == 0
        g = heipa.HostGraph(cx, ca, cw, cv)
    return g


@pytest.mark.parametrize("name,make", TINY)
def test_match_equals_serial_simulation(name, make):
    g = make()
    mp, nc, st = orc.match(g)
    smp, snc = refsim.match(g)
    assert nc == snc
    assert list(mp) == smp


@pytest.mark.parametrize("name,make", TINY)
def test_match_invariants(name, make):
    g = make()
    mp, nc, st = orc.match(g)
    assert set(mp.tolist()) == set(range(nc))  # onto [0, nc)
    sizes = np.bincount(mp, minlength=nc)
    assert sizes.max() <= 2
    first = {}
    for u in range(g.n):  # coarse ids follow representative (min id) order
        first.setdefault(int(mp[u]), u)
    assert list(first.keys()) == sorted(first.keys())


@pytest.mark.parametrize("name,make", TINY)
def test_contract_equals_brute_force(name, make):
    g = make()
    mp, nc, _ = orc.match(g)
    cx, ca, cw, cv, rc = orc.contract(g, mp, nc)
    assert rc == 0
    x, a, w, v = refsim.contract(g, [int(t) for t in mp], nc)
    assert list(cx) == x and list(ca) == a and list(cw) == w and list(cv) == v
    # the cut of any coarse partition equals the cut of its projection
    cp = random_partition(nc, 3, 1)
    coarse = heipa.HostGraph(cx, ca, cw, cv)
    assert orc.edge_cut(coarse, cp) == orc.edge_cut(g, orc.project(mp, cp))


def test_contract_overflow_is_reported():
    g = csr_from_edges(4, [(0, 2), (1, 2), (0, 3)], [2**31 - 1, 2**31 - 1, 1])
    mp = np.array([0, 0, 1, 2], np.uint32)  # edges (0,2) and (1,2) merge
    *_, rc = orc.contract(g, mp, 3)
    assert rc == 3


# ---------------------------------------------------------------- Jet / Solver / HM
@pytest.mark.parametrize("ultra", [False, True])
def test_jet_invariants(ultra):
    g = heipa.gen_rgg(3000, 10.0, 2, 10)
    k = 6
    lmax = math.ceil(1.03 * g.n / k)
    for skew in (0.0, 0.8):
        part0 = random_partition(g.n, k, 3, skew)
        cut0, maxw0 = orc.edge_cut(g, part0), int(block_weights(g, part0, k).max())
        p, cut, maxw, tr = orc.jet_refine(g, part0, k, lmax, ultra, 0, cut0, maxw0, trace_cap=10000)
        assert cut == orc.edge_cut(g, p)  # incremental cut == recount (solver.h:522-527)
        assert maxw == block_weights(g, p, k).max()
        assert maxw <= lmax and cut < cut0
        for t in tr:  # every traced iteration reports a consistent working cut
            assert t.n_moves > 0 and t.mode in (0, 1, 2)


@pytest.mark.parametrize("kind,k,ultra", [("grid", 8, False), ("rgg", 16, True), ("rmat", 5, False)])
def test_solver_invariants(kind, k, ultra):
    g = {"grid": lambda: heipa.gen_grid(60, 70), "rgg": lambda: heipa.gen_rgg(5000, 10.0, 1, 10),
         "rmat": lambda: heipa.gen_rmat(11, 0, 8.0, 0.57, 0.19, 0.19, 9, 20, True)}[kind]()
    p, cut, maxw, lv = orc.partition(g, k, 0.03, 1, ultra)
    lmax = math.ceil(1.03 * g.g_weight / k)
    assert p.max() < k
    assert cut == orc.edge_cut(g, p)
    assert maxw == block_weights(g, p, k).max() <= lmax
    assert lv[-1].n <= 8 * k and all(a.n > 8 * k for a in lv[:-1])  # solver.h:343
    p2, cut2, _, _ = orc.partition(g, k, 0.03, 1, ultra)  # deterministic
    assert np.array_equal(p, p2) and cut == cut2


@pytest.mark.parametrize("kind,k,ultra,imb", [("rgg", 4, False, 0.03), ("rmat", 6, True, 0.05),
                                               ("grid", 3, False, 0.01)])
def test_solver_level_loop_restated(kind, k, ultra, imb):
    """internal_solve (solver.h:340-400) restated over the oracle's pinned
    pieces: coarsen (two-hop matching with seed 12345, two_hop_matching.h:590)
    while n > 8k, the host partitioner on the coarsest graph (no refinement
    there), then per level from the coarsest-but-one down to 0: project the
    partition through the level's mapping and run Jet with that level's index
    and lmax = ceil((1 + imbalance) W / k) (solver.h:410), carrying the cut and
    the maximum block weight from level to level."""
    g = {"rgg": lambda: heipa.gen_rgg(4000, 9.0, 4, 10),
         "rmat": lambda: heipa.gen_rmat(11, 0, 6.0, 0.57, 0.19, 0.19, 3, 20, True),
         "grid": lambda: heipa.gen_grid(40, 45)}[kind]()
    seed = 5
    lmax = math.ceil((1.0 + imb) * (int(g.g_weight) / k))
    graphs, maps = [g], []
    while graphs[-1].n > 8 * k:
        mp, nc, _ = orc.match(graphs[-1], 12345)
        cx, ca, cw, cv, rc = orc.contract(graphs[-1], mp, nc)
        assert rc == 0
        maps.append(mp)
        graphs.append(heipa.HostGraph(cx, ca, cw, cv))
    assert len(maps) >= 2
    part = heipa.initial_partition(graphs[-1], k, imb, seed)
    cut, maxw = orc.edge_cut(graphs[-1], part), int(block_weights(graphs[-1], part, k).max())
    levels = {len(maps): (cut, maxw)}
    for lv in range(len(maps) - 1, -1, -1):
        part = part[maps[lv]]
        part, cut, maxw, _ = orc.jet_refine(graphs[lv], part, k, lmax, ultra, lv, cut, maxw)
        levels[lv] = (cut, maxw)
    p, c, mw, lt = orc.partition(g, k, imb, seed, ultra)
    assert np.array_equal(p, part) and (c, mw) == (cut, maxw)
    assert len(lt) == len(graphs)
    for lv, t in enumerate(lt):
        assert (t.n, t.m) == (graphs[lv].n, len(graphs[lv].adjncy))
        assert (t.cut, t.maxw) == levels[lv]
        if lv < len(maps):
            assert t.nc == graphs[lv + 1].n
            assert t.mapping_hash == sum(orc.mix64((u << 32) | int(maps[lv][u])) for u in range(graphs[lv].n)) & ((1 << 64) - 1)


def test_grid_quality_near_optimal():
    # 64x64 grid, k=8: a 2x4 tiling cuts 4*64 = 256 edges; the result must be close
    g = heipa.gen_grid(64, 64)
    _, cut, maxw, _ = orc.partition(g, 8, 0.03, 1, False)
    assert cut <= 1.3 * 256 and maxw <= math.ceil(1.03 * 512)


def test_hm_partition_structure():
    g = heipa.gen_rmat(10, 0, 8.0, 0.57, 0.19, 0.19, 4, 10, True)
    hier = [2, 3, 2]
    p = orc.hm_partition(g, hier, 0.03, 1, False)
    assert p.max() < 12
    # adaptive imbalance (global_multisection.h:39-48)
    e = orc.adaptive_imbalance(0.03, 1000, 12, 1000, 12, 3)
    assert abs(e - (1.03 ** (1 / 3) - 1)) < 1e-12
    assert abs(heipa.adaptive_imbalance(0.03, 1000, 12, 500, 6, 2) - ((1.03 * (6 * 1000) / (12 * 500)) ** 0.5 - 1)) < 1e-12
    # each top-level block (ids 0..5 and 6..11) is a union of leaves
    bw = block_weights(g, p, 12)
    W = g.g_weight
    assert bw.max() <= math.ceil(1.03 * W / 12) * 1.05
