"""CPU tests of the product library's host side: the C-ABI loads and exports
what include/heipa.h declares, generators, METIS I/O, the partition writer,
the shared initial partitioner, and error codes without a GPU."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import heipa
from graphs import csr_from_edges

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = heipa.lib()
    names = heipa.exported_symbols_from_header()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", heipa.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_binding_constants_match_header():
    """Every status code, option and job kind that include/heipa.h defines
    has the same value in the Python binding."""
    import re
    text = open(os.path.join(ROOT, "include", "heipa.h")).read()
    defs = dict(re.findall(r"^#define HEIPA_((?:OK|ERR_|OPT_|JOB_)\w*)\s+(\d+)", text, re.M))
    assert "OPT_REGION_EDGES" in defs and "ERR_IO" in defs and "JOB_HM" in defs
    for name, value in defs.items():
        assert getattr(heipa, name) == int(value), name


def test_null_context_is_an_error_not_a_crash():
    L = heipa.lib()
    assert L.heipa_edge_cut(None, None, None, None) == heipa.ERR_INVALID
    assert L.heipa_ctx_destroy(None) == heipa.ERR_INVALID
    assert L.heipa_ctx_last_error(None) == b"null context"
    assert L.heipa_distance_matrix(None, None, 0, None) == heipa.ERR_INVALID


def test_grid_generator_closed_form():
    g = heipa.gen_grid(7, 5)
    assert g.n == 35 and g.m == 2 * (6 * 5 + 7 * 4)
    assert g.adjwgt is None and g.vwgt is None
    for u in range(g.n):
        r, c = divmod(u, 5)
        want = sorted(v for v, ok in ((u - 5, r > 0), (u - 1, c > 0), (u + 1, c < 4), (u + 5, r < 6)) if ok)
        assert list(g.adjncy[g.xadj[u]:g.xadj[u + 1]]) == want


def _xoshiro_points(n, seed):
    """xoshiro256** seeded by splitmix64, doubles (x >> 11) * 2^-53."""
    M = (1 << 64) - 1
    st = seed

    def sm():
        nonlocal st
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    s = [sm() for _ in range(4)]
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M  # noqa: E731

    def nxt():
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return r

    pts = []
    for _ in range(n):
        x = (nxt() >> 11) * 2.0**-53
        y = (nxt() >> 11) * 2.0**-53
        pts.append((x, y))
    return pts


def test_rgg_generator_brute_force():
    n, deg, seed = 400, 8.0, 11
    g = heipa.gen_rgg(n, deg, seed, 0)
    pts = _xoshiro_points(n, seed)
    r2 = (math.sqrt(deg / (math.pi * n))) ** 2
    for u in range(n):
        want = [v for v in range(n) if v != u and (pts[u][0] - pts[v][0]) ** 2 + (pts[u][1] - pts[v][1]) ** 2 <= r2]
        assert list(g.adjncy[g.xadj[u]:g.xadj[u + 1]]) == want


def test_rmat_generator_properties():
    g = heipa.gen_rmat(10, 0, 8.0, 0.57, 0.19, 0.19, 5, 100, True)
    assert g.n == 1024 and g.vwgt is not None and (g.vwgt == 1).all()
    assert g.adjwgt.min() >= 1 and g.adjwgt.max() <= 100
    seen = set()
    for u in range(g.n):
        row = g.adjncy[g.xadj[u]:g.xadj[u + 1]]
        assert (np.diff(row.astype(np.int64)) > 0).all()  # sorted, no duplicates
        assert u not in row  # no self loops
        for j, v in enumerate(row):
            seen.add((u, int(v), int(g.adjwgt[g.xadj[u] + j])))
    assert all((v, u, w) in seen for (u, v, w) in seen)  # symmetric with equal weights
    deg = np.diff(g.xadj.astype(np.int64))
    assert deg[0] == deg.max()  # a=0.57 concentrates edges on low ids


def test_metis_round_trip(tmp_path):
    for g in (heipa.gen_grid(6, 9), heipa.gen_rgg(200, 6.0, 2, 10),
              heipa.gen_rmat(8, 0, 4.0, 0.57, 0.19, 0.19, 1, 10, True)):
        p = str(tmp_path / "g.graph")
        heipa.write_metis(p, g)
        h = heipa.read_metis(p)
        assert np.array_equal(h.xadj, g.xadj) and np.array_equal(h.adjncy, g.adjncy)
        assert (h.adjwgt is None) == (g.adjwgt is None) and (h.vwgt is None) == (g.vwgt is None)
        if g.adjwgt is not None:
            assert np.array_equal(h.adjwgt, g.adjwgt)
        if g.vwgt is not None:
            assert np.array_equal(h.vwgt, g.vwgt)


@pytest.mark.parametrize("fmt,body,ew,vw", [
    ("", "2 3\n1\n1\n", None, None),
    ("001", "2 5 3 7\n1 5\n1 7\n", [5, 7, 5, 7], None),
    ("010", "4 2 3\n6 1\n9 1\n", None, [4, 6, 9]),
    ("011", "4 2 5 3 7\n6 1 5\n9 1 7\n", [5, 7, 5, 7], [4, 6, 9]),
    ("1", "2 5 3 7\n1 5\n1 7\n", [5, 7, 5, 7], None),
    ("11", "4 2 5 3 7\n6 1 5\n9 1 7\n", [5, 7, 5, 7], [4, 6, 9]),
])
def test_metis_reader_formats(tmp_path, fmt, body, ew, vw):
    # path 2-1-3 (3 vertices, 2 undirected edges) with comment lines
    p = tmp_path / "f.graph"
    p.write_text(f"% comment\n3 2 {fmt}\n".replace(" \n", "\n") + "% inner comment\n" + body)
    g = heipa.read_metis(str(p))
    assert list(g.xadj) == [0, 2, 3, 4]
    assert list(g.adjncy) == [1, 2, 0, 0]
    assert (None if g.adjwgt is None else list(g.adjwgt)) == ew
    assert (None if g.vwgt is None else list(g.vwgt)) == vw


def test_metis_reader_errors(tmp_path):
    L = heipa.lib()
    h = ctypes.c_void_p()
    assert L.heipa_read_metis(str(tmp_path / "missing.graph").encode(), ctypes.byref(h)) == heipa.ERR_IO
    bad = tmp_path / "bad.graph"
    bad.write_text("3 5\n2\n1\n\n")  # header announces more edges than present
    assert L.heipa_read_metis(str(bad).encode(), ctypes.byref(h)) == heipa.ERR_INVALID
    bad.write_text("2 1\n3\n1\n")  # neighbour id out of range
    assert L.heipa_read_metis(str(bad).encode(), ctypes.byref(h)) == heipa.ERR_INVALID


def test_partition_writer_format(tmp_path):
    p = str(tmp_path / "part.txt")
    heipa.write_partition(p, np.array([3, 0, 12, 7], np.uint32))
    assert open(p).read() == "3\n0\n12\n7\n"


@pytest.mark.parametrize("k", [1, 2, 5, 8, 64])
def test_initial_partitioner(k):
    g = heipa.gen_rgg(2000, 8.0, 3, 10)
    p = heipa.initial_partition(g, k, 0.03, 7)
    assert p.max() < k
    if k > 1:
        lmax = math.ceil(1.03 * g.n / k)
        assert np.bincount(p, minlength=k).max() <= lmax
    assert np.array_equal(p, heipa.initial_partition(g, k, 0.03, 7))


def test_initial_partitioner_weighted_and_disconnected():
    g = csr_from_edges(10, [(0, 1), (2, 3), (4, 5), (6, 7)], [3, 1, 4, 1], [5, 1, 1, 1, 1, 1, 1, 1, 1, 1])
    p = heipa.initial_partition(g, 2, 0.1, 1)
    w = np.bincount(p, weights=g.vwgt, minlength=2)
    assert w.max() <= math.ceil(1.1 * 14 / 2)


def test_job_layout_matches_header(tmp_path):
    """The ctypes mirror of heipa_job has the C layout (offsets from gcc)."""
    fields = [f for f, _ in heipa.Job._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "heipa.h"\nint main(void) {\n'
                   + "".join(f'  printf("{f} %zu\\n", offsetof(heipa_job, {f}));\n' for f in fields)
                   + '  printf("sizeof %zu\\n", sizeof(heipa_job));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for f in fields:
        assert int(out[f]) == getattr(heipa.Job, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(heipa.Job)


def test_job_estimate_grows_with_size():
    L = heipa.lib()
    def est(n, m, k, hier=None):
        j = heipa.Job()
        j.graph.n, j.graph.m = n, m
        if hier:
            j.kind, j.l = heipa.JOB_HM, len(hier)
            for i, a in enumerate(hier):
                j.hier[i] = a
        else:
            j.kind, j.k = heipa.JOB_PARTITION, k
        return L.heipa_job_estimate(ctypes.byref(j))
    assert est(1000, 8000, 8) < est(2000, 8000, 8) < est(2000, 16000, 8)
    assert est(1 << 20, 12 << 20, 64) > 56 * (12 << 20)
    assert est(1 << 16, 1 << 20, 8, [4, 8, 6]) > est(1 << 16, 1 << 20, 8)
