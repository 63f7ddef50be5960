"""ctypes binding of the CPU oracle (liboracle) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this,
as the checker.  All arrays are numpy host arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libheipa_oracle.so")

u32p = ctypes.POINTER(ctypes.c_uint32)
s32p = ctypes.POINTER(ctypes.c_int32)
s64p = ctypes.POINTER(ctypes.c_int64)


class OrcGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("xadj", u32p), ("adjncy", u32p),
                ("adjwgt", s32p), ("vwgt", s32p), ("g_weight", ctypes.c_int64)]


class OrcJetTrace(ctypes.Structure):
    _fields_ = [("ratio_idx", ctypes.c_uint32), ("iteration", ctypes.c_uint32), ("mode", ctypes.c_uint32),
                ("n_cand", ctypes.c_uint32), ("n_moves", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("cut", ctypes.c_int64), ("maxw", ctypes.c_int64),
                ("moves_hash", ctypes.c_uint64), ("gain_hash", ctypes.c_uint64)]


class OrcLevelTrace(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("nc", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("mapping_hash", ctypes.c_uint64), ("cut", ctypes.c_int64), ("maxw", ctypes.c_int64)]


class OrcMatchStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in
                ("rounds", "after_hem", "after_leaf", "after_twin", "after_relative",
                 "ran_leaf", "ran_twin", "ran_relative")]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        G = ctypes.POINTER(OrcGraph)
        c_u32, c_u64, c_int, c_dbl = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        sig = {
            "orc_set_row_mode": (None, [c_int]),
            "orc_set_row_check": (None, [c_int]),
            "orc_row_error": (c_int, []),
            "orc_murmur3_x86_32": (c_u32, [ctypes.c_void_p, ctypes.c_int32, c_u32]),
            "orc_pod_hash_u32": (c_u32, [c_u32]),
            "orc_xorshift": (c_u32, [c_u32]),
            "orc_mix64": (c_u64, [c_u64]),
            "orc_gain_bucket": (c_u32, [ctypes.c_int64, ctypes.c_int64]),
            "orc_edge_cut": (ctypes.c_int64, [G, u32p]),
            "orc_comm_cost": (ctypes.c_int64, [G, u32p, s32p, c_u32]),
            "orc_distance_matrix": (c_int, [u32p, s32p, c_u32, s32p]),
            "orc_match": (c_int, [G, c_u32, u32p, u32p, ctypes.POINTER(OrcMatchStats)]),
            "orc_contract": (c_int, [G, u32p, c_u32, u32p, u32p, s32p, s32p, u32p]),
            "orc_project": (None, [u32p, c_u32, u32p, u32p]),
            "orc_jet_refine": (c_int, [G, u32p, c_u32, ctypes.c_int64, c_int, c_u32, s64p, s64p,
                                       ctypes.POINTER(OrcJetTrace), c_u32, u32p]),
            "orc_partition": (c_int, [G, c_u32, c_dbl, c_u64, c_int, u32p, s64p, s64p,
                                      ctypes.POINTER(OrcLevelTrace), c_u32, u32p]),
            "orc_hm_partition": (c_int, [G, u32p, c_u32, c_dbl, c_u64, c_int, u32p]),
            "orc_adaptive_imbalance": (c_dbl, [c_dbl, ctypes.c_int64, c_u32, ctypes.c_int64, c_u32, c_u64]),
            "orc_map_refine": (c_int, [G, u32p, s32p, c_u32, ctypes.c_int64, s64p, s64p,
                                       ctypes.POINTER(OrcJetTrace), c_u32, u32p]),
            "orc_global_multisection": (c_int, [G, u32p, c_u32, c_dbl, c_u64, u32p]),
            "orc_im_partition": (c_int, [G, u32p, s32p, c_u32, c_dbl, c_u64, u32p, s64p, s64p,
                                         ctypes.POINTER(OrcLevelTrace), c_u32, u32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class _G:
    """Keeps numpy arrays alive for an OrcGraph view."""

    def __init__(self, g):
        self.xadj = np.ascontiguousarray(g.xadj, np.uint32)
        self.adjncy = np.ascontiguousarray(g.adjncy, np.uint32)
        if len(self.adjncy) == 0:
            self.adjncy = np.zeros(1, np.uint32)
        self.adjwgt = None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, np.int32)
        if self.adjwgt is not None and len(self.adjwgt) == 0:
            self.adjwgt = np.zeros(1, np.int32)
        self.vwgt = None if g.vwgt is None else np.ascontiguousarray(g.vwgt, np.int32)
        n = len(self.xadj) - 1
        gw = int(self.vwgt[:n].astype(np.int64).sum()) if self.vwgt is not None else n
        self.s = OrcGraph(n, int(self.xadj[-1]), _p(self.xadj, ctypes.c_uint32), _p(self.adjncy, ctypes.c_uint32),
                          None if self.adjwgt is None else _p(self.adjwgt, ctypes.c_int32),
                          None if self.vwgt is None else _p(self.vwgt, ctypes.c_int32), gw)

    @property
    def ref(self):
        return ctypes.byref(self.s)


def set_row_mode(mode):
    """1 (default): block connectivity rows as the reference keeps them
    (block_conn.h), entries of weight 0 included; 0: positive connections only."""
    lib().orc_set_row_mode(int(mode))


def set_row_check(on):
    """Compare every row with the adjacency after each update (slow)."""
    lib().orc_set_row_check(int(on))


def row_error():
    return lib().orc_row_error()


def murmur3(data: bytes, seed: int) -> int:
    buf = ctypes.create_string_buffer(data, len(data))
    return lib().orc_murmur3_x86_32(buf, len(data), seed)


def pod_hash(x):
    return lib().orc_pod_hash_u32(x)


def xorshift(x):
    return lib().orc_xorshift(x)


def mix64(x):
    return lib().orc_mix64(x)


def gain_bucket(gx, vw):
    return lib().orc_gain_bucket(gx, vw)


def edge_cut(g, part):
    G = _G(g)
    p = np.ascontiguousarray(part, np.uint32)
    return lib().orc_edge_cut(G.ref, _p(p, ctypes.c_uint32))


def comm_cost(g, part, D):
    G = _G(g)
    p = np.ascontiguousarray(part, np.uint32)
    D = np.ascontiguousarray(D, np.int32)
    k = int(round(np.sqrt(D.size)))
    return lib().orc_comm_cost(G.ref, _p(p, ctypes.c_uint32), _p(D.reshape(-1), ctypes.c_int32), k)


def distance_matrix(hier, dist):
    hier = np.ascontiguousarray(hier, np.uint32)
    dist = np.ascontiguousarray(dist, np.int32)
    k = int(np.prod(hier.astype(np.int64)))
    D = np.zeros(k * k, np.int32)
    lib().orc_distance_matrix(_p(hier, ctypes.c_uint32), _p(dist, ctypes.c_int32), len(hier), _p(D, ctypes.c_int32))
    return D.reshape(k, k)


def match(g, seed=12345):
    G = _G(g)
    n = len(G.xadj) - 1
    mp = np.zeros(max(n, 1), np.uint32)
    nc = ctypes.c_uint32()
    st = OrcMatchStats()
    lib().orc_match(G.ref, seed, _p(mp, ctypes.c_uint32), ctypes.byref(nc), ctypes.byref(st))
    return mp[:n], nc.value, st


def contract(g, mapping, nc):
    """Returns (xadj, adjncy, adjwgt, vwgt, status)."""
    G = _G(g)
    m = int(G.xadj[-1])
    mp = np.ascontiguousarray(mapping, np.uint32)
    cx = np.zeros(nc + 1, np.uint32)
    ca = np.zeros(max(m, 1), np.uint32)
    cw = np.zeros(max(m, 1), np.int32)
    cv = np.zeros(max(nc, 1), np.int32)
    cm = ctypes.c_uint32()
    rc = lib().orc_contract(G.ref, _p(mp, ctypes.c_uint32), nc, _p(cx, ctypes.c_uint32), _p(ca, ctypes.c_uint32),
                            _p(cw, ctypes.c_int32), _p(cv, ctypes.c_int32), ctypes.byref(cm))
    return cx, ca[:cm.value].copy(), cw[:cm.value].copy(), cv[:nc].copy(), rc


def project(mapping, coarse_part):
    mp = np.ascontiguousarray(mapping, np.uint32)
    cp = np.ascontiguousarray(coarse_part, np.uint32)
    out = np.zeros(len(mp), np.uint32)
    lib().orc_project(_p(mp, ctypes.c_uint32), len(mp), _p(cp, ctypes.c_uint32), _p(out, ctypes.c_uint32))
    return out


def jet_refine(g, part, k, lmax, ultra, level, cut, maxw, trace_cap=0):
    G = _G(g)
    p = np.ascontiguousarray(part, np.uint32).copy()
    c, mw = ctypes.c_int64(cut), ctypes.c_int64(maxw)
    tr = (OrcJetTrace * trace_cap)() if trace_cap else None
    tl = ctypes.c_uint32()
    lib().orc_jet_refine(G.ref, _p(p, ctypes.c_uint32), k, lmax, int(ultra), level, ctypes.byref(c), ctypes.byref(mw),
                         tr, trace_cap, ctypes.byref(tl))
    trace = [tr[i] for i in range(min(tl.value, trace_cap))] if trace_cap else None
    return p, c.value, mw.value, trace


def partition(g, k, imbalance=0.03, seed=1, ultra=False, level_cap=64):
    G = _G(g)
    n = len(G.xadj) - 1
    p = np.zeros(max(n, 1), np.uint32)
    c, mw = ctypes.c_int64(), ctypes.c_int64()
    lv = (OrcLevelTrace * level_cap)()
    nl = ctypes.c_uint32()
    rc = lib().orc_partition(G.ref, k, imbalance, seed, int(ultra), _p(p, ctypes.c_uint32), ctypes.byref(c),
                             ctypes.byref(mw), lv, level_cap, ctypes.byref(nl))
    if rc != 0:
        raise RuntimeError(f"orc_partition failed: {rc}")
    return p[:n], c.value, mw.value, [lv[i] for i in range(min(nl.value, level_cap))]


def hm_partition(g, hier, imbalance=0.03, seed=1, ultra=False):
    G = _G(g)
    n = len(G.xadj) - 1
    p = np.zeros(max(n, 1), np.uint32)
    h = np.ascontiguousarray(hier, np.uint32)
    rc = lib().orc_hm_partition(G.ref, _p(h, ctypes.c_uint32), len(h), imbalance, seed, int(ultra),
                                _p(p, ctypes.c_uint32))
    if rc != 0:
        raise RuntimeError(f"orc_hm_partition failed: {rc}")
    return p[:n]


def adaptive_imbalance(eps, gW, gk, lW, krem, depth):
    return lib().orc_adaptive_imbalance(eps, gW, gk, lW, krem, depth)


def map_refine(g, part, D, lmax, J, maxw, trace_cap=0):
    """ProMap_jet_refine: returns (part, J, maxw, trace); trace cut fields hold J."""
    G = _G(g)
    p = np.ascontiguousarray(part, np.uint32).copy()
    D = np.ascontiguousarray(D, np.int32)
    k = int(round(np.sqrt(D.size)))
    j, mw = ctypes.c_int64(J), ctypes.c_int64(maxw)
    tr = (OrcJetTrace * trace_cap)() if trace_cap else None
    tl = ctypes.c_uint32()
    lib().orc_map_refine(G.ref, _p(p, ctypes.c_uint32), _p(D.reshape(-1), ctypes.c_int32), k, lmax, ctypes.byref(j),
                         ctypes.byref(mw), tr, trace_cap, ctypes.byref(tl))
    trace = [tr[i] for i in range(min(tl.value, trace_cap))] if trace_cap else None
    return p, j.value, mw.value, trace


def global_multisection(g, hier, imbalance=0.03, seed=1):
    G = _G(g)
    n = len(G.xadj) - 1
    p = np.zeros(max(n, 1), np.uint32)
    h = np.ascontiguousarray(hier, np.uint32)
    rc = lib().orc_global_multisection(G.ref, _p(h, ctypes.c_uint32), len(h), imbalance, seed,
                                       _p(p, ctypes.c_uint32))
    if rc != 0:
        raise RuntimeError(f"orc_global_multisection failed: {rc}")
    return p[:n]


def im_partition(g, hier, dist, imbalance=0.03, seed=1, level_cap=64):
    """ProMapSolver::solve (--config IM): returns (part, J, maxw, levels)."""
    G = _G(g)
    n = len(G.xadj) - 1
    p = np.zeros(max(n, 1), np.uint32)
    h = np.ascontiguousarray(hier, np.uint32)
    d = np.ascontiguousarray(dist, np.int32)
    j, mw = ctypes.c_int64(), ctypes.c_int64()
    lv = (OrcLevelTrace * level_cap)()
    nl = ctypes.c_uint32()
    rc = lib().orc_im_partition(G.ref, _p(h, ctypes.c_uint32), _p(d, ctypes.c_int32), len(h), imbalance, seed,
                                _p(p, ctypes.c_uint32), ctypes.byref(j), ctypes.byref(mw), lv, level_cap,
                                ctypes.byref(nl))
    if rc != 0:
        raise RuntimeError(f"orc_im_partition failed: {rc}")
    return p[:n], j.value, mw.value, [lv[i] for i in range(min(nl.value, level_cap))]
