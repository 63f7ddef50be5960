"""ctypes binding of libheipa.so (include/heipa.h) for tests, the benchmark and
Python callers.

Device arrays are torch tensors on the HIP device; torch only provides the
memory.  Every library call is synchronous at return; the wrapper
synchronises torch's stream before a call so tensors written by torch are
complete.  The binding fails loudly when the native library is missing: there
is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PKG_ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libheipa.so")
HEADER_PATH = os.path.abspath(os.path.join(PKG_ROOT, "..", "include", "heipa.h"))

OK, ERR_OOM, ERR_INVALID, ERR_OVERFLOW, ERR_DEVICE, ERR_IO = 0, 1, 2, 3, 4, 5
OPT_RESOLVE_CAP = 1
OPT_BLOCK_JET = 2
OPT_GRAPHS = 3
OPT_PROF_SAMPLE = 4
OPT_BUCKET_COPIES = 5
OPT_BLOCK_LDS = 6
OPT_HM_WORKERS = 7
OPT_HM_WORKER_BYTES = 8
OPT_RACE = 9
OPT_HIER_CACHE = 10
OPT_GRAPH_KEY = 11
OPT_GROUP_CAP = 12
OPT_BLOCK_GROUP_CAP = 13
OPT_REGION_EDGES = 14
UNIFORM_VW, UNIFORM_EW = 1, 2

u32p = ctypes.POINTER(ctypes.c_uint32)
s32p = ctypes.POINTER(ctypes.c_int32)
s64p = ctypes.POINTER(ctypes.c_int64)
u64p = ctypes.POINTER(ctypes.c_uint64)


class HeipaError(RuntimeError):
    def __init__(self, what, code, msg=""):
        super().__init__(f"{what} failed with status {code}: {msg}")
        self.code = code


class Graph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("g_weight", ctypes.c_int64),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("xadj", ctypes.c_void_p), ("adjncy", ctypes.c_void_p),
                ("adjwgt", ctypes.c_void_p), ("vwgt", ctypes.c_void_p)]


class MatchStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in
                ("rounds", "after_hem", "after_leaf", "after_twin", "after_relative",
                 "ran_leaf", "ran_twin", "ran_relative")]


class JetTrace(ctypes.Structure):
    _fields_ = [("ratio_idx", ctypes.c_uint32), ("iteration", ctypes.c_uint32), ("mode", ctypes.c_uint32),
                ("n_cand", ctypes.c_uint32), ("n_moves", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("cut", ctypes.c_int64), ("maxw", ctypes.c_int64),
                ("moves_hash", ctypes.c_uint64), ("gain_hash", ctypes.c_uint64)]


JOB_PARTITION, JOB_HM = 0, 1


class Job(ctypes.Structure):
    """heipa_job: one partition or HM request for a Pool."""
    _fields_ = [("graph", Graph), ("kind", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("hier", ctypes.c_uint32 * 8), ("l", ctypes.c_uint32), ("ultra", ctypes.c_int32),
                ("imbalance", ctypes.c_double), ("seed", ctypes.c_uint64), ("part", ctypes.c_void_p),
                ("status", ctypes.c_int32), ("slot", ctypes.c_int32), ("cut", ctypes.c_int64),
                ("maxw", ctypes.c_int64), ("estimate", ctypes.c_uint64), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double), ("state", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("peak", ctypes.c_uint64), ("graph_key", ctypes.c_uint64)]


class Tune(ctypes.Structure):
    _fields_ = [("group_cap", ctypes.c_uint32), ("block_group_cap", ctypes.c_uint32), ("block_jet", ctypes.c_uint32),
                ("runs", ctypes.c_uint32), ("ms_start", ctypes.c_double), ("ms_best", ctypes.c_double)]


class LevelTrace(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("nc", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("mapping_hash", ctypes.c_uint64), ("cut", ctypes.c_int64), ("maxw", ctypes.c_int64)]


class SolveStats(ctypes.Structure):
    _fields_ = [("cut", ctypes.c_int64), ("maxw", ctypes.c_int64), ("lmax", ctypes.c_int64),
                ("initial_cut", ctypes.c_int64), ("levels", ctypes.c_uint32), ("jet_iterations", ctypes.c_uint32),
                ("ms_coarsening", ctypes.c_double), ("ms_contraction", ctypes.c_double),
                ("ms_initial", ctypes.c_double), ("ms_uncontraction", ctypes.c_double),
                ("ms_refinement", ctypes.c_double)]


_lib = None


def lib():
    """Loads libheipa.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built; run __graft_entry__.build() or make -C gpu-heipa_amd")
        # torch bundles its own HIP runtime under an unversioned file name.
        # Loading it first lets libheipa's libamdhip64.so.7 dependency bind to
        # that same runtime; two HIP/HSA runtimes in one process cannot share
        # the device.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        vp, c_int, c_u32, c_u64, c_dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
        GP = ctypes.POINTER(Graph)
        sig = {
            "heipa_ctx_create": (c_int, [c_int, c_u64, ctypes.POINTER(vp)]),
            "heipa_ctx_destroy": (c_int, [vp]),
            "heipa_ctx_last_error": (ctypes.c_char_p, [vp]),
            "heipa_ctx_stream": (vp, [vp]),
            "heipa_arena_mark": (c_u64, [vp]),
            "heipa_arena_release": (c_int, [vp, c_u64]),
            "heipa_arena_peak": (c_u64, [vp]),
            "heipa_arena_reset_peak": (c_int, [vp]),
            "heipa_ctx_share_hier_cache": (c_int, [vp, vp]),
            "heipa_hier_cache_stats": (c_int, [vp, u64p, u64p, u64p, u64p]),
            "heipa_pool_hier_cache_stats": (c_int, [vp, u64p, u64p, u64p, u64p]),
            "heipa_download": (c_int, [vp, vp, vp, c_u64]),
            "heipa_upload": (c_int, [vp, vp, vp, c_u64]),
            "heipa_edge_cut": (c_int, [vp, GP, vp, s64p]),
            "heipa_edge_cut_noalloc": (c_int, [GP, vp, vp, vp]),
            "heipa_comm_cost": (c_int, [vp, GP, vp, vp, c_u32, s64p]),
            "heipa_distance_matrix": (c_int, [u32p, s32p, c_u32, s32p]),
            "heipa_block_weights": (c_int, [vp, GP, vp, c_u32, s64p, s64p]),
            "heipa_match": (c_int, [vp, GP, c_u32, vp, u32p, ctypes.POINTER(MatchStats)]),
            "heipa_contract": (c_int, [vp, GP, vp, c_u32, GP]),
            "heipa_project": (c_int, [vp, vp, c_u32, vp, vp]),
            "heipa_jet_refine": (c_int, [vp, GP, vp, vp, c_u32, ctypes.c_int64, c_int, c_u32, s64p, s64p]),
            "heipa_jet_refine_traced": (c_int, [vp, GP, vp, vp, c_u32, ctypes.c_int64, c_int, c_u32, s64p, s64p,
                                                ctypes.POINTER(JetTrace), c_u32, u32p]),
            "heipa_partition": (c_int, [vp, GP, c_u32, c_dbl, c_u64, c_int, vp, ctypes.POINTER(SolveStats),
                                        ctypes.POINTER(LevelTrace), c_u32]),
            "heipa_hm_partition": (c_int, [vp, GP, u32p, c_u32, c_dbl, c_u64, c_int, vp]),
            "heipa_map_refine": (c_int, [vp, GP, vp, vp, vp, c_u32, ctypes.c_int64, c_u32, s64p, s64p]),
            "heipa_map_refine_traced": (c_int, [vp, GP, vp, vp, vp, c_u32, ctypes.c_int64, c_u32, s64p, s64p,
                                                ctypes.POINTER(JetTrace), c_u32, u32p]),
            "heipa_im_partition": (c_int, [vp, GP, u32p, s32p, c_u32, c_dbl, c_u64, vp, ctypes.POINTER(SolveStats),
                                           ctypes.POINTER(LevelTrace), c_u32]),
            "heipa_global_multisection": (c_int, [c_u32, u32p, u32p, s32p, s32p, u32p, c_u32, c_dbl, c_u64, u32p]),
            "heipa_adaptive_imbalance": (c_dbl, [c_dbl, ctypes.c_int64, c_u32, ctypes.c_int64, c_u32, c_u64]),
            "heipa_ctx_set_option": (c_int, [vp, c_int, c_u64]),
            "heipa_pool_create": (c_int, [c_int, c_u32, c_u64, c_u64, ctypes.POINTER(vp)]),
            "heipa_pool_destroy": (c_int, [vp]),
            "heipa_job_estimate": (c_u64, [ctypes.POINTER(Job)]),
            "heipa_pool_run": (c_int, [vp, ctypes.POINTER(Job), c_u32]),
            "heipa_pool_submit": (c_int, [vp, ctypes.POINTER(Job)]),
            "heipa_pool_wait": (c_int, [vp, ctypes.POINTER(Job), c_dbl]),
            "heipa_pool_clock": (c_dbl, [vp]),
            "heipa_pool_set_option": (c_int, [vp, c_int, c_u64]),
            "heipa_pool_autotune": (c_int, [vp, ctypes.POINTER(Job), c_u32, ctypes.POINTER(Tune)]),
            "heipa_prof_enable": (c_int, [vp, c_int]),
            "heipa_prof_reset": (c_int, [vp]),
            "heipa_prof_filter": (c_int, [vp, ctypes.c_char_p]),
            "heipa_prof_count": (c_u32, [vp]),
            "heipa_prof_get": (c_int, [vp, c_u32, ctypes.c_char_p, c_u32, ctypes.POINTER(c_u64),
                                       ctypes.POINTER(c_dbl), ctypes.POINTER(c_dbl)]),
            "heipa_read_metis": (c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
            "heipa_write_metis": (c_int, [ctypes.c_char_p, vp]),
            "heipa_write_partition": (c_int, [ctypes.c_char_p, u32p, c_u32]),
            "heipa_gen_grid": (c_int, [c_u32, c_u32, ctypes.POINTER(vp)]),
            "heipa_gen_rgg": (c_int, [c_u32, c_dbl, c_u64, c_u32, ctypes.POINTER(vp)]),
            "heipa_gen_rmat": (c_int, [c_u32, c_u32, c_dbl, c_dbl, c_dbl, c_dbl, c_u64, c_u32, c_int,
                                       ctypes.POINTER(vp)]),
            "heipa_host_graph_from_arrays": (c_int, [c_u32, u32p, u32p, s32p, s32p, ctypes.POINTER(vp)]),
            "heipa_host_graph_view": (c_int, [vp, u32p, u32p, s64p, u32p, ctypes.POINTER(u32p), ctypes.POINTER(u32p),
                                              ctypes.POINTER(s32p), ctypes.POINTER(s32p)]),
            "heipa_host_graph_free": (None, [vp]),
            "heipa_initial_partition": (c_int, [c_u32, u32p, u32p, s32p, s32p, c_u32, c_dbl, c_u64, u32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols_from_header(path=HEADER_PATH):
    """Names of the functions include/heipa.h declares."""
    import re
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(heipa_\w+)\s*\(", txt, re.M)))


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# ---------------------------------------------------------------- host graphs
@dataclass
class HostGraph:
    """A CSR graph in host memory (numpy arrays).  adjwgt / vwgt are None when uniform."""
    xadj: np.ndarray
    adjncy: np.ndarray
    adjwgt: np.ndarray | None = None
    vwgt: np.ndarray | None = None

    @property
    def n(self):
        return len(self.xadj) - 1

    @property
    def m(self):
        return int(self.xadj[-1])

    @property
    def g_weight(self):
        return int(self.vwgt.sum()) if self.vwgt is not None else self.n

    @property
    def flags(self):
        return (UNIFORM_VW if self.vwgt is None else 0) | (UNIFORM_EW if self.adjwgt is None else 0)


def _from_handle(h):
    L = lib()
    n, m, flags = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    gw = ctypes.c_int64()
    xa, ad = u32p(), u32p()
    ew, vw = s32p(), s32p()
    L.heipa_host_graph_view(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(gw), ctypes.byref(flags),
                            ctypes.byref(xa), ctypes.byref(ad), ctypes.byref(ew), ctypes.byref(vw))
    N, M = n.value, m.value
    xadj = np.ctypeslib.as_array(xa, shape=(N + 1,)).copy()
    adjncy = np.ctypeslib.as_array(ad, shape=(M,)).copy() if M else np.zeros(0, np.uint32)
    adjwgt = np.ctypeslib.as_array(ew, shape=(M,)).copy() if (ew and M) else (np.zeros(0, np.int32) if ew else None)
    vwgt = np.ctypeslib.as_array(vw, shape=(N,)).copy() if (vw and N) else None
    L.heipa_host_graph_free(h)
    return HostGraph(xadj, adjncy, adjwgt, vwgt)


def _gen(fn, *args):
    h = ctypes.c_void_p()
    rc = fn(*args, ctypes.byref(h))
    if rc != OK:
        raise HeipaError(fn.__name__, rc)
    return _from_handle(h)


def gen_grid(rows, cols):
    return _gen(lib().heipa_gen_grid, rows, cols)


def gen_rgg(n, avg_deg, seed, wmod=0):
    return _gen(lib().heipa_gen_rgg, n, avg_deg, seed, wmod)


def gen_rmat(scale, n=0, edge_factor=16.0, a=0.57, b=0.19, c=0.19, seed=1, wmod=0, explicit_vw=False):
    return _gen(lib().heipa_gen_rmat, scale, n, edge_factor, a, b, c, seed, wmod, int(explicit_vw))


def read_metis(path):
    return _gen(lib().heipa_read_metis, path.encode())


def _handle_of(g: HostGraph):
    L = lib()
    h = ctypes.c_void_p()
    xadj = np.ascontiguousarray(g.xadj, np.uint32)
    adjncy = np.ascontiguousarray(g.adjncy, np.uint32)
    ew = None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, np.int32)
    vw = None if g.vwgt is None else np.ascontiguousarray(g.vwgt, np.int32)
    rc = L.heipa_host_graph_from_arrays(g.n, _ptr(xadj, ctypes.c_uint32), _ptr(adjncy, ctypes.c_uint32),
                                        None if ew is None else _ptr(ew, ctypes.c_int32),
                                        None if vw is None else _ptr(vw, ctypes.c_int32), ctypes.byref(h))
    if rc != OK:
        raise HeipaError("heipa_host_graph_from_arrays", rc)
    return h


def write_metis(path, g: HostGraph):
    L = lib()
    h = _handle_of(g)
    rc = L.heipa_write_metis(path.encode(), h)
    L.heipa_host_graph_free(h)
    if rc != OK:
        raise HeipaError("heipa_write_metis", rc)


def write_partition(path, part):
    part = np.ascontiguousarray(part, np.uint32)
    rc = lib().heipa_write_partition(path.encode(), _ptr(part, ctypes.c_uint32), len(part))
    if rc != OK:
        raise HeipaError("heipa_write_partition", rc)


def distance_matrix(hier, dist):
    hier = np.ascontiguousarray(hier, np.uint32)
    dist = np.ascontiguousarray(dist, np.int32)
    k = int(np.prod(hier.astype(np.int64)))
    D = np.zeros(k * k, np.int32)
    rc = lib().heipa_distance_matrix(_ptr(hier, ctypes.c_uint32), _ptr(dist, ctypes.c_int32), len(hier),
                                     _ptr(D, ctypes.c_int32))
    if rc != OK:
        raise HeipaError("heipa_distance_matrix", rc)
    return D.reshape(k, k)


def initial_partition(g: HostGraph, k, imbalance, seed):
    part = np.zeros(g.n, np.uint32)
    xadj = np.ascontiguousarray(g.xadj, np.uint32)
    adjncy = np.ascontiguousarray(g.adjncy, np.uint32)
    ew = None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, np.int32)
    vw = None if g.vwgt is None else np.ascontiguousarray(g.vwgt, np.int32)
    rc = lib().heipa_initial_partition(g.n, _ptr(xadj, ctypes.c_uint32), _ptr(adjncy, ctypes.c_uint32),
                                       None if ew is None else _ptr(ew, ctypes.c_int32),
                                       None if vw is None else _ptr(vw, ctypes.c_int32), k, imbalance, seed,
                                       _ptr(part, ctypes.c_uint32))
    if rc != OK:
        raise HeipaError("heipa_initial_partition", rc)
    return part


def global_multisection(g: HostGraph, hier, imbalance, seed):
    """The IM path's initial mapping (host): ids < prod(hier)."""
    part = np.zeros(max(g.n, 1), np.uint32)
    xadj = np.ascontiguousarray(g.xadj, np.uint32)
    adjncy = np.ascontiguousarray(g.adjncy, np.uint32)
    ew = None if g.adjwgt is None else np.ascontiguousarray(g.adjwgt, np.int32)
    vw = None if g.vwgt is None else np.ascontiguousarray(g.vwgt, np.int32)
    h = np.ascontiguousarray(hier, np.uint32)
    rc = lib().heipa_global_multisection(g.n, _ptr(xadj, ctypes.c_uint32), _ptr(adjncy, ctypes.c_uint32),
                                         None if ew is None else _ptr(ew, ctypes.c_int32),
                                         None if vw is None else _ptr(vw, ctypes.c_int32), _ptr(h, ctypes.c_uint32),
                                         len(h), imbalance, seed, _ptr(part, ctypes.c_uint32))
    if rc != OK:
        raise HeipaError("heipa_global_multisection", rc)
    return part[:g.n]


def adaptive_imbalance(eps, gW, gk, lW, krem, depth):
    return lib().heipa_adaptive_imbalance(eps, gW, gk, lW, krem, depth)


# ---------------------------------------------------------------- device side
def _torch():
    import torch
    return torch


class DeviceGraph:
    """Device copy of a HostGraph (torch tensors on cuda:N) plus its heipa_graph descriptor."""

    def __init__(self, g: HostGraph, device="cuda:0"):
        torch = _torch()
        self.host = g
        self.xadj = torch.from_numpy(np.ascontiguousarray(g.xadj, np.uint32).view(np.int32)).to(device)
        self.adjncy = torch.from_numpy(np.ascontiguousarray(g.adjncy, np.uint32).view(np.int32)).to(device)
        self.adjwgt = None if g.adjwgt is None else torch.from_numpy(np.ascontiguousarray(g.adjwgt, np.int32)).to(device)
        self.vwgt = None if g.vwgt is None else torch.from_numpy(np.ascontiguousarray(g.vwgt, np.int32)).to(device)
        self.desc = Graph(g.n, g.m, g.g_weight, g.flags, 0, self.xadj.data_ptr(), self.adjncy.data_ptr(),
                          self.adjwgt.data_ptr() if self.adjwgt is not None else None,
                          self.vwgt.data_ptr() if self.vwgt is not None else None)
        torch.cuda.synchronize()

    @property
    def n(self):
        return self.desc.n


class Context:
    """A heipa_ctx on one device (stream + LIFO arena)."""

    def __init__(self, device=0, arena_bytes=0):
        self.L = lib()
        self.h = ctypes.c_void_p()
        rc = self.L.heipa_ctx_create(device, arena_bytes, ctypes.byref(self.h))
        if rc != OK:
            raise HeipaError("heipa_ctx_create", rc)

    def close(self):
        if self.h:
            self.L.heipa_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, name, rc):
        if rc != OK:
            raise HeipaError(name, rc, self.L.heipa_ctx_last_error(self.h).decode())

    @staticmethod
    def _sync():
        _torch().cuda.synchronize()

    def edge_cut(self, g: DeviceGraph, part):
        self._sync()
        out = ctypes.c_int64()
        self._check("heipa_edge_cut", self.L.heipa_edge_cut(self.h, ctypes.byref(g.desc), part.data_ptr(), ctypes.byref(out)))
        return out.value

    def comm_cost(self, g: DeviceGraph, part, D, k):
        self._sync()
        out = ctypes.c_int64()
        self._check("heipa_comm_cost", self.L.heipa_comm_cost(self.h, ctypes.byref(g.desc), part.data_ptr(),
                                                              D.data_ptr(), k, ctypes.byref(out)))
        return out.value

    def block_weights(self, g: DeviceGraph, part, k):
        self._sync()
        bw = np.zeros(k, np.int64)
        mx = ctypes.c_int64()
        self._check("heipa_block_weights", self.L.heipa_block_weights(self.h, ctypes.byref(g.desc), part.data_ptr(), k,
                                                                      _ptr(bw, ctypes.c_int64), ctypes.byref(mx)))
        return bw, mx.value

    def match(self, g: DeviceGraph, seed=12345):
        torch = _torch()
        mapping = torch.empty(max(g.n, 1), dtype=torch.int32, device=g.xadj.device)
        nc = ctypes.c_uint32()
        st = MatchStats()
        self._sync()
        self._check("heipa_match", self.L.heipa_match(self.h, ctypes.byref(g.desc), seed, mapping.data_ptr(),
                                                      ctypes.byref(nc), ctypes.byref(st)))
        return mapping[:g.n], nc.value, st

    def contract(self, g: DeviceGraph, mapping, nc):
        """Returns (HostGraph of the coarse graph); the device copy is released."""
        self._sync()
        mark = self.L.heipa_arena_mark(self.h)
        out = Graph()
        self._check("heipa_contract", self.L.heipa_contract(self.h, ctypes.byref(g.desc), mapping.data_ptr(), nc,
                                                            ctypes.byref(out)))
        hg = self.download_graph(out)
        self._check("heipa_arena_release", self.L.heipa_arena_release(self.h, mark))
        return hg

    def project(self, mapping, coarse_part, fine_n):
        torch = _torch()
        fine = torch.empty(max(fine_n, 1), dtype=torch.int32, device=coarse_part.device)
        self._sync()
        self._check("heipa_project", self.L.heipa_project(self.h, mapping.data_ptr(), fine_n, coarse_part.data_ptr(),
                                                          fine.data_ptr()))
        return fine[:fine_n]

    def jet_refine(self, g: DeviceGraph, part, k, lmax, ultra, level, cut, maxw, trace_cap=0, bweights=None):
        """jet_refine on the device; bweights (optional device int64 tensor of k entries) receives the
        returned partition's block weights.  Returns (cut, maxw, trace)."""
        self._sync()
        c, mw = ctypes.c_int64(cut), ctypes.c_int64(maxw)
        bw = bweights.data_ptr() if bweights is not None else None
        if trace_cap:
            tr = (JetTrace * trace_cap)()
            tl = ctypes.c_uint32()
            self._check("heipa_jet_refine_traced", self.L.heipa_jet_refine_traced(
                self.h, ctypes.byref(g.desc), part.data_ptr(), bw, k, lmax, int(ultra), level, ctypes.byref(c),
                ctypes.byref(mw), tr, trace_cap, ctypes.byref(tl)))
            return c.value, mw.value, [tr[i] for i in range(min(tl.value, trace_cap))]
        self._check("heipa_jet_refine", self.L.heipa_jet_refine(self.h, ctypes.byref(g.desc), part.data_ptr(), bw, k,
                                                                lmax, int(ultra), level, ctypes.byref(c),
                                                                ctypes.byref(mw)))
        return c.value, mw.value, None

    def partition(self, g: DeviceGraph, k, imbalance=0.03, seed=1, ultra=False, part=None, level_cap=0):
        torch = _torch()
        if part is None:
            part = torch.empty(max(g.n, 1), dtype=torch.int32, device=g.xadj.device)
        st = SolveStats()
        lv = (LevelTrace * level_cap)() if level_cap else None
        self._sync()
        self._check("heipa_partition", self.L.heipa_partition(self.h, ctypes.byref(g.desc), k, imbalance, seed,
                                                              int(ultra), part.data_ptr(), ctypes.byref(st), lv,
                                                              level_cap))
        levels = [lv[i] for i in range(min(st.levels, level_cap))] if level_cap else None
        return part[:g.n], st, levels

    def map_refine(self, g: DeviceGraph, part, D, k, lmax, level, J, maxw, trace_cap=0, bweights=None):
        """ProMap_jet_refine on the device: D is a device int32 k x k tensor; bweights as for jet_refine;
        returns (J, maxw, trace)."""
        self._sync()
        j, mw = ctypes.c_int64(J), ctypes.c_int64(maxw)
        bw = bweights.data_ptr() if bweights is not None else None
        if trace_cap:
            tr = (JetTrace * trace_cap)()
            tl = ctypes.c_uint32()
            self._check("heipa_map_refine_traced", self.L.heipa_map_refine_traced(
                self.h, ctypes.byref(g.desc), part.data_ptr(), bw, D.data_ptr(), k, lmax, level, ctypes.byref(j),
                ctypes.byref(mw), tr, trace_cap, ctypes.byref(tl)))
            return j.value, mw.value, [tr[i] for i in range(min(tl.value, trace_cap))]
        self._check("heipa_map_refine", self.L.heipa_map_refine(self.h, ctypes.byref(g.desc), part.data_ptr(), bw,
                                                                D.data_ptr(), k, lmax, level, ctypes.byref(j),
                                                                ctypes.byref(mw)))
        return j.value, mw.value, None

    def im_partition(self, g: DeviceGraph, hier, dist, imbalance=0.03, seed=1, part=None, level_cap=0):
        """ProMapSolver::solve (--config IM); returns (part, stats, levels) with J in the cut fields."""
        torch = _torch()
        if part is None:
            part = torch.empty(max(g.n, 1), dtype=torch.int32, device=g.xadj.device)
        h = np.ascontiguousarray(hier, np.uint32)
        d = np.ascontiguousarray(dist, np.int32)
        st = SolveStats()
        lv = (LevelTrace * level_cap)() if level_cap else None
        self._sync()
        self._check("heipa_im_partition", self.L.heipa_im_partition(
            self.h, ctypes.byref(g.desc), _ptr(h, ctypes.c_uint32), _ptr(d, ctypes.c_int32), len(h), imbalance, seed,
            part.data_ptr(), ctypes.byref(st), lv, level_cap))
        levels = [lv[i] for i in range(min(st.levels, level_cap))] if level_cap else None
        return part[:g.n], st, levels

    def hm_partition(self, g: DeviceGraph, hier, imbalance=0.03, seed=1, ultra=False, part=None):
        torch = _torch()
        if part is None:
            part = torch.empty(max(g.n, 1), dtype=torch.int32, device=g.xadj.device)
        h = np.ascontiguousarray(hier, np.uint32)
        self._sync()
        self._check("heipa_hm_partition", self.L.heipa_hm_partition(self.h, ctypes.byref(g.desc),
                                                                    _ptr(h, ctypes.c_uint32), len(h), imbalance, seed,
                                                                    int(ultra), part.data_ptr()))
        return part[:g.n]

    def download(self, host_array: np.ndarray, dev_ptr):
        self._check("heipa_download", self.L.heipa_download(self.h, host_array.ctypes.data, dev_ptr, host_array.nbytes))
        return host_array

    def download_graph(self, desc: Graph) -> HostGraph:
        """Host copy of a device heipa_graph."""
        n, m = desc.n, desc.m

        def dl(ptr, count, dtype):
            buf = np.zeros(count, dtype)
            if count and ptr:
                self.download(buf, ptr)
            return buf

        xadj = dl(desc.xadj, n + 1, np.uint32)
        adjncy = dl(desc.adjncy, m, np.uint32)
        adjwgt = dl(desc.adjwgt, m, np.int32) if desc.adjwgt else None
        vwgt = dl(desc.vwgt, n, np.int32) if desc.vwgt else None
        return HostGraph(xadj, adjncy, adjwgt, vwgt)

    def set_option(self, option, value):
        self._check("heipa_ctx_set_option", self.L.heipa_ctx_set_option(self.h, int(option), int(value)))

    def hier_cache_stats(self):
        """(graphs held, device bytes, levels reused, levels built) of the context's hierarchy cache"""
        v = [ctypes.c_uint64() for _ in range(4)]
        self._check("heipa_hier_cache_stats", self.L.heipa_hier_cache_stats(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    # per-kernel timing
    def prof_enable(self, on=True):
        self._check("heipa_prof_enable", self.L.heipa_prof_enable(self.h, int(on)))

    def prof_filter(self, prefix=""):
        self._check("heipa_prof_filter", self.L.heipa_prof_filter(self.h, prefix.encode()))

    def prof_reset(self):
        self._check("heipa_prof_reset", self.L.heipa_prof_reset(self.h))

    def prof_stats(self):
        out = {}
        cnt = self.L.heipa_prof_count(self.h)
        buf = ctypes.create_string_buffer(128)
        for i in range(cnt):
            ln, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
            self._check("heipa_prof_get", self.L.heipa_prof_get(self.h, i, buf, 128, ctypes.byref(ln), ctypes.byref(ms),
                                                                ctypes.byref(by)))
            out[buf.value.decode()] = {"launches": ln.value, "ms": ms.value, "bytes": by.value}
        return out


class Pool:
    """heipa_pool: `slots` workers, each with its own context (stream, arena of
    arena_bytes), running jobs from one FIFO queue under a device-memory budget."""

    def __init__(self, slots, arena_bytes, budget=0, device=0):
        self.L = lib()
        self.h = ctypes.c_void_p()
        rc = self.L.heipa_pool_create(device, slots, arena_bytes, budget, ctypes.byref(self.h))
        if rc != OK:
            raise HeipaError("heipa_pool_create", rc)
        self._keep = {}

    @staticmethod
    def job(g: "DeviceGraph", k=0, hier=None, imbalance=0.03, seed=1, ultra=False, part=None, graph_key=0):
        """A job for g (partition into k blocks, or HM with hierarchy `hier`); the
        device partition tensor is allocated unless given; graph_key names g for
        resident hierarchies (OPT_HIER_CACHE)."""
        torch = _torch()
        j = Job()
        j.graph = g.desc
        if hier:
            j.kind, j.l = JOB_HM, len(hier)
            for i, a in enumerate(hier):
                j.hier[i] = a
        else:
            j.kind, j.k = JOB_PARTITION, k
        j.ultra, j.imbalance, j.seed = int(ultra), imbalance, seed
        j.graph_key = graph_key
        if part is None:
            part = torch.empty(max(g.n, 1), dtype=torch.int32, device=g.xadj.device)
        j.part = part.data_ptr()
        j._part, j._graph = part, g  # keep the tensors alive with the job
        return j

    def estimate(self, job):
        return self.L.heipa_job_estimate(ctypes.byref(job))

    def run(self, jobs):
        """Batched form: all jobs, returns when every one is done."""
        _torch().cuda.synchronize()
        arr = (Job * len(jobs))(*jobs)
        rc = self.L.heipa_pool_run(self.h, arr, len(jobs))
        if rc != OK:
            raise HeipaError("heipa_pool_run", rc)
        for i, j in enumerate(jobs):  # copy the results back into the caller's objects
            ctypes.memmove(ctypes.addressof(j), ctypes.addressof(arr[i]), ctypes.sizeof(Job))
        return jobs

    def autotune(self, sample):
        """heipa_pool_autotune over the sample jobs: the launch-shape setting it
        leaves on the workers, as a dict (the sample jobs hold their last run's
        results)."""
        _torch().cuda.synchronize()
        arr = (Job * len(sample))(*sample)
        t = Tune()
        rc = self.L.heipa_pool_autotune(self.h, arr, len(sample), ctypes.byref(t))
        if rc != OK:
            raise HeipaError("heipa_pool_autotune", rc)
        for i, j in enumerate(sample):
            ctypes.memmove(ctypes.addressof(j), ctypes.addressof(arr[i]), ctypes.sizeof(Job))
        return {"group_cap": t.group_cap, "block_group_cap": t.block_group_cap, "block_jet": t.block_jet,
                "runs": t.runs, "ms_start": round(t.ms_start, 3), "ms_best": round(t.ms_best, 3)}

    def submit(self, job):
        """Served form: queue one job (it must stay alive until done)."""
        self._keep[id(job)] = job
        rc = self.L.heipa_pool_submit(self.h, ctypes.byref(job))
        if rc != OK:
            raise HeipaError("heipa_pool_submit", rc)

    def wait(self, job, timeout_ms=-1.0):
        done = self.L.heipa_pool_wait(self.h, ctypes.byref(job), float(timeout_ms)) == 1
        if done:
            self._keep.pop(id(job), None)
        return done

    def clock(self):
        return self.L.heipa_pool_clock(self.h)

    def hier_cache_stats(self):
        """(graphs held, device bytes, levels reused, levels built) of the shared hierarchy cache"""
        v = [ctypes.c_uint64() for _ in range(4)]
        rc = self.L.heipa_pool_hier_cache_stats(self.h, *[ctypes.byref(x) for x in v])
        if rc != OK:
            raise HeipaError("heipa_pool_hier_cache_stats", rc)
        return tuple(x.value for x in v)

    def set_option(self, option, value):
        rc = self.L.heipa_pool_set_option(self.h, int(option), int(value))
        if rc != OK:
            raise HeipaError("heipa_pool_set_option", rc)

    def close(self):
        if self.h:
            self.L.heipa_pool_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_numpy_u32(t):
    """torch int32 tensor holding u32 values -> numpy uint32."""
    return t.cpu().numpy().view(np.uint32)


def from_numpy_u32(a, device="cuda:0"):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to(device)
