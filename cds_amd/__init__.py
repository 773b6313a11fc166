"""Python host side of the MI355X k-clique densest-subgraph engine.

Thin ctypes bindings over the C ABI in include/cds.h (libcds_amd.so, built
in-tree by cds_amd/build.py).  Mirrors the reference's entry point, the
`main <graph> <k> <pSize> <cpSize> <glBufferSize> <partitionSize>
<earlyStop>` process of cudaCode/main.cu:865-887: `solve()` takes the same
parameters and returns the same report fields plus the CPU oracle's result
block (cpuCode/c++/src/helpers.cpp:1553-1588).

There is no CPU fallback: if the library is missing, importing works but
every call raises `EngineUnavailable`.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcds_amd.so")
# launch shapes found by `python -m cds_amd.tune` (loaded by Engine by default)
TUNE_CACHE = os.path.join(HERE, "tune_cache.txt")
CLI_PATH = os.path.join(HERE, "bin", "cds_main")

STATUS = {
    0: "OK", 1: "INVALID", 2: "IO", 3: "HIP", 4: "OOM", 5: "RANGE",
    6: "NOT_CONVERGED", 7: "UNSUPPORTED", 8: "BUSY", 9: "CANCELLED",
    10: "INTERNAL",
}


class EngineUnavailable(RuntimeError):
    pass


class CdsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)


class CGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint64),
                ("m_header", ctypes.c_uint64), ("offsets", u32p),
                ("neighbors", u32p), ("max_degree", ctypes.c_uint32)]


class CParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("p_size", ctypes.c_uint32),
                ("cp_size", ctypes.c_uint32),
                ("gl_buffer_size", ctypes.c_uint32),
                ("partition_size", ctypes.c_uint32),
                ("early_stop", ctypes.c_int32), ("ub1", ctypes.c_uint32),
                ("ub2", ctypes.c_uint32),
                ("want_vertex_arrays", ctypes.c_uint32)]


class CTimes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "upload", "dag", "listing", "core_decompose", "densest_core", "prune",
        "components", "exact", "total")]


class CResult(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_uint32), ("n", ctypes.c_uint32), ("m", ctypes.c_uint64),
        ("max_degree", ctypes.c_uint32), ("total_cliques", ctypes.c_uint64),
        ("clique_degree", u64p), ("core", u32p), ("kmax", ctypes.c_uint64),
        ("k_prime", ctypes.c_uint64), ("peel_max_density", ctypes.c_double),
        ("core_density", ctypes.c_double), ("core_size", ctypes.c_uint32),
        ("core_edges", ctypes.c_uint64), ("core_cliques", ctypes.c_uint64),
        ("deleted_cliques", ctypes.c_uint64), ("pruned_edges", ctypes.c_uint64),
        ("num_components", ctypes.c_uint32),
        ("paper_lower_bound", ctypes.c_double), ("ub1_val", ctypes.c_float),
        ("ub2_val", ctypes.c_float), ("density", ctypes.c_float),
        ("size", ctypes.c_uint32), ("vertices", u32p),
        ("set_cliques", ctypes.c_uint64), ("set_size", ctypes.c_uint32),
        ("trace_len", ctypes.c_uint32), ("trace_lb", f32p), ("trace_ub", f32p),
        ("trace_alpha", f32p), ("trace_component", u32p),
        ("trace_margin", f64p), ("flow_networks_solved", ctypes.c_uint32),
        ("flow_sweeps", ctypes.c_uint64), ("peel_levels", ctypes.c_uint32),
        ("replay_phases", ctypes.c_uint32), ("replay_grid_phases", ctypes.c_uint32),
        ("oracle_quirk_flags", ctypes.c_uint32), ("dag_arcs", ctypes.c_uint64),
        ("list_work", ctypes.c_uint64), ("max_out_degree", ctypes.c_uint32),
        ("list_items", ctypes.c_uint64), ("t_list_count_ms", ctypes.c_double),
        ("t_list_write_ms", ctypes.c_double),
        ("peak_device_bytes", ctypes.c_uint64), ("peel_grid", ctypes.c_uint32),
        ("flow_launches", ctypes.c_uint32),
        ("density_certificates", ctypes.c_uint32),
        ("peel_distinct_levels", ctypes.c_uint32),
        ("peel_coarse_shift", ctypes.c_uint32),
        ("peel_refined_steps", ctypes.c_uint32),
        ("peel_launches", ctypes.c_uint32),
        ("peel_alive_scanned", ctypes.c_uint64),
        ("peel_split_ms", ctypes.c_double * 4),
        ("t_incidence_ms", ctypes.c_double), ("t_peel_kernel_ms", ctypes.c_double),
        ("t_replay_ms", ctypes.c_double), ("t_replay_prep_ms", ctypes.c_double),
        ("t_flow_kernel_ms", ctypes.c_double), ("flow_clique_visits", ctypes.c_uint64),
        ("flow_arc_visits", ctypes.c_uint64), ("flow_relabels", ctypes.c_uint64),
        ("times", CTimes),
        ("flow_grid", ctypes.c_uint32), ("replay_grid", ctypes.c_uint32),
        ("coop_shrinks", ctypes.c_uint32), ("fragile_decisions", ctypes.c_uint32),
        ("trace_fragile", ctypes.POINTER(ctypes.c_uint8)),
        ("trace_drift", ctypes.POINTER(ctypes.c_double)),
        ("prune_tau", ctypes.c_uint64), ("prune_vertices", ctypes.c_uint32),
        ("prune_passes", ctypes.c_uint32), ("dense_peel", ctypes.c_uint32),
        ("flow_implicit", ctypes.c_uint32), ("flow_graph_launches", ctypes.c_uint32),
        ("replay_bits_phases", ctypes.c_uint32), ("refine_bits_steps", ctypes.c_uint32),
        ("list_staged", ctypes.c_uint32), ("prune_unlinks", ctypes.c_uint32)]


class COpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_uint64),
                ("profile_kernels", ctypes.c_uint32), ("tune_cache", ctypes.c_char_p)]


TUNING_FIELDS = ("list_threads", "list_chunk", "peel_blocks", "peel_coarse_shift",
                 "replay_blocks", "flow_blocks")


class CTuning(ctypes.Structure):
    _fields_ = [("list_threads", ctypes.c_uint32), ("list_chunk", ctypes.c_uint32),
                ("peel_blocks", ctypes.c_uint32), ("peel_coarse_shift", ctypes.c_int32),
                ("replay_blocks", ctypes.c_uint32), ("flow_blocks", ctypes.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f in TUNING_FIELDS}

    @classmethod
    def from_dict(cls, d):
        t = cls(0, 0, 0, -1, 0, 0)
        for f, v in d.items():
            if f not in TUNING_FIELDS:
                raise ValueError(f"unknown tuning field {f}")
            setattr(t, f, int(v))
        return t


class CServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32),
                ("max_wait_us", ctypes.c_uint32), ("large_vertices", ctypes.c_uint32)]


class CServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in
                ("submitted", "admitted", "completed", "failed", "batches",
                 "max_batch_seen", "peak_admitted_bytes", "queued")]


class CRequestTimes(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("service_ms", ctypes.c_double)]


# Every function include/cds.h declares, with its ctypes signature.
_SIGNATURES = {
    "cds_last_error": (ctypes.c_char_p, []),
    "cds_version": (ctypes.c_char_p, []),
    "cds_struct_size": (ctypes.c_size_t, [ctypes.c_int]),
    "cds_graph_load_adjlist": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CGraph)]),
    "cds_graph_from_csr": (ctypes.c_int, [ctypes.c_uint32, u32p, u32p, ctypes.POINTER(CGraph)]),
    "cds_graph_write_adjlist": (ctypes.c_int, [ctypes.POINTER(CGraph), ctypes.c_char_p]),
    "cds_result_json": (ctypes.c_size_t, [ctypes.POINTER(CResult), ctypes.c_char_p, ctypes.c_size_t]),
    "cds_graph_save_csr": (ctypes.c_int, [ctypes.POINTER(CGraph), ctypes.c_char_p]),
    "cds_graph_load_csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CGraph)]),
    "cds_graph_free": (None, [ctypes.POINTER(CGraph)]),
    "cds_engine_create": (ctypes.c_int, [ctypes.POINTER(COpts), ctypes.POINTER(ctypes.c_void_p)]),
    "cds_engine_destroy": (None, [ctypes.c_void_p]),
    "cds_engine_device": (ctypes.c_int, [ctypes.c_void_p]),
    "cds_graph_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CGraph), ctypes.POINTER(ctypes.c_void_p)]),
    "cds_dgraph_free": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "cds_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CParams), ctypes.POINTER(CResult)]),
    "cds_result_free": (None, [ctypes.POINTER(CResult)]),
    "cds_count_cliques": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, u64p, u64p, f64p]),
    "cds_count_cliques_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, u64p, u64p, f64p]),
    "cds_core_decompose": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CParams), ctypes.POINTER(CResult)]),
    "cds_solve_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(CParams), ctypes.c_size_t, ctypes.POINTER(CResult)]),
    "cds_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CServerOpts), ctypes.POINTER(ctypes.c_void_p)]),
    "cds_server_destroy": (None, [ctypes.c_void_p]),
    "cds_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CGraph), ctypes.POINTER(CParams), ctypes.POINTER(ctypes.c_uint64)]),
    "cds_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "cds_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(CResult), ctypes.POINTER(CRequestTimes)]),
    "cds_cancel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "cds_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CServerStats)]),
    "cds_engine_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CTuning)]),
    "cds_engine_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(CTuning)]),
    "cds_engine_load_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]),
    "cds_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(CParams), ctypes.c_char_p, ctypes.POINTER(CTuning), f64p, f64p]),
    "cds_tuning_lookup": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(CTuning)]),
    "cds_gen_chung_lu": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(CGraph)]),
    "cds_gen_rmat": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(CGraph)]),
}

_lib = None


def lib():
    """The loaded libcds_amd.so (raises EngineUnavailable if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineUnavailable(
                f"{LIB_PATH} is missing: run `python -m cds_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        for i, cls in enumerate((CGraph, CParams, CResult, COpts, CTuning)):
            if L.cds_struct_size(i) != ctypes.sizeof(cls):
                raise EngineUnavailable(
                    f"ABI mismatch for {cls.__name__}: C {L.cds_struct_size(i)}"
                    f" vs ctypes {ctypes.sizeof(cls)}")
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise CdsError(status, lib().cds_last_error().decode())


def tuning_lookup(path, arch, cus, k, m):
    """cds_tuning_lookup: the cache entry nearest to (arch, cus, k, m), as a
    dict, or None if no entry is within 2 edge classes."""
    t = CTuning()
    st = lib().cds_tuning_lookup(os.fsencode(path), arch.encode(), cus, k, m, ctypes.byref(t))
    if st == 1:  # CDS_ERR_INVALID: no entry near enough
        return None
    _check(st)
    return t.as_dict()


class Graph:
    """Host CSR graph owned by the library (cds_graph)."""

    def __init__(self):
        self._g = CGraph()

    @classmethod
    def load(cls, path):
        g = cls()
        _check(lib().cds_graph_load_adjlist(os.fsencode(path), ctypes.byref(g._g)))
        return g

    @classmethod
    def load_csr(cls, path):
        """Graph from a binary CSR file written by save_csr."""
        g = cls()
        _check(lib().cds_graph_load_csr(os.fsencode(path), ctypes.byref(g._g)))
        return g

    @classmethod
    def from_csr(cls, offsets, neighbors):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        neighbors = np.ascontiguousarray(neighbors, dtype=np.uint32)
        g = cls()
        _check(lib().cds_graph_from_csr(
            len(offsets) - 1, offsets.ctypes.data_as(u32p),
            neighbors.ctypes.data_as(u32p), ctypes.byref(g._g)))
        return g

    @classmethod
    def from_edges(cls, n, edges):
        """Undirected graph from an iterable of (u, v) pairs."""
        adj = [set() for _ in range(n)]
        for u, v in edges:
            if u != v:
                adj[u].add(v)
                adj[v].add(u)
        off = np.zeros(n + 1, dtype=np.uint32)
        for i in range(n):
            off[i + 1] = off[i] + len(adj[i])
        nb = np.array([w for i in range(n) for w in sorted(adj[i])], dtype=np.uint32)
        return cls.from_csr(off, nb)

    @classmethod
    def chung_lu(cls, n, exponent, mean_degree, planted_size=0, planted_p=0.0,
                 seed=1):
        g = cls()
        _check(lib().cds_gen_chung_lu(n, exponent, mean_degree, planted_size,
                                      planted_p, seed, ctypes.byref(g._g)))
        return g

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, planted_count=0,
             planted_min=0, planted_max=0, planted_p=0.0, seed=1):
        g = cls()
        _check(lib().cds_gen_rmat(scale, edge_factor, a, b, c, planted_count,
                                  planted_min, planted_max, planted_p, seed,
                                  ctypes.byref(g._g)))
        return g

    @property
    def n(self):
        return self._g.n

    @property
    def m(self):
        return self._g.m

    @property
    def max_degree(self):
        return self._g.max_degree

    def csr(self):
        """(offsets, neighbors) as numpy views (valid while self lives)."""
        off = np.ctypeslib.as_array(self._g.offsets, (self._g.n + 1,))
        nb = np.ctypeslib.as_array(self._g.neighbors, (max(1, int(off[-1])),))[: int(off[-1])]
        return off, nb

    def write(self, path):
        _check(lib().cds_graph_write_adjlist(ctypes.byref(self._g), os.fsencode(path)))

    def save_csr(self, path):
        _check(lib().cds_graph_save_csr(ctypes.byref(self._g), os.fsencode(path)))

    def __del__(self):
        if _lib is not None and self._g.offsets:
            _lib.cds_graph_free(ctypes.byref(self._g))


def _arr(ptr, count, dtype):
    if not ptr or count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, (count,)).astype(dtype, copy=True)


def result_json(r):
    """cds_result_json of a CResult: the result record as a JSON string."""
    need = lib().cds_result_json(ctypes.byref(r), None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib().cds_result_json(ctypes.byref(r), buf, need + 1)
    return buf.value.decode()


def result_to_dict(r, with_json=False):
    d = {f: getattr(r, f) for f, _ in CResult._fields_
         if f not in ("clique_degree", "core", "vertices", "trace_lb", "trace_ub",
                      "trace_alpha", "trace_component", "trace_margin", "times",
                      "peel_split_ms", "trace_fragile", "trace_drift")}
    d["vertices"] = _arr(r.vertices, r.set_size, np.uint32).tolist()
    d["clique_degree"] = _arr(r.clique_degree, r.n, np.uint64) if r.clique_degree else None
    d["core"] = _arr(r.core, r.n, np.uint32) if r.core else None
    n = r.trace_len
    d["trace_lb"] = _arr(r.trace_lb, n, np.float32).tolist()
    d["trace_ub"] = _arr(r.trace_ub, n, np.float32).tolist()
    d["trace_alpha"] = _arr(r.trace_alpha, n, np.float32).tolist()
    d["trace_component"] = _arr(r.trace_component, n, np.uint32).tolist()
    d["trace_margin"] = _arr(r.trace_margin, n, np.float64).tolist()
    d["trace_fragile"] = _arr(r.trace_fragile, n, np.uint8).tolist()
    d["trace_drift"] = _arr(r.trace_drift, n, np.float64).tolist()
    d["times"] = {f: getattr(r.times, f) for f, _ in CTimes._fields_}
    d["peel_split_ms"] = dict(zip(("barrier", "scan", "process", "flush"),
                                  list(r.peel_split_ms)))
    if with_json:
        d["json"] = result_json(r)
    return d


class Engine:
    """One engine per GPU (cds_engine): owns a HIP stream and device memory."""

    def __init__(self, device=0, hbm_budget_bytes=0, profile_kernels=False,
                 tune_cache=None):
        # tune_cache None: the file CDS_TUNE_CACHE names, else the shipped
        # cache (cds_amd/tune_cache.txt) if present, the order cds_main uses;
        # "": none (the CDS_TUNE_CACHE environment variable still applies)
        if tune_cache is None and not os.environ.get("CDS_TUNE_CACHE") and \
                os.path.exists(TUNE_CACHE):
            tune_cache = TUNE_CACHE
        opts = COpts(device, hbm_budget_bytes, 1 if profile_kernels else 0,
                     os.fsencode(tune_cache) if tune_cache else None)
        h = ctypes.c_void_p()
        _check(lib().cds_engine_create(ctypes.byref(opts), ctypes.byref(h)))
        self._h = h

    def set_tuning(self, tuning=None):
        """Default launch shapes (dict of TUNING_FIELDS; None: built-in rules)."""
        t = CTuning.from_dict(tuning) if tuning is not None else None
        _check(lib().cds_engine_set_tuning(self._h, ctypes.byref(t) if t else None))

    def get_tuning(self, k, m):
        t = CTuning()
        _check(lib().cds_engine_get_tuning(self._h, k, m, ctypes.byref(t)))
        return t.as_dict()

    def load_tuning(self, path):
        n = ctypes.c_uint32()
        _check(lib().cds_engine_load_tuning(self._h, os.fsencode(path), ctypes.byref(n)))
        return n.value

    def autotune(self, dgraph, k, path=None):
        """cds_autotune on one graph; returns (tuning dict, best ms, first ms)."""
        p = CParams(k, 8000, 16000, 60000, 30000, -1, 1, 1, 0)
        t = CTuning()
        best, base = ctypes.c_double(), ctypes.c_double()
        _check(lib().cds_autotune(self._h, dgraph._h, ctypes.byref(p),
                                  os.fsencode(path) if path else None, ctypes.byref(t),
                                  ctypes.byref(best), ctypes.byref(base)))
        return t.as_dict(), best.value, base.value

    def upload(self, graph):
        return DeviceGraph(self, graph)

    def solve(self, dgraph, k, p_size=8000, cp_size=16000, gl_buffer_size=60000,
              partition_size=30000, early_stop=-1, ub1=True, ub2=True,
              vertex_arrays=False, with_json=False):
        p = CParams(k, p_size, cp_size, gl_buffer_size, partition_size,
                    early_stop, int(ub1), int(ub2), int(vertex_arrays))
        r = CResult()
        _check(lib().cds_solve(self._h, dgraph._h, ctypes.byref(p), ctypes.byref(r)))
        try:
            return result_to_dict(r, with_json)
        finally:
            lib().cds_result_free(ctypes.byref(r))

    def solve_batch(self, dgraphs, ks, vertex_arrays=False):
        """cds_solve_batch: every graph solved with its own k (default
        parameters, ub1 = ub2 = on); returns one result dict per graph."""
        n = len(dgraphs)
        if len(ks) != n:
            raise ValueError("one k per graph")
        if n == 0:
            return []
        hs = (ctypes.c_void_p * n)(*[d._h for d in dgraphs])
        ps = (CParams * n)(*[CParams(int(k), 8000, 16000, 60000, 30000, -1, 1, 1,
                                     int(vertex_arrays)) for k in ks])
        rs = (CResult * n)()
        _check(lib().cds_solve_batch(self._h, hs, ps, n, rs))
        try:
            return [result_to_dict(r) for r in rs]
        finally:
            for r in rs:
                lib().cds_result_free(ctypes.byref(r))

    def core_decompose(self, dgraph, k, vertex_arrays=False):
        """Listing + clique-core peel + densest core only (no flow)."""
        p = CParams(k, 8000, 16000, 60000, 30000, -1, 1, 1, int(vertex_arrays))
        r = CResult()
        _check(lib().cds_core_decompose(self._h, dgraph._h, ctypes.byref(p),
                                        ctypes.byref(r)))
        try:
            return result_to_dict(r)
        finally:
            lib().cds_result_free(ctypes.byref(r))

    def count_cliques(self, dgraph, k, clique_degree=True):
        T = ctypes.c_uint64()
        ms = ctypes.c_double()
        cd = np.zeros(dgraph.n, dtype=np.uint64) if clique_degree else None
        _check(lib().cds_count_cliques(
            self._h, dgraph._h, k, ctypes.byref(T),
            cd.ctypes.data_as(u64p) if cd is not None else None, ctypes.byref(ms)))
        return T.value, cd, ms.value

    def count_cliques_noalloc(self, dgraph, k, clique_degree=True):
        """cds_count_cliques_noalloc: count_cliques with no device allocation
        (scratch in the library's static device array)."""
        T = ctypes.c_uint64()
        ms = ctypes.c_double()
        cd = np.zeros(dgraph.n, dtype=np.uint64) if clique_degree else None
        _check(lib().cds_count_cliques_noalloc(
            self._h, dgraph._h, k, ctypes.byref(T),
            cd.ctypes.data_as(u64p) if cd is not None else None, ctypes.byref(ms)))
        return T.value, cd, ms.value

    def close(self):
        """Destroys the engine (its stream and device memory).  Call it
        before the interpreter exits: at shutdown __del__ makes no HIP call,
        the runtime and any profiler attached to it may already be gone."""
        if getattr(self, "_h", None):
            lib().cds_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if not sys.is_finalizing():
            self.close()


class DeviceGraph:
    """A graph resident in HBM (cds_dgraph)."""

    def __init__(self, engine, graph):
        h = ctypes.c_void_p()
        _check(lib().cds_graph_upload(engine._h, ctypes.byref(graph._g), ctypes.byref(h)))
        self._h = h
        self._engine = engine
        self.n = graph.n
        self.m = graph.m

    def free(self):
        """Releases the graph's device memory (its engine must be open)."""
        if getattr(self, "_h", None) and self._engine._h:
            lib().cds_dgraph_free(self._engine._h, self._h)
        self._h = None

    def __del__(self):
        if not sys.is_finalizing():
            self.free()


class Server:
    """cds_server: a request loop over one engine with HBM-budget admission
    and continuous batching.  submit() is thread-safe; wait() returns the
    result dict with the request's queue_ms / service_ms."""

    def __init__(self, engine, hbm_budget_bytes=0, max_batch=1024, max_wait_us=0,
                 large_vertices=65536):
        self.engine = engine
        o = CServerOpts(int(hbm_budget_bytes), int(max_batch), int(max_wait_us),
                        int(large_vertices))
        h = ctypes.c_void_p()
        _check(lib().cds_server_create(engine._h, ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def submit(self, graph, k, ub1=True, ub2=True, vertex_arrays=False):
        p = CParams(int(k), 8000, 16000, 60000, 30000, -1, int(ub1), int(ub2),
                    int(vertex_arrays))
        t = ctypes.c_uint64()
        _check(lib().cds_submit(self._h, ctypes.byref(graph._g), ctypes.byref(p),
                                ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        d = ctypes.c_int()
        _check(lib().cds_poll(self._h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket, timeout_s=None):
        r = CResult()
        tm = CRequestTimes()
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        _check(lib().cds_wait(self._h, ticket, us, ctypes.byref(r), ctypes.byref(tm)))
        try:
            d = result_to_dict(r)
        finally:
            lib().cds_result_free(ctypes.byref(r))
        d["queue_ms"] = tm.queue_ms
        d["service_ms"] = tm.service_ms
        return d

    def cancel(self, ticket):
        _check(lib().cds_cancel(self._h, ticket))

    def stats(self):
        st = CServerStats()
        _check(lib().cds_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in CServerStats._fields_}

    def close(self):
        if self._h:
            lib().cds_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
