"""Test-side access to the oracle (oracle/liboracle.so, the C++ restatement
of the reference CPU pipeline, and oracle/_ref/, the reference's own sources
built by oracle/Makefile).  Only tests, smoke() and bench.py's cpu_baseline
leg use this module."""
import ctypes
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "dsd_harness")
REF_DSD = os.path.join(ROOT, "oracle", "_ref", "dsd")

PEEL_FAITHFUL, PEEL_FAST = 0, 1
FLOW_FLOAT_EK, FLOW_EXACT = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(ORACLE_LIB)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_solve_json.restype = ctypes.c_void_p
        L.oracle_solve_json.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_peel_rows_json.restype = ctypes.c_void_p
        L.oracle_peel_rows_json.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32]
        L.oracle_count_cliques.restype = ctypes.c_long
        L.oracle_count_cliques.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_long)]
        L.oracle_component_flow.restype = ctypes.c_longlong
        L.oracle_component_flow.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int),
            ctypes.POINTER(ctypes.c_long), ctypes.c_longlong, ctypes.c_longlong,
            ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def oracle_solve(graph, k, peel=PEEL_FAST, flow=FLOW_EXACT, ub1=True, ub2=True,
                 rows=False):
    """Full oracle pipeline on a cds_amd.Graph; returns the JSON dict."""
    g = graph._g
    p = lib().oracle_solve_json(g.n, g.offsets, g.neighbors, k, peel, flow,
                                int(ub1), int(ub2), int(rows))
    try:
        return json.loads(ctypes.string_at(p).decode())
    finally:
        lib().oracle_free(p)


def oracle_peel_rows(graph, k):
    """Listing + fast peel only: total_cliques, removal_vertex,
    removal_value (rows 1..n, index 0 unused) and row_remaining."""
    g = graph._g
    p = lib().oracle_peel_rows_json(g.n, g.offsets, g.neighbors, k)
    try:
        return json.loads(ctypes.string_at(p).decode())
    finally:
        lib().oracle_free(p)


def oracle_count(graph, k):
    g = graph._g
    cd = np.zeros(g.n, dtype=np.int64)
    T = lib().oracle_count_cliques(g.n, g.offsets, g.neighbors, k,
                                   cd.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
    return T, cd


def component_flow(V, cliques, k, cd, cap_t, unit):
    """Exact max flow of one vertex/clique network (oracle Dinic)."""
    T = len(cliques)
    mem = np.ascontiguousarray(np.asarray(cliques, dtype=np.int32).reshape(-1))
    cdv = np.ascontiguousarray(np.asarray(cd, dtype=np.int64))
    reach = np.zeros(V, dtype=np.int32)
    F = lib().oracle_component_flow(
        V, T, k, mem.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        cdv.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), cap_t, unit,
        reach.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return F, reach


def reference_available():
    return os.path.exists(REF_HARNESS)


def reference_solve(path, k, ub1=True, ub2=True, rows=False):
    """The reference's own pipeline (oracle/_ref/dsd_harness) on a graph file."""
    out = subprocess.run([REF_HARNESS, path, str(k), str(int(ub1)), str(int(ub2)),
                          str(int(rows))], capture_output=True, text=True,
                         timeout=600, check=True).stdout
    return json.loads(out)
