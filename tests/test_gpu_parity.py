"""GPU parity: the HIP engine (through the C ABI) against the oracle.

The oracle is the restatement of the reference CPU pipeline with its
max-flow evaluated exactly (FLOW_EXACT); the engine must match it
bit-exactly on every reported quantity: clique count and per-vertex clique
degrees, clique-core numbers, k', kmax, the densest core, pruning,
components, the float bounds and alpha sequence, and the final vertex set
with its float density and exact clique fraction.
"""
import math
import os
import sys

import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve

pytestmark = pytest.mark.gpu

SCALARS = ["total_cliques", "k_prime", "kmax", "core_size", "core_edges",
           "core_cliques", "deleted_cliques", "pruned_edges", "num_components",
           "size", "set_cliques"]
FLOATS = ["core_density", "paper_lower_bound", "density"]


def _same_float(a, b):
    if isinstance(a, float) and math.isnan(a):
        return isinstance(b, float) and math.isnan(b)
    return float(a) == float(b)


def compare(got, want, label):
    errs = []
    for key in SCALARS:
        if int(got[key]) != int(want[key]):
            errs.append(f"{key}: {got[key]} != {want[key]}")
    for key in FLOATS:
        if not _same_float(got[key], want[key]):
            errs.append(f"{key}: {got[key]!r} != {want[key]!r}")
    for gk, wk in (("ub1_val", "ub1"), ("ub2_val", "ub2")):
        if not _same_float(got[gk], want[wk]):
            errs.append(f"{gk}: {got[gk]!r} != {want[wk]!r}")
    if list(got["vertices"]) != list(want["vertices"]):
        errs.append(f"vertices differ ({len(got['vertices'])} vs {len(want['vertices'])})")
    if got["clique_degree"] is not None:
        if not np.array_equal(got["clique_degree"].astype(np.int64),
                              np.asarray(want["clique_degree"], dtype=np.int64)):
            errs.append("clique_degree differs")
        if not np.array_equal(got["core"].astype(np.int64),
                              np.asarray(want["core"], dtype=np.int64)):
            errs.append("core numbers differ")
    if got["trace_alpha"] != [np.float32(a) for a in want["trace_alpha"]]:
        errs.append(f"alpha trace differs: {got['trace_alpha'][:6]} vs {want['trace_alpha'][:6]}")
    assert not errs, f"{label}: " + "; ".join(errs)


def _announce(label):
    print(f"[case] {label}", flush=True)
    sys.stdout.flush()


SMALL = [(seed, k) for seed in range(12) for k in (2, 3, 4)]


@pytest.mark.parametrize("seed,k", SMALL)
def test_small_chung_lu(engine, seed, k):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                               float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                               0.8, seed=seed)
    label = f"chung_lu seed={seed} n={n} k={k}"
    _announce(label)
    got = engine.solve(engine.upload(g), k, vertex_arrays=True)
    compare(got, oracle_solve(g, k), label)


def test_reference_data_graph(engine):
    """cudaCode/data/test.txt: K4 on {6,7,8,9} plus a path; T3=4, T4=1."""
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_data_test.txt")
    g = cds_amd.Graph.load(path)
    for k in (2, 3, 4):
        _announce(f"ref_data_test k={k}")
        got = engine.solve(engine.upload(g), k, vertex_arrays=True)
        compare(got, oracle_solve(g, k), f"ref_data_test k={k}")


def test_c1_chung_lu_10k(engine):
    """The CPU-runnable headline configuration: Chung-Lu n=10,000, exponent
    2.5, mean degree 10, planted G(40, 0.9), k=3."""
    g = cds_amd.Graph.chung_lu(10000, 2.5, 10.0, 40, 0.9, seed=1)
    _announce("C1 k=3")
    got = engine.solve(engine.upload(g), 3, vertex_arrays=True)
    compare(got, oracle_solve(g, 3), "C1 k=3")
