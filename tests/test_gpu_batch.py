"""GPU parity of the batched entry point (cds_solve_batch).

Every batched answer must be bit-identical to the plain one-graph path
(cds_solve) on the same graph, and a sample is checked against the oracle.
Graphs follow the batched workload's generator (Chung-Lu with a planted
G(s, 0.9) community, exponent in [2.1, 3.0], mean degree in [4, 24]) at
sizes the oracle finishes quickly.
"""
import math

import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu

FIELDS = ("total_cliques", "k_prime", "kmax", "core_size", "core_cliques",
          "num_components", "density", "size", "set_cliques", "set_size",
          "peel_max_density", "core_density", "ub1_val", "ub2_val")


def mixed_graphs(count, seed, nmax=3000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(math.exp(rng.uniform(math.log(200), math.log(nmax))))
        g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                   float(rng.uniform(4, 24)),
                                   int(rng.integers(8, 61)), 0.9,
                                   seed=seed * 1000 + i)
        k = int(rng.choice([2, 3, 4], p=[0.2, 0.5, 0.3]))
        out.append((g, k))
    return out


def same(a, b):
    for f in FIELDS:
        if f in ("density", "ub1_val", "ub2_val", "peel_max_density", "core_density"):
            if not (np.float64(a[f]).tobytes() == np.float64(b[f]).tobytes()):
                return f
        elif a[f] != b[f]:
            return f
    if list(a["vertices"]) != list(b["vertices"]):
        return "vertices"
    if a["trace_alpha"] != b["trace_alpha"]:
        return "trace_alpha"
    return None


def test_batch_equals_plain(engine):
    items = mixed_graphs(40, seed=5)
    dgs = [engine.upload(g) for g, _ in items]
    ks = [k for _, k in items]
    _announce(f"batch of {len(items)} graphs")
    batch = engine.solve_batch(dgs, ks)
    for i, ((g, k), dg, rb) in enumerate(zip(items, dgs, batch)):
        rp = engine.solve(dg, k)
        diff = same(rb, rp)
        assert diff is None, f"graph {i} (n={g.n}, k={k}): batched {diff} differs"


def test_batch_matches_oracle(engine):
    items = mixed_graphs(8, seed=9, nmax=800)
    dgs = [engine.upload(g) for g, _ in items]
    ks = [k for _, k in items]
    _announce(f"batch of {len(items)} graphs vs oracle")
    batch = engine.solve_batch(dgs, ks, vertex_arrays=True)
    for i, ((g, k), got) in enumerate(zip(items, batch)):
        compare(got, oracle_solve(g, k), f"batch item {i} n={g.n} k={k}")


def test_batch_error_reports_item(engine):
    items = mixed_graphs(4, seed=13, nmax=400)
    dgs = [engine.upload(g) for g, _ in items]
    ks = [3, 3, 9, 3]  # k = 9 is outside the device lister
    _announce("batch with one invalid item")
    with pytest.raises(cds_amd.CdsError) as ei:
        engine.solve_batch(dgs, ks)
    assert "batch item 2" in str(ei.value)
    # the engine stays usable
    again = engine.solve_batch(dgs, [3, 3, 3, 3])
    assert len(again) == 4


def test_batch_empty(engine):
    assert engine.solve_batch([], []) == []


@pytest.mark.parametrize("count,kk", [(120, 2), (300, 2), (40, 4)])
def test_batch_level_runs(engine, count, kk):
    """The per-graph level tables come from a radix sort of (graph, core)
    keys over the bits that vary, 8 at a time: 120 small k = 2 graphs need
    one pass over the core bits and one over the graph bits (an even count,
    which starts from a copy), 300 graphs a second graph pass (odd), and
    the k = 4 graphs deeper cores.  Every answer equals the plain path's."""
    rng = np.random.default_rng(count + kk)
    items = []
    for i in range(count):
        n = int(rng.integers(30, 160))
        items.append(cds_amd.Graph.chung_lu(n, float(rng.uniform(2.2, 3.0)),
                                            float(rng.uniform(3, 8)),
                                            int(rng.integers(6, 14)), 0.9,
                                            seed=77000 + 1000 * kk + i))
    dgs = [engine.upload(g) for g in items]
    _announce(f"batch of {count} graphs, k={kk}: level-run sort")
    batch = engine.solve_batch(dgs, [kk] * count)
    for i, (g, dg, rb) in enumerate(zip(items, dgs, batch)):
        rp = engine.solve(dg, kk)
        diff = same(rb, rp)
        assert diff is None, f"graph {i} (n={g.n}, k={kk}): batched {diff} differs"
    for dg in dgs:
        dg.free()


def test_batch_flow_steps_as_graph_launches(engine):
    """The batched search launches each step's single-workgroup max flows as
    one HIP graph (built on the first step, re-aimed at every later step's
    items through the executable graph's node parameters).  Its answers
    equal the eager launches' (CDS_BATCH_GRAPH=0) and the plain path's over
    the steps of two calls in a row, with the engine's other work in
    between."""
    import os

    items = mixed_graphs(24, seed=17, nmax=2000)
    dgs = [engine.upload(g) for g, _ in items]
    ks = [k for _, k in items]
    _announce(f"batch of {len(items)} graphs: flow steps as graph launches")
    first = engine.solve_batch(dgs, ks)
    second = engine.solve_batch(dgs[::-1], ks[::-1])[::-1]
    os.environ["CDS_BATCH_GRAPH"] = "0"
    try:
        eager = engine.solve_batch(dgs, ks)
    finally:
        del os.environ["CDS_BATCH_GRAPH"]
    # (a call's union of one k may need no flow at all: bounds decide it)
    assert max(r["flow_graph_launches"] for r in first) > 0
    assert max(r["flow_graph_launches"] for r in eager) == 0
    for i, ((g, k), dg) in enumerate(zip(items, dgs)):
        rp = engine.solve(dg, k)
        for name, rb in (("graph", first[i]), ("graph again", second[i]), ("eager", eager[i])):
            diff = same(rb, rp)
            assert diff is None, f"graph {i} (n={g.n}, k={k}): {name} batched {diff} differs"
    for dg in dgs:
        dg.free()
