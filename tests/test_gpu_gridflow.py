"""GPU parity of the whole-GPU kernels and of R-MAT inputs.

The max flow runs on as many workgroups as a component's size needs, one
for the small graphs of the oracle comparison; CDS_FLOW_BLOCKS=3 forces
three, so the grid barriers and cross-workgroup pushes of that kernel are
checked on every component (including components that need one flow per
alpha query).  Likewise CDS_REPLAY_GRID_WORK=0 sends every exact phase
replay with at most 16384 vertices to the whole-GPU replay kernel.
"""
import math
import os

import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu


def _engine_with(env):
    """An engine created while `env` is set (the engine reads it once)."""
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return cds_amd.Engine(device=0)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def grid_engine():
    eng = _engine_with({"CDS_FLOW_BLOCKS": "3"})
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def replay_engine():
    eng = _engine_with({"CDS_REPLAY_GRID_WORK": "0"})
    yield eng
    eng.close()


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(8) for k in (3, 4)])
def test_grid_flow_small(grid_engine, seed, k):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                               float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                               0.8, seed=seed)
    label = f"grid flow chung_lu seed={seed} n={n} k={k}"
    _announce(label)
    got = grid_engine.solve(grid_engine.upload(g), k, vertex_arrays=True)
    compare(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("seed", [10, 11, 12, 17])
def test_grid_flow_non_certified(grid_engine, seed):
    """Graphs whose densest subgraph is a proper subset of its component, so
    every alpha query below the optimum needs its own exact flow and
    minimal min-cut side."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                               float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                               0.8, seed=seed)
    for k in (3, 4):
        label = f"grid flow seed={seed} k={k}"
        _announce(label)
        got = grid_engine.solve(grid_engine.upload(g), k, vertex_arrays=True)
        compare(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("scale", [9, 10, 11])
def test_rmat_parity(engine, scale):
    g = cds_amd.Graph.rmat(scale, 16, seed=2)
    label = f"rmat scale={scale} k=3"
    _announce(label)
    got = engine.solve(engine.upload(g), 3, vertex_arrays=True)
    compare(got, oracle_solve(g, 3), label)


_REPLAYED = []


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(12) for k in (3, 4)])
def test_grid_replay_small(replay_engine, seed, k):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 400))
    g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                               float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                               0.8, seed=100 + seed)
    label = f"grid replay chung_lu seed={100 + seed} n={n} k={k}"
    _announce(label)
    # the bitset replay would take the k = 3 levels: off here
    got = _with_env({"CDS_REPLAY_BITS": "0"},
                    lambda: replay_engine.solve(replay_engine.upload(g), k, vertex_arrays=True))
    compare(got, oracle_solve(g, k), label)
    assert got["replay_grid_phases"] == got["replay_phases"], label
    _REPLAYED.append(got["replay_grid_phases"])


@pytest.mark.parametrize("scale", [9, 10])
def test_grid_replay_rmat(replay_engine, scale):
    g = cds_amd.Graph.rmat(scale, 16, seed=3)
    label = f"grid replay rmat scale={scale} k=3"
    _announce(label)
    got = _with_env({"CDS_REPLAY_BITS": "0"},
                    lambda: replay_engine.solve(replay_engine.upload(g), 3, vertex_arrays=True))
    compare(got, oracle_solve(g, 3), label)
    _REPLAYED.append(got["replay_grid_phases"])


def test_grid_replay_was_exercised():
    assert sum(_REPLAYED) > 0, "no phase went through the whole-GPU replay"


def test_batched_flows_on_several_workgroups(grid_engine):
    """cds_solve_batch with three workgroups per max flow: the cooperative
    batched launch and its per-problem barriers, against the oracle."""
    rng = np.random.default_rng(4242)
    graphs, ks = [], []
    for i in range(10):
        n = int(rng.integers(40, 400))
        graphs.append(cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                             float(rng.uniform(4, 12)),
                                             int(rng.integers(8, 30)), 0.85,
                                             seed=4242 + i))
        ks.append(3 if i % 3 else 4)
    _announce("batched flows on three workgroups each")
    got = grid_engine.solve_batch([grid_engine.upload(g) for g in graphs], ks,
                                  vertex_arrays=True)
    for i, (g, k, r) in enumerate(zip(graphs, ks, got)):
        compare(r, oracle_solve(g, k), f"batched multi-workgroup flow item {i}")


def _with_env(env, fn):
    """Runs fn() with the variables of `env` set (solves read them)."""
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _with_wide(fn):
    """Runs fn() with CDS_FLOW_WIDE=1: every max flow on clique-id incidence
    (the member's position found in the clique's record), which otherwise
    serves only networks of 2^32 clique-member slots or more."""
    return _with_env({"CDS_FLOW_WIDE": "1"}, fn)


@pytest.mark.parametrize("blocks", ["one", "three"])
@pytest.mark.parametrize("seed,k", [(s, k) for s in range(6) for k in (3, 4)])
def test_wide_flow_small(engine, grid_engine, seed, k, blocks):
    eng = engine if blocks == "one" else grid_engine
    rng = np.random.default_rng(seed + 40)
    n = int(rng.integers(20, 600))
    g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                               float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                               0.8, seed=seed + 40)
    label = f"wide flow ({blocks} workgroups) chung_lu seed={seed} n={n} k={k}"
    _announce(label)
    got = _with_wide(lambda: eng.solve(eng.upload(g), k, vertex_arrays=True))
    compare(got, oracle_solve(g, k), label)


def _small_graph(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 500))
    return cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                  float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                                  0.8, seed=seed), n


@pytest.mark.parametrize("blocks", ["one", "three"])
@pytest.mark.parametrize("seed,k", [(s, k) for s in range(6) for k in (2, 3, 4)])
def test_implicit_flow_small(engine, grid_engine, seed, k, blocks):
    """The implicit-clique max flow (the default) on one and on three
    workgroups, k = 2 (16-byte records) to 4, against the oracle; every
    whole-network flow of the solve ran it."""
    eng = engine if blocks == "one" else grid_engine
    g, n = _small_graph(seed + 70)
    label = f"implicit flow ({blocks} workgroups) chung_lu seed={seed + 70} n={n} k={k}"
    _announce(label)
    got = eng.solve(eng.upload(g), k, vertex_arrays=True)
    compare(got, oracle_solve(g, k), label)
    assert got["flow_implicit"] == got["flow_launches"], label


@pytest.mark.parametrize("blocks,wide", [("one", "0"), ("three", "0"), ("three", "1")])
@pytest.mark.parametrize("seed,k", [(s, k) for s in range(6) for k in (3, 4)])
def test_explicit_flow_small(engine, grid_engine, seed, k, blocks, wide):
    """CDS_FLOW_IMPLICIT=0: the explicit clique network (clique heights,
    excess and work lists), which serves flows whose unit is too large for
    32-bit residuals, against the oracle; wide = 1 runs its clique-id
    incidence variant."""
    eng = engine if blocks == "one" else grid_engine
    g, n = _small_graph(seed + 90)
    label = (f"explicit flow ({blocks} workgroups, wide={wide}) chung_lu "
             f"seed={seed + 90} n={n} k={k}")
    _announce(label)
    got = _with_env({"CDS_FLOW_IMPLICIT": "0", "CDS_FLOW_WIDE": wide},
                    lambda: eng.solve(eng.upload(g), k, vertex_arrays=True))
    compare(got, oracle_solve(g, k), label)
    assert got["flow_implicit"] == 0, label


def _served_mix(count, seed, first):
    """Graphs of the batched workload's generator (Chung-Lu, n log-uniform
    in [200, 20000], exponent [2.1, 3.0], mean degree [4, 24], one planted
    G(s, 0.9), s in [8, 60]) with k drawn from {2, 3, 4} (30/50/20)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(math.exp(rng.uniform(math.log(200), math.log(20000))))
        g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)), float(rng.uniform(4, 24)),
                                   int(rng.integers(8, 61)), 0.9, seed=first + i)
        k = int(rng.choice([2, 3, 4], p=[0.3, 0.5, 0.2]))
        out.append((g, k))
    return out


def test_implicit_flow_mixed_stream(engine):
    """300 graphs of the served stream's small share: every max flow finishes
    on the implicit-clique network (no fall-back to the explicit one), the
    answers equal the explicit network's, and those of up to 3000 vertices
    equal the oracle's.  The sample holds a graph whose flow did not finish
    in 20000 rounds while an admissible arc limited by its clique's lowest
    members could still trigger a relabel."""
    _announce("implicit flow on 300 mixed-stream graphs")
    for i, (g, k) in enumerate(_served_mix(300, 5, 9000)):
        dg = engine.upload(g)
        got = engine.solve(dg, k, vertex_arrays=True)
        label = f"mixed graph {i} n={g.n} k={k}"
        assert got["flow_implicit"] == got["flow_launches"], label
        ref = _with_env({"CDS_FLOW_IMPLICIT": "0"},
                        lambda: engine.solve(dg, k, vertex_arrays=True))
        assert ref["flow_implicit"] == 0, label
        for f in ("total_cliques", "k_prime", "core_size", "size", "set_cliques",
                  "set_size", "num_components", "trace_alpha", "trace_margin", "vertices"):
            assert got[f] == ref[f], f"{label}: {f}"
        if g.n <= 3000:
            compare(got, oracle_solve(g, k), label)
        dg.free()
