"""GPU parity of the whole-GPU flow kernel and of R-MAT inputs.

The engine switches to the cooperative grid-wide push-relabel for
components above a node threshold; CDS_GRID_FLOW_NODES=0 forces it for
every component, so the small graphs of the oracle comparison exercise it
too (including components that need one flow per alpha query).
"""
import os

import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid_engine():
    old = os.environ.get("CDS_GRID_FLOW_NODES")
    os.environ["CDS_GRID_FLOW_NODES"] = "0"
    try:
        eng = cds_amd.Engine(device=0)
    finally:
        if old is None:
            del os.environ["CDS_GRID_FLOW_NODES"]
        else:
            os.environ["CDS_GRID_FLOW_NODES"] = old
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
