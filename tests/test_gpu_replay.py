"""GPU parity of the bitset replay: a level with few vertices alive at its
start (k = 2, 3) is replayed by one workgroup on adjacency bitsets instead of
the incidence lists.  CDS_REPLAY_BITS=1 sends every level it can take there,
0 none; the answers must match the oracle and each other bit for bit."""
import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_gridflow import _engine_with, _with_env
from test_gpu_parity import SCALARS, _announce, compare

pytestmark = pytest.mark.gpu


def _graph(seed):
    rng = np.random.default_rng(seed + 700)
    n = int(rng.integers(20, 600))
    return cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                  float(rng.uniform(3, 14)), int(rng.integers(0, 40)),
                                  0.85, seed=seed + 700)


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(12) for k in (2, 3)])
def test_bitset_replay_small(engine, seed, k):
    g = _graph(seed)
    label = f"bitset replay chung_lu seed={seed} n={g.n} k={k}"
    _announce(label)
    got = _with_env({"CDS_REPLAY_BITS": "1"},
                    lambda: engine.solve(engine.upload(g), k, vertex_arrays=True))
    compare(got, oracle_solve(g, k), label)
    assert got["replay_bits_phases"] == got["replay_phases"], label


def test_bitset_replay_k4_not_taken(engine):
    g = _graph(3)
    got = _with_env({"CDS_REPLAY_BITS": "1"},
                    lambda: engine.solve(engine.upload(g), 4, vertex_arrays=True))
    compare(got, oracle_solve(g, 4), "bitset replay k=4 (lists)")
    assert got["replay_bits_phases"] == 0


@pytest.mark.parametrize("scale,k,seed", [(13, 3, 53), (14, 3, 54), (14, 2, 55), (15, 3, 56)])
def test_bitset_replay_matches_lists(engine, scale, k, seed):
    """R-MAT graphs too large for the oracle: the bitset replay (default and
    forced) and the incidence-list replays give the same answer."""
    g = cds_amd.Graph.rmat(scale, 16, seed=seed)
    dg = engine.upload(g)
    _announce(f"bitset vs list replay R-MAT scale {scale} k={k}")
    lists = _with_env({"CDS_REPLAY_BITS": "0"}, lambda: engine.solve(dg, k))
    assert lists["replay_bits_phases"] == 0
    for mode in ("", "1"):
        got = _with_env({"CDS_REPLAY_BITS": mode}, lambda: engine.solve(dg, k))
        for f in SCALARS + ["kmax", "set_size", "replay_phases"]:
            assert got[f] == lists[f], (mode, f)
        assert np.float32(got["density"]) == np.float32(lists["density"])
        assert got["vertices"] == lists["vertices"]
        if mode == "1":
            assert got["replay_bits_phases"] == got["replay_phases"]


@pytest.fixture(scope="module")
def coarse_engine():
    eng = _engine_with({"CDS_PEEL_COARSE": "1", "CDS_PEEL_COARSE_SHIFT": "1"})
    yield eng
    eng.close()


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(12) for k in (2, 3)])
def test_bitset_refine_small(coarse_engine, seed, k):
    """Coarse peel steps re-peeled on bitsets (their levels' replays taken
    from the removal sequence): every core number and the answer against
    the oracle."""
    g = _graph(seed + 50)
    label = f"bitset refine chung_lu seed={seed + 50} n={g.n} k={k}"
    _announce(label)
    got = coarse_engine.solve(coarse_engine.upload(g), k, vertex_arrays=True)
    compare(got, oracle_solve(g, k), label)
    assert got["refine_bits_steps"] == got["peel_refined_steps"], label


@pytest.mark.parametrize("scale,k,seed", [(14, 3, 61), (15, 3, 62), (14, 2, 63)])
def test_bitset_refine_matches_grid_refine(coarse_engine, scale, k, seed):
    """R-MAT graphs: the bitset re-peel and the whole-GPU re-peel of the
    coarse steps give the same answer and peel summary."""
    g = cds_amd.Graph.rmat(scale, 16, seed=seed)
    dg = coarse_engine.upload(g)
    _announce(f"bitset vs grid refine R-MAT scale {scale} k={k}")
    grid = _with_env({"CDS_REPLAY_BITS": "0"}, lambda: coarse_engine.solve(dg, k))
    bits = coarse_engine.solve(dg, k)
    assert grid["refine_bits_steps"] == 0
    assert bits["refine_bits_steps"] > 0
    for f in SCALARS + ["kmax", "set_size", "peel_distinct_levels"]:
        assert bits[f] == grid[f], f
    assert np.float32(bits["density"]) == np.float32(grid["density"])
    assert bits["vertices"] == grid["vertices"]
