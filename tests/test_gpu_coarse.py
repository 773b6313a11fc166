"""GPU parity of the coarse-step peel: the first pass removes vertices in
steps [m, m + m >> s] and only the steps the answer depends on are re-peeled
exactly.  With every core number requested all steps are re-peeled, so the
core numbers are checked too; without, the answer must still match the
oracle and the exact-level peel."""
import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_gridflow import _engine_with
from test_gpu_parity import SCALARS, _announce, compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 3])
def coarse_engine(request):
    eng = _engine_with({"CDS_PEEL_COARSE": "1",
                        "CDS_PEEL_COARSE_SHIFT": str(request.param)})
    eng.shift = request.param
    yield eng
    eng.close()


def small_graph(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    return cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                  float(rng.uniform(3, 12)), int(rng.integers(0, 25)),
                                  0.8, seed=seed)


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(8) for k in (2, 3, 4)])
def test_coarse_peel_every_core_number(coarse_engine, seed, k):
    g = small_graph(seed)
    label = f"coarse s={coarse_engine.shift} seed={seed} n={g.n} k={k} all cores"
    _announce(label)
    got = coarse_engine.solve(coarse_engine.upload(g), k, vertex_arrays=True)
    assert got["peel_coarse_shift"] == coarse_engine.shift
    compare(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("seed,k", [(s, k) for s in range(8, 16) for k in (3, 4)])
def test_coarse_peel_answer(coarse_engine, seed, k):
    g = small_graph(seed)
    label = f"coarse s={coarse_engine.shift} seed={seed} n={g.n} k={k} answer"
    _announce(label)
    got = coarse_engine.solve(coarse_engine.upload(g), k)
    compare(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("scale,k", [(13, 3), (14, 3), (13, 4)])
def test_coarse_peel_matches_exact_levels(coarse_engine, scale, k):
    """R-MAT graphs too large for the oracle: the coarse-step peel and the
    exact-level peel must give the same answer, bit for bit."""
    exact_engine = _engine_with({"CDS_PEEL_COARSE": "0"})
    try:
        g = cds_amd.Graph.rmat(scale, 16, seed=40 + scale)
        _announce(f"coarse vs exact R-MAT scale {scale} k={k}")
        a = coarse_engine.solve(coarse_engine.upload(g), k)
        b = exact_engine.solve(exact_engine.upload(g), k)
    finally:
        exact_engine.close()
    assert a["peel_coarse_shift"] == coarse_engine.shift and b["peel_coarse_shift"] == 0
    for f in SCALARS + ["kmax", "set_size"]:
        assert a[f] == b[f], f
    assert np.float32(a["density"]) == np.float32(b["density"])
    assert a["vertices"] == b["vertices"]
    assert a["trace_alpha"] == b["trace_alpha"]
    assert a["peel_distinct_levels"] <= b["peel_distinct_levels"]
