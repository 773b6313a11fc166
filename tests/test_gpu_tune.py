"""Launch shapes and the autotuner on the GPU: every listing tile shape,
peel width, coarse-step choice, replay and flow width gives the same
answer (clique counts and clique degrees bit-exact, the oracle's final
set), and cds_autotune keeps an answer-preserving winner in the engine and
in its cache file."""
import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve

pytestmark = pytest.mark.gpu

SHAPES = [(256, 64), (256, 32), (512, 64), (512, 32), (128, 32)]
ANSWER = ("total_cliques", "k_prime", "kmax", "core_size", "core_cliques",
          "num_components", "size", "set_cliques", "set_size")


def answer(r):
    return tuple(int(r[f]) for f in ANSWER) + (float(r["density"]), tuple(r["vertices"]))


@pytest.fixture(scope="module")
def graphs():
    return {
        "chung_lu": cds_amd.Graph.chung_lu(4000, 2.3, 10.0, 30, 0.9, seed=11),
        "rmat": cds_amd.Graph.rmat(12, 12, 0.57, 0.19, 0.19, seed=5),
    }


@pytest.mark.parametrize("k", [3, 4])
@pytest.mark.parametrize("name", ["chung_lu", "rmat"])
def test_listing_shapes_count_the_same(graphs, name, k):
    eng = cds_amd.Engine(device=0, tune_cache="")
    try:
        dg = eng.upload(graphs[name])
        T0, cd0, _ = eng.count_cliques(dg, k)
        assert T0 > 0 and int(cd0.sum()) == k * T0
        for th, ch in SHAPES:
            eng.set_tuning({"list_threads": th, "list_chunk": ch})
            T, cd, _ = eng.count_cliques(dg, k)
            assert T == T0, (th, ch)
            assert np.array_equal(cd, cd0), (th, ch)
            r = eng.solve(dg, k, vertex_arrays=True)
            assert r["total_cliques"] == T0
            assert np.array_equal(r["clique_degree"], cd0), (th, ch)
    finally:
        eng.close()


def test_forced_shapes_keep_the_answer(graphs):
    g = graphs["chung_lu"]
    want = oracle_solve(g, 3)
    eng = cds_amd.Engine(device=0, tune_cache="")
    try:
        dg = eng.upload(g)
        base = eng.solve(dg, 3, vertex_arrays=True)
        assert base["vertices"] == want["vertices"] and base["k_prime"] == want["k_prime"]
        for t in ({"peel_coarse_shift": 0}, {"peel_coarse_shift": 1},
                  {"peel_coarse_shift": 3}, {"peel_blocks": 1}, {"peel_blocks": 8},
                  {"replay_blocks": 4}, {"flow_blocks": 2},
                  {"list_threads": 512, "list_chunk": 32, "peel_blocks": 4,
                   "peel_coarse_shift": 2, "flow_blocks": 1}):
            eng.set_tuning(t)
            r = eng.solve(dg, 3, vertex_arrays=True)
            assert answer(r) == answer(base), t
            assert np.array_equal(r["core"], base["core"]), t
        eng.set_tuning(None)
        assert eng.get_tuning(3, g.m)["peel_coarse_shift"] == -1
    finally:
        eng.close()


def test_autotune_writes_cache_and_keeps_answer(graphs, tmp_path):
    g = graphs["rmat"]
    path = str(tmp_path / "tune.txt")
    eng = cds_amd.Engine(device=0, tune_cache="")
    try:
        dg = eng.upload(g)
        base = eng.solve(dg, 3)
        best, best_ms, first_ms = eng.autotune(dg, 3, path)
        assert best_ms > 0 and first_ms > 0 and best_ms <= first_ms
        assert eng.get_tuning(3, g.m) == best
        assert answer(eng.solve(dg, 3)) == answer(base)
        lines = [l for l in open(path) if l.startswith("cds-tune")]
        assert len(lines) == 1
        # a second tune of the same (k, class) replaces the line
        eng.autotune(dg, 3, path)
        assert len([l for l in open(path) if l.startswith("cds-tune")]) == 1
    finally:
        eng.close()
    # a new engine loads the file at creation and uses the entry
    eng2 = cds_amd.Engine(device=0, tune_cache=path)
    try:
        got = eng2.get_tuning(3, g.m)
        arch_line = open(path).read().split("\n")[1].split()
        assert got == cds_amd.tuning_lookup(path, arch_line[2], int(arch_line[3]), 3, g.m)
        assert eng2.get_tuning(4, g.m)["peel_coarse_shift"] == -1  # no k=4 entry
        assert answer(eng2.solve(eng2.upload(g), 3)) == answer(base)
    finally:
        eng2.close()


def mixed_params(index, seed=4):
    """(n, exponent, mean degree, planted size) of graph `index` of the
    batched workload's generator sequence (numpy default_rng(seed))."""
    import math

    rng = np.random.default_rng(seed)
    for _ in range(index + 1):
        n = int(math.exp(rng.uniform(math.log(200), math.log(20000))))
        ex = float(rng.uniform(2.1, 3.0))
        md = float(rng.uniform(4, 24))
        s = int(rng.integers(8, 61))
    return n, ex, md, s


def test_coarse_step_straddling_kprime_raised_by_replay():
    """Regression: n=2009 Chung-Lu (exponent 2.18, mean degree 17.6, planted
    G(30, 0.9)), k=3, peeled in coarse steps [m, 1.25 m] on 16 workgroups.
    A replay raises k' inside a coarse step; that step must then be re-peeled
    exactly, or the densest core counts 112 vertices instead of the oracle's
    110."""
    n, ex, md, s = mixed_params(94)
    assert n == 2009
    g = cds_amd.Graph.chung_lu(n, ex, md, s, 0.9, seed=1_000_000 + 94)
    want = oracle_solve(g, 3)
    eng = cds_amd.Engine(device=0, tune_cache="")
    try:
        dg = eng.upload(g)
        for t in ({"list_threads": 512, "list_chunk": 64, "peel_blocks": 16,
                   "peel_coarse_shift": 2}, {"peel_coarse_shift": 2}, {}):
            eng.set_tuning(t)
            r = eng.solve(dg, 3)
            for f in ("core_size", "core_cliques", "k_prime", "kmax", "size",
                      "total_cliques"):
                assert r[f] == want[f], (t, f, r[f], want[f])
            assert r["vertices"] == want["vertices"], t
    finally:
        eng.close()


@pytest.fixture(scope="module")
def mixed_cases():
    cases = []
    for idx in range(80, 104, 2):
        n, ex, md, s = mixed_params(idx)
        g = cds_amd.Graph.chung_lu(n, ex, md, s, 0.9, seed=1_000_000 + idx)
        for k in (3, 4):
            cases.append((idx, k, g, oracle_solve(g, k)))
    return cases


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_coarse_shifts_match_oracle_on_mixed_graphs(mixed_cases, shift):
    """Twelve graphs of the batched workload's distribution, every one
    peeled in coarse steps [m, m + m >> shift], against the oracle."""
    eng = cds_amd.Engine(device=0, tune_cache="")
    eng.set_tuning({"peel_coarse_shift": shift, "peel_blocks": 16})
    try:
        for idx, k, g, want in mixed_cases:
            r = eng.solve(eng.upload(g), k)
            for f in ("core_size", "core_cliques", "k_prime", "kmax", "size",
                      "set_cliques", "num_components"):
                assert r[f] == want[f], (idx, k, f, r[f], want[f])
            assert r["vertices"] == want["vertices"], (idx, k)
    finally:
        eng.close()
