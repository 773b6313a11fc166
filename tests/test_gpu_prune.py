"""The pruned path against the oracle: the solve cut to the tau-clique-core
by clique-count passes, peeled with the persistent peel or, for a dense
core, on adjacency bitsets (CDS_CORE_PRUNE=1 forces the cut,
CDS_DENSE_PEEL=1/0 the peel), must report what the oracle reports for the
whole graph: T and the clique degrees, k', kmax, the densest core, prune,
components, bounds, the alpha trace and the answer.  Core numbers are known
inside the core (equal to the oracle's there); outside it they are below
prune_tau.  The full-size fixtures (R-MAT scale 20 k=3, 18 k=3, 16 k=4)
run through both peels of the pruned path as well.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCALE = sorted(glob.glob(os.path.join(HERE, "golden", "scale_*.json")))
NONE = 0xFFFFFFFF


def _solve(engine, dg, k, dense, vertex_arrays=True, unlink=None, bits=None):
    saved = {v: os.environ.get(v) for v in ("CDS_CORE_PRUNE", "CDS_DENSE_PEEL",
                                            "CDS_PRUNE_UNLINK", "CDS_PRUNE_BITS")}
    os.environ["CDS_CORE_PRUNE"] = "1"
    os.environ["CDS_DENSE_PEEL"] = "1" if dense else "0"
    if unlink is not None:
        os.environ["CDS_PRUNE_UNLINK"] = unlink
    if bits is not None:
        os.environ["CDS_PRUNE_BITS"] = bits
    try:
        return engine.solve(dg, k, vertex_arrays=vertex_arrays)
    finally:
        for v, old in saved.items():
            if old is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = old


def _check(got, want, label):
    assert got["prune_vertices"] > 0 and got["prune_tau"] <= got["k_prime"], label
    cd, core = got["clique_degree"], got["core"].astype(np.int64)
    wcore = np.asarray(want["core"], dtype=np.int64)
    inside = core != NONE
    assert inside.sum() == got["prune_vertices"], label
    assert np.array_equal(core[inside], wcore[inside]), f"{label}: core numbers in the core"
    assert (wcore[~inside] < got["prune_tau"]).all(), f"{label}: pruned vertices"
    assert (wcore[inside] >= got["prune_tau"]).all(), f"{label}: kept vertices"
    got = dict(got, clique_degree=None)  # core numbers checked above
    compare(got, want, label)
    assert np.array_equal(cd.astype(np.int64), np.asarray(want["clique_degree"], np.int64))


GRAPHS = [("chung_lu", s, k) for s in range(6) for k in (3, 4)] + \
         [("rmat", s, k) for s in range(3) for k in (3, 4)]


@pytest.mark.parametrize("dense", [False, True], ids=["persistent", "bitset"])
@pytest.mark.parametrize("kind,seed,k", GRAPHS)
def test_pruned_path_matches_oracle(engine, kind, seed, k, dense):
    if kind == "chung_lu":
        rng = np.random.default_rng(seed + 300)
        n = int(rng.integers(200, 3000))
        g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 2.8)), float(rng.uniform(6, 14)),
                                   int(rng.integers(10, 40)), 0.85, seed=seed + 300)
    else:
        g = cds_amd.Graph.rmat(10, 16, seed=seed + 70)
    label = f"pruned {kind} seed={seed} k={k} {'bitset' if dense else 'persistent'}"
    _announce(label)
    dg = engine.upload(g)
    got = _solve(engine, dg, k, dense)
    dg.free()
    assert got["dense_peel"] == int(dense)
    _check(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("kind,seed,k", GRAPHS)
def test_pruned_path_in_place_drops(engine, kind, seed, k):
    """Every drop after the first count pass done in place (the cliques
    through the dropped vertices enumerated from their neighbourhoods and
    taken from the counts of the rest): the same answer as the oracle, and
    the solve's own count of G[S] equals the kept count (checked inside)."""
    if kind == "chung_lu":
        rng = np.random.default_rng(seed + 300)
        n = int(rng.integers(200, 3000))
        g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 2.8)), float(rng.uniform(6, 14)),
                                   int(rng.integers(10, 40)), 0.85, seed=seed + 300)
    else:
        g = cds_amd.Graph.rmat(10, 16, seed=seed + 70)
    label = f"pruned in place {kind} seed={seed} k={k}"
    _announce(label)
    dg = engine.upload(g)
    got = _solve(engine, dg, k, False, unlink="1")
    dg.free()
    if got["max_degree"] <= 2048:
        assert got["prune_passes"] <= 1, label
    _check(got, oracle_solve(g, k), label)


@pytest.mark.parametrize("scale,ef,k,seed", [(14, 16, 3, 81), (14, 16, 4, 82), (15, 12, 4, 83),
                                              (13, 24, 4, 84)])
def test_in_place_drops_match_count_passes(engine, scale, ef, k, seed):
    """R-MAT graphs beyond the oracle: drops in place (forced, default),
    subgraphs counted on bitsets, and a listing count pass after every drop
    give the same pruned core and answer."""
    g = cds_amd.Graph.rmat(scale, ef, seed=seed)
    dg = engine.upload(g)
    _announce(f"in-place drops vs count passes R-MAT scale {scale} ef {ef} k={k}")
    base = _solve(engine, dg, k, False, vertex_arrays=False, unlink="0", bits="0")
    assert base["prune_unlinks"] == 0
    for mode, bits in (("1", "0"), ("0", None), ("1", None), (None, None)):
        got = _solve(engine, dg, k, False, vertex_arrays=False, unlink=mode, bits=bits)
        for f in ("prune_tau", "prune_vertices", "kmax", "set_size", "total_cliques",
                  "k_prime", "core_size", "core_edges", "core_cliques", "deleted_cliques",
                  "pruned_edges", "num_components", "size", "set_cliques"):
            assert got[f] == base[f], (mode, f, got[f], base[f])
        if mode == "0":
            assert got["prune_passes"] == base["prune_passes"]
        assert got["prune_passes"] + got["prune_unlinks"] >= base["prune_passes"]
        if mode == "1" and base["prune_passes"] > 1 and got["max_degree"] <= 2048:
            assert got["prune_unlinks"] > 0
        assert np.float32(got["density"]) == np.float32(base["density"])
        assert got["vertices"] == base["vertices"]
    dg.free()


@pytest.mark.parametrize("dense", [False, True], ids=["persistent", "bitset"])
@pytest.mark.parametrize("path", SCALE, ids=[os.path.basename(p)[6:-5] for p in SCALE])
def test_pruned_path_on_full_size_fixtures(engine, path, dense):
    with open(path) as f:
        fx = json.load(f)
    _announce(f"pruned {fx['name']} {'bitset' if dense else 'persistent'}")
    g = cds_amd.Graph.rmat(**fx["args"])
    dg = engine.upload(g)
    got = _solve(engine, dg, fx["k"], dense)
    dg.free()
    errs = []
    for key in ("total_cliques", "k_prime", "kmax", "core_size", "core_edges", "core_cliques",
                "deleted_cliques", "pruned_edges", "num_components", "size", "set_cliques"):
        if int(got[key]) != int(fx[key]):
            errs.append(f"{key}: {got[key]} != {fx[key]}")
    for gk, wk in (("core_density", "core_density"), ("density", "density"),
                   ("ub1_val", "ub1"), ("ub2_val", "ub2"),
                   ("paper_lower_bound", "paper_lower_bound")):
        if float(got[gk]) != float(fx[wk]):
            errs.append(f"{gk}: {got[gk]!r} != {fx[wk]!r}")
    if list(got["vertices"]) != fx["vertices"]:
        errs.append("vertex set differs")
    if got["trace_alpha"] != [np.float32(a) for a in fx["trace_alpha"]]:
        errs.append("alpha trace differs")
    cd = got["clique_degree"]
    if hashlib.sha256(np.ascontiguousarray(cd, "<u8").tobytes()).hexdigest() != \
            fx["clique_degree_sha256"]:
        errs.append("clique degrees differ")
    core = got["core"].astype(np.int64)
    inside = core != NONE
    if inside.sum() != got["prune_vertices"] or (core[inside] < got["prune_tau"]).any():
        errs.append("core numbers inside the pruned core")
    assert not errs, f"{fx['name']}: " + "; ".join(errs)
    assert got["dense_peel"] == int(dense) and got["prune_tau"] <= got["k_prime"]
