"""R-MAT scale 18, edge factor 24, plus 32 planted G(s, 0.85) communities
(s in [30, 120]), k = 4: 7.5e9 4-cliques, beyond what the oracle (or any
materialising solve) can hold, so its answer is checked by properties
instead of a fixture:

  * the clique degrees from the count pass sum to k T;
  * the solve ran on the pruned tau-clique-core (tau <= k'), peeled exactly;
  * the k-cliques inside the reported set, recounted by a separate listing
    of the set's induced subgraph, equal the reported numerator, and the
    float density is that fraction rounded;
  * the set is at least as dense as the whole graph and as the densest core
    (level-boundary densities), and k' = ceil(max peel density);
  * every component of the densest core was solved (here one network,
    certified in one max flow).
"""
import math

import numpy as np
import pytest

import cds_amd
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu


def c3_graph():
    return cds_amd.Graph.rmat(18, 24, 0.57, 0.19, 0.19, 32, 30, 120, 0.85, seed=3)


def induced(g, verts):
    off, nb = g.csr()
    off = off.astype(np.int64)
    nb = nb.astype(np.int64)
    idx = np.full(g.n, -1, dtype=np.int64)
    verts = np.asarray(verts, dtype=np.int64)
    idx[verts] = np.arange(len(verts))
    rows, cols = [], []
    for i, v in enumerate(verts):
        nbr = nb[off[v]:off[v + 1]]
        keep = idx[nbr]
        keep = keep[keep >= 0]
        rows.append(np.full(len(keep), i, dtype=np.int64))
        cols.append(keep)
    cols = np.concatenate(cols)
    cnt = np.bincount(np.concatenate(rows), minlength=len(verts))
    o2 = np.zeros(len(verts) + 1, dtype=np.uint32)
    o2[1:] = np.cumsum(cnt)
    return cds_amd.Graph.from_csr(o2, cols.astype(np.uint32))


def test_c3_rmat18_ef24_planted_k4(engine):
    _announce("C3 R-MAT 18 ef 24 + 32 planted, k=4")
    g = c3_graph()
    dg = engine.upload(g)
    T, cd, _ = engine.count_cliques(dg, 4)
    assert int(cd.astype(np.uint64).sum()) == 4 * T
    r = engine.solve(dg, 4)
    dg.free()
    assert r["total_cliques"] == T > 2 ** 32
    assert 0 < r["prune_tau"] <= r["k_prime"] and r["dense_peel"] == 1
    assert r["k_prime"] == math.ceil(r["peel_max_density"])
    assert r["core_cliques"] + r["deleted_cliques"] == T
    # the reported set, recounted from its own induced subgraph
    S = list(r["vertices"])
    assert len(S) == r["size"] == r["set_size"] and S == sorted(S)
    h = induced(g, S)
    dh = engine.upload(h)
    ts, cds, _ = engine.count_cliques(dh, 4)
    dh.free()
    assert ts == r["set_cliques"] and int(cds.astype(np.uint64).sum()) == 4 * ts
    assert np.float32(r["density"]) == np.float32(ts / len(S))
    # at least as dense as the whole graph and the densest core
    assert ts * g.n >= T * len(S)
    assert ts * r["core_size"] >= r["core_cliques"] * len(S)
    assert math.ceil(ts / len(S)) <= r["k_prime"]
    assert r["num_components"] >= 1 and r["flow_launches"] >= 1
    assert r["flow_networks_solved"] <= r["num_components"]
