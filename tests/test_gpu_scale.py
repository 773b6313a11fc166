"""GPU parity at full size: the engine on graphs too large for the tests to
re-solve on the CPU, against the oracle's answers committed as fixtures
(tests/golden/scale_*.json, written by tests/golden/make_scale_fixtures.py
with the oracle's fast peel and exact flows).

Each graph is regenerated from the fixture's generator call and checked
against its CSR digest first.  Then every field is compared: T, the clique
degrees and core numbers (SHA-256 of the arrays, plus their sums), k',
kmax, the densest core (size, edges, cliques, deleted cliques), the pruned
edge count, the components, the float bounds, the alpha trace, and the
final vertex set with its float density and exact clique fraction.

rmat20_k3 is the plain-method headline graph: R-MAT scale 20, edge factor
16, (a, b, c) = (.57, .19, .19), seed 2, k = 3 (4.24e8 triangles).
"""
import glob
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import cds_amd

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "scale_*.json")))


def _load(path):
    with open(path) as f:
        return json.load(f)


def _graph(fx):
    if fx["generator"] == "rmat":
        return cds_amd.Graph.rmat(**fx["args"])
    return cds_amd.Graph.chung_lu(**fx["args"])


def _digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


@pytest.mark.parametrize("wide", [False, True], ids=["flow", "wide_flow"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[6:-5] for p in FIXTURES])
def test_full_size_against_oracle_fixture(engine, path, wide):
    """wide: the max flows on the kernel for networks of 2^32 slots or more
    (CDS_FLOW_WIDE=1), which R-MAT scale 18 edge factor 24 at k = 4 needs."""
    fx = _load(path)
    print(f"[case] {fx['name']} k={fx['k']} {'wide flow' if wide else ''}", flush=True)
    sys.stdout.flush()
    g = _graph(fx)
    off, nb = g.csr()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(nb, dtype="<u4").tobytes())
    assert (g.n, g.m, h.hexdigest()) == (fx["n"], fx["m"], fx["csr_sha256"]), \
        "generator output changed: the fixture no longer describes this graph"
    dg = engine.upload(g)
    old = os.environ.get("CDS_FLOW_WIDE")
    os.environ["CDS_FLOW_WIDE"] = "1" if wide else "0"
    try:
        got = engine.solve(dg, fx["k"], vertex_arrays=True)
    finally:
        if old is None:
            del os.environ["CDS_FLOW_WIDE"]
        else:
            os.environ["CDS_FLOW_WIDE"] = old
    dg.free()
    errs = []
    for key in ("total_cliques", "k_prime", "kmax", "core_size", "core_edges",
                "core_cliques", "deleted_cliques", "pruned_edges", "num_components",
                "size", "set_cliques"):
        if int(got[key]) != int(fx[key]):
            errs.append(f"{key}: {got[key]} != {fx[key]}")
    for gk, wk in (("core_density", "core_density"), ("paper_lower_bound",
                   "paper_lower_bound"), ("density", "density"), ("ub1_val", "ub1"),
                   ("ub2_val", "ub2")):
        if float(got[gk]) != float(fx[wk]):
            errs.append(f"{gk}: {got[gk]!r} != {fx[wk]!r}")
    cd, core = got["clique_degree"], got["core"]
    if int(cd.astype(np.uint64).sum()) != fx["clique_degree_sum"] or \
            _digest(cd, "<u8") != fx["clique_degree_sha256"]:
        errs.append("clique degrees differ")
    if int(core.astype(np.uint64).sum()) != fx["core_sum"] or \
            _digest(core, "<u4") != fx["core_sha256"]:
        errs.append("core numbers differ")
    if list(got["vertices"]) != fx["vertices"]:
        errs.append(f"vertex set differs ({len(got['vertices'])} vs {len(fx['vertices'])})")
    if got["set_size"] != fx["size"]:
        errs.append(f"set size {got['set_size']} != {fx['size']}")
    if got["trace_alpha"] != [np.float32(a) for a in fx["trace_alpha"]]:
        errs.append("alpha trace differs")
    if got["trace_component"] != fx["trace_component"]:
        errs.append("trace components differ")
    # the exact margins kT - F(alpha) of the engine's envelope and of the
    # oracle's Dinic flows at every alpha step
    if [float(x) for x in got["trace_margin"]] != [float(x) for x in fx["trace_margin"]]:
        errs.append("trace margins differ")
    assert not errs, f"{fx['name']}: " + "; ".join(errs)
