"""Writes the full-size fixtures tests/golden/scale_<name>.json.

Each fixture is the oracle's answer (oracle/liboracle.so, FLOW_EXACT with
the fast peel: the restatement of cpuCode/c++ pinned by tests/test_oracle.py)
on one seeded synthetic graph too large for the tests to re-solve on the
CPU.  It holds the generator call (so a test regenerates the same graph), a
digest of the CSR (so the test knows it did), and every field of the answer:
T, the clique degrees and core numbers as SHA-256 digests plus their sums,
k', kmax, the densest core, prune, components, bounds, the vertex set and
its exact fraction, and the density-search trace.

    python tests/golden/make_scale_fixtures.py            # every fixture
    python tests/golden/make_scale_fixtures.py rmat20_k3  # one of them

R-MAT scale 20 at k = 3 (4.24e8 triangles) takes about ten minutes and
30 GB of host memory.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cds_amd  # noqa: E402
import oracle_ref as O  # noqa: E402

# name -> generator call and k.  rmat20_k3 is the plain-method headline
# graph: R-MAT scale 20, edge factor 16, (a, b, c) = (.57, .19, .19), seed 2.
FIXTURES = {
    "rmat20_k3": dict(gen="rmat", args=dict(scale=20, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=2), k=3),
    "rmat18_k3": dict(gen="rmat", args=dict(scale=18, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=5), k=3),
    "rmat16_k4": dict(gen="rmat", args=dict(scale=16, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=11), k=4),
}

ANSWER_FIELDS = ["total_cliques", "k_prime", "kmax", "max_density", "core_density",
                 "core_size", "core_edges", "core_cliques", "deleted_cliques",
                 "pruned_edges", "num_components", "component_size",
                 "component_cliques", "paper_lower_bound", "ub1", "ub2", "density",
                 "size", "vertices", "set_cliques", "trace_component", "trace_lb",
                 "trace_ub", "trace_alpha", "trace_equal", "trace_margin",
                 "quirk_flags"]


def make_graph(spec):
    if spec["gen"] == "rmat":
        return cds_amd.Graph.rmat(**spec["args"])
    return cds_amd.Graph.chung_lu(**spec["args"])


def csr_digest(g):
    off, nb = g.csr()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(nb, dtype="<u4").tobytes())
    return h.hexdigest()


def array_digest(values, dtype):
    return hashlib.sha256(np.ascontiguousarray(values, dtype=dtype).tobytes()).hexdigest()


def fixture(name):
    spec = FIXTURES[name]
    g = make_graph(spec)
    t0 = time.time()
    r = O.oracle_solve(g, spec["k"], peel=O.PEEL_FAST, flow=O.FLOW_EXACT)
    cd = np.asarray(r["clique_degree"], dtype=np.uint64)
    core = np.asarray(r["core"], dtype=np.uint32)
    out = dict(name=name, generator=spec["gen"], args=spec["args"], k=spec["k"],
               n=g.n, m=g.m, csr_sha256=csr_digest(g),
               clique_degree_sha256=array_digest(cd, "<u8"),
               clique_degree_sum=int(cd.sum()), clique_degree_max=int(cd.max()),
               core_sha256=array_digest(core, "<u4"), core_sum=int(core.astype(np.uint64).sum()),
               oracle_seconds=round(time.time() - t0, 1))
    for f in ANSWER_FIELDS:
        out[f] = r[f]
    return out


def main(names):
    for name in names:
        fx = fixture(name)
        path = os.path.join(HERE, f"scale_{name}.json")
        with open(path + ".tmp", "w") as f:
            json.dump(fx, f, indent=1)
            f.write("\n")
        os.replace(path + ".tmp", path)
        print(f"{name}: T={fx['total_cliques']} k'={fx['k_prime']} "
              f"set={fx['set_cliques']}/{fx['size']} in {fx['oracle_seconds']} s",
              flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(FIXTURES))
