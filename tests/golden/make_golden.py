"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference sources; oracle/_ref/ must be built: `make -C oracle ref`).

Inputs: the reference's own data graph (cudaCode/data/test.txt, copied as
ref_data_test.txt), the 10-vertex graph logged in
cpuCode/c++/build/testnew.log:1-12 (ref_log_graph.txt), the example of
notes.ipynb cell 15 (notes_example.txt), and small seeded Chung-Lu / R-MAT
graphs written by the engine's generators (graphs/*.txt).

Outputs: reference_outputs.json = the reference's own pipeline
(oracle/_ref/dsd_harness, i.e. cpuCode/c++/src/*.cpp unmodified) on each
graph for k = 2, 3, 4: clique count and degrees, core numbers, densest core,
pruning, components, bounds, final density, size and vertex set.

reference_fragile.json = the same on float-fragile graphs (R-MAT scale 8,
edge factor 16, seeds 7, 8 and 16, k = 3), where the reference's float
Edmonds-Karp (cpuCode/c++/src/helpers.cpp:1343-1405) decides an alpha step
the other way than exact arithmetic and prints another densest subgraph.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import cds_amd  # noqa: E402
from oracle_ref import reference_solve  # noqa: E402

KEYS = ["total_cliques", "clique_degree", "core", "k_prime", "kmax",
        "core_density", "core_size", "core_edges", "core_cliques",
        "deleted_cliques", "pruned_edges", "num_components", "component_size",
        "component_cliques", "paper_lower_bound", "ub1", "ub2", "density",
        "size", "vertices"]


def seeded_graphs():
    out = []
    for seed in range(16):
        rng = np.random.default_rng(1000 + seed)
        n = int(rng.integers(20, 150))
        g = cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                   float(rng.uniform(3, 10)),
                                   int(rng.integers(0, 18)), 0.85, seed=1000 + seed)
        out.append((f"cl_{seed:02d}", g, {"kind": "chung_lu", "seed": 1000 + seed}))
    for seed in range(4):
        g = cds_amd.Graph.rmat(7, 8, seed=2000 + seed)
        out.append((f"rmat7_{seed}", g, {"kind": "rmat", "scale": 7, "seed": 2000 + seed}))
    return out


FRAGILE = [(7, 3), (8, 3), (16, 3)]


def fragile_graphs():
    return [(f"rmat8_{seed}", cds_amd.Graph.rmat(8, 16, seed=seed), k)
            for seed, k in FRAGILE]


def main():
    fragile = {}
    os.makedirs(os.path.join(HERE, "graphs"), exist_ok=True)
    for name, g, k in fragile_graphs():
        path = os.path.join(HERE, "graphs", name + ".txt")
        g.write(path)
        ref = reference_solve(path, k)
        fragile[f"{name}/k{k}"] = {"graph": os.path.relpath(path, HERE), "k": k,
                                   **{key: ref[key] for key in KEYS}}
    with open(os.path.join(HERE, "reference_fragile.json"), "w") as f:
        json.dump(fragile, f, indent=0, sort_keys=True, allow_nan=True)
    fixtures = {}
    files = [("ref_data_test", os.path.join(HERE, "ref_data_test.txt")),
             ("ref_log_graph", os.path.join(HERE, "ref_log_graph.txt")),
             ("notes_example", os.path.join(HERE, "notes_example.txt"))]
    os.makedirs(os.path.join(HERE, "graphs"), exist_ok=True)
    for name, g, meta in seeded_graphs():
        path = os.path.join(HERE, "graphs", name + ".txt")
        g.write(path)
        files.append((name, path))
    for name, path in files:
        for k in (2, 3, 4):
            ref = reference_solve(path, k)
            fixtures[f"{name}/k{k}"] = {
                "graph": os.path.relpath(path, HERE),
                "k": k,
                **{key: ref[key] for key in KEYS},
            }
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(fixtures, f, indent=0, sort_keys=True, allow_nan=True)
    print(f"wrote {len(fixtures)} fixtures")


if __name__ == "__main__":
    main()
