"""Float fragility of the density search, and the drop-in binary's printed
answer against the reference's.

The reference decides each alpha step with a float Edmonds-Karp
(cpuCode/c++/src/helpers.cpp:1279-1290); the engine decides exactly and
flags the steps whose exact margin lies within its float-drift bound of the
1e-2 threshold (cds_result.trace_fragile).  On the fragile fixtures
(tests/golden/reference_fragile.json, the reference built from its sources)
the reference prints another densest subgraph: the engine must equal the
oracle's exact mode there, and flag the step where the reference's float
search (the oracle's float mode, which reproduces the fixture) parts from
it.

cds_main's printed answer block (K:, Density:, Size:, densest subgraph:)
must equal the block oracle/_ref/dsd (cpuCode/c++ built from source)
prints for the same graph, on committed graphs where the two searches take
the same decisions; on the fragile graphs the blocks differ and the JSON
record names the fragile steps.
"""
import json
import os
import subprocess

import pytest

import cds_amd
from oracle_ref import FLOW_EXACT, FLOW_FLOAT_EK, REF_DSD, oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
with open(os.path.join(GOLDEN, "reference_fragile.json")) as f:
    FRAGILE = json.load(f)
with open(os.path.join(GOLDEN, "reference_outputs.json")) as f:
    STABLE = json.load(f)


@pytest.mark.parametrize("key", sorted(FRAGILE))
def test_engine_flags_the_step_the_reference_takes_the_other_way(engine, key):
    fx = FRAGILE[key]
    _announce(f"fragile {key}")
    g = cds_amd.Graph.load(os.path.join(GOLDEN, fx["graph"]))
    got = engine.solve(engine.upload(g), fx["k"], vertex_arrays=True)
    ex = oracle_solve(g, fx["k"], flow=FLOW_EXACT)
    compare(got, ex, key)
    fl = oracle_solve(g, fx["k"], flow=FLOW_FLOAT_EK)
    assert (float(fl["density"]), fl["vertices"]) == (float(fx["density"]), fx["vertices"])
    assert list(got["vertices"]) != fx["vertices"]
    assert got["fragile_decisions"] > 0
    # the first step where the float search decides otherwise is flagged
    for i, (a, b) in enumerate(zip(fl["trace_alpha"], got["trace_alpha"])):
        assert a == b, f"alpha sequences part before any decision differs (step {i})"
        if fl["trace_equal"][i] != ex["trace_equal"][i]:
            assert got["trace_fragile"][i] == 1, f"step {i} not flagged"
            assert got["trace_drift"][i] >= abs(fl["trace_margin"][i] - got["trace_margin"][i])
            break
    else:
        pytest.fail("the float and exact searches never part")


def _cli(path, k, env=None):
    r = subprocess.run([cds_amd.CLI_PATH, path, str(k), "8000", "16000", "60000", "30000",
                        "-1"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _ref(path, k, tmp_path):
    motif = tmp_path / f"motif{k}.txt"
    edges = [f"{a} {b}" for a in range(k) for b in range(a + 1, k)]
    motif.write_text(f"{k} 1 {len(edges)}\n" + "\n".join(edges) + "\n")
    r = subprocess.run([REF_DSD, path, str(motif), "1", "1"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _block(out):
    keys = ("K:", "Density:", "Size:", "densest subgraph:")
    return [" ".join(line.split()) for line in out.splitlines() if line.startswith(keys)]


CLI_CASES = sorted(k for k in STABLE
                   if k.startswith(("cl_00", "cl_01", "cl_02", "cl_03", "cl_04", "rmat7_0",
                                    "ref_data")))


@pytest.mark.skipif(not os.path.exists(REF_DSD), reason="oracle/_ref/dsd not built")
def test_cli_prints_the_reference_answer_block(tmp_path):
    if not os.path.exists(cds_amd.CLI_PATH):
        pytest.fail("cds_main not built")
    for key in CLI_CASES:
        fx = STABLE[key]
        path = os.path.join(GOLDEN, fx["graph"])
        _announce(f"cds_main vs reference {key}")
        ours, ref = _block(_cli(path, fx["k"])), _block(_ref(path, fx["k"], tmp_path))
        assert len(ref) == 4 and ours == ref, f"{key}: {ours} != {ref}"


@pytest.mark.skipif(not os.path.exists(REF_DSD), reason="oracle/_ref/dsd not built")
def test_cli_on_fragile_graphs_differs_and_says_why(tmp_path):
    for key in sorted(FRAGILE):
        fx = FRAGILE[key]
        path = os.path.join(GOLDEN, fx["graph"])
        _announce(f"cds_main vs reference (fragile) {key}")
        env = dict(os.environ, CDS_JSON=str(tmp_path / "r.json"))
        ours, ref = _block(_cli(path, fx["k"], env)), _block(_ref(path, fx["k"], tmp_path))
        assert ours != ref
        rec = json.loads((tmp_path / "r.json").read_text())
        assert rec["search"]["fragile_decisions"] > 0
        assert any(step["fragile"] for step in rec["search"]["trace"])
