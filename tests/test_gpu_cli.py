"""GPU checks of the drop-in binary's output side: the report lines, the
JSON record (CDS_JSON) and the binary CSR cache (CDS_CSR_CACHE) agree with
the library's own result for the same graph."""
import json
import os
import subprocess

import numpy as np
import pytest

import cds_amd
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu


def run_cli(path, k, env):
    r = subprocess.run([cds_amd.CLI_PATH, path, str(k), "8000", "16000", "60000",
                        "30000", "-1"], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cli_json_record_and_csr_cache(engine, tmp_path):
    if not os.path.exists(cds_amd.CLI_PATH):
        pytest.fail("cds_main not built")
    g = cds_amd.Graph.chung_lu(3000, 2.4, 8.0, 25, 0.9, seed=31)
    path = str(tmp_path / "g.txt")
    g.write(path)
    cache = tmp_path / "cache"
    cache.mkdir()
    env = dict(os.environ, CDS_JSON=str(tmp_path / "r.json"), CDS_CSR_CACHE=str(cache))
    _announce("cds_main with CDS_JSON and CDS_CSR_CACHE, k=3, twice")
    first = run_cli(path, 3, env)
    rec = json.loads((tmp_path / "r.json").read_text())
    cached = list(cache.iterdir())
    assert len(cached) == 1 and cached[0].name.endswith(".csr")
    h = cds_amd.Graph.load_csr(str(cached[0]))
    assert np.array_equal(h.csr()[1], g.csr()[1])
    second = run_cli(path, 3, env)

    def answer_lines(out):
        keep = ("K:", "Density:", "Size:", "densest subgraph:", "Density_exact=",
                "Total Cliques=", "n=", "Max Degree=", "Mac Clique Core=")
        return [l for l in out.splitlines() if l.startswith(keep)]

    assert answer_lines(first) == answer_lines(second)
    want = engine.solve(engine.upload(g), 3)
    a = rec["answer"]
    assert np.float32(a["density"]) == np.float32(want["density"])
    assert (a["size"], a["set_cliques"], a["set_size"]) == \
        (want["size"], want["set_cliques"], want["set_size"])
    assert a["vertices"] == list(want["vertices"])
    assert rec["total_cliques"] == want["total_cliques"]
    assert rec["core"]["k_prime"] == want["k_prime"]
    assert f"Density_exact={want['set_cliques']}/{want['set_size']}" in first
    assert len(rec["search"]["trace"]) == len(want["trace_alpha"])
