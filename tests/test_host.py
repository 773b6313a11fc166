"""Host-only checks of the C ABI library (no GPU calls): it loads, exports
every function include/cds.h declares with matching struct layouts, parses
and validates the reference's adjacency-list format, and its generators are
deterministic."""
import os
import re
import subprocess

import numpy as np
import pytest

import cds_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cds.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cds_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = cds_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(cds_amd._SIGNATURES), \
        "bindings out of sync with include/cds.h"


def test_struct_layouts_match():
    L = cds_amd.lib()
    for i, cls in enumerate((cds_amd.CGraph, cds_amd.CParams, cds_amd.CResult,
                             cds_amd.COpts)):
        assert L.cds_struct_size(i) == __import__("ctypes").sizeof(cls)


def test_load_roundtrip_and_format(tmp_path):
    g = cds_amd.Graph.chung_lu(500, 2.5, 6.0, 10, 0.9, seed=3)
    path = str(tmp_path / "g.txt")
    g.write(path)
    first = open(path).readline().split()
    assert first == [str(g.n), str(g.m)]
    h = cds_amd.Graph.load(path)
    a, b = g.csr(), h.csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_loader_drops_self_loops_and_sorts(tmp_path):
    path = tmp_path / "g.txt"
    path.write_text("4 3\n0 2 1 0\n1 0\n2 0 3\n3 2\n")
    g = cds_amd.Graph.load(str(path))
    off, nb = g.csr()
    assert off.tolist() == [0, 2, 3, 5, 6]
    assert nb.tolist() == [1, 2, 0, 0, 3, 2]


@pytest.mark.parametrize("text,msg", [
    ("3 1\n0 1\n1\n", "not symmetric"),
    ("3 1\n0 5\n5 0\n", ">= n"),
    ("3 1\n0 x\n", "bad neighbor"),
])
def test_loader_rejects_malformed(tmp_path, text, msg):
    path = tmp_path / "bad.txt"
    path.write_text(text)
    with pytest.raises(cds_amd.CdsError) as ei:
        cds_amd.Graph.load(str(path))
    assert msg in str(ei.value)


def test_missing_file_is_an_io_error():
    with pytest.raises(cds_amd.CdsError) as ei:
        cds_amd.Graph.load("/nonexistent/graph.txt")
    assert ei.value.status == 2


def test_generators_are_deterministic():
    a = cds_amd.Graph.chung_lu(2000, 2.5, 10.0, 40, 0.9, seed=11)
    b = cds_amd.Graph.chung_lu(2000, 2.5, 10.0, 40, 0.9, seed=11)
    assert np.array_equal(a.csr()[1], b.csr()[1])
    r1 = cds_amd.Graph.rmat(10, 8, seed=5)
    r2 = cds_amd.Graph.rmat(10, 8, seed=5)
    assert np.array_equal(r1.csr()[1], r2.csr()[1])
    off, nb = r1.csr()
    # symmetric, no self loops, sorted lists
    for v in range(r1.n):
        lst = nb[off[v]:off[v + 1]]
        assert np.all(np.diff(lst.astype(np.int64)) > 0)
        assert v not in lst


def test_cli_rejects_wrong_argc():
    if not os.path.exists(cds_amd.CLI_PATH):
        pytest.skip("cds_main not built")
    r = subprocess.run([cds_amd.CLI_PATH, "a", "3"], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stdout.strip() == "Server wrong input parameters!"
