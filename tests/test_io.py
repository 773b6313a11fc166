"""Host-only checks of the input pipeline and the result record: the
parallel adjacency-list loader gives the same CSR and the same first error
for any thread count, the binary CSR file round-trips and rejects damaged
files, and cds_result_json renders a result as valid JSON."""
import ctypes
import json
import os

import numpy as np
import pytest

import cds_amd


def reference_csr(n, lines):
    """CSR the loader must produce from (v, [w...]) lines: lists merged per
    vertex, self loops dropped, sorted, de-duplicated."""
    adj = [set() for _ in range(n)]
    for v, ws in lines:
        adj[v].update(w for w in ws if w != v)
    off = np.zeros(n + 1, np.int64)
    for v in range(n):
        off[v + 1] = off[v] + len(adj[v])
    nb = np.array([w for v in range(n) for w in sorted(adj[v])], np.int64)
    return off, nb


def messy_text(g, rng):
    """The graph's lines shuffled, split in two for some vertices, with
    duplicated neighbours, self loops, blank lines, tabs and CRLF endings."""
    off, nb = g.csr()
    lines = []
    for v in range(g.n):
        ws = [int(w) for w in nb[off[v]:off[v + 1]]]
        if ws and rng.random() < 0.2:
            ws.append(ws[0])
        if rng.random() < 0.1:
            ws.append(v)
        if len(ws) > 2 and rng.random() < 0.3:
            cut = len(ws) // 2
            lines += [(v, ws[:cut]), (v, ws[cut:])]
        else:
            lines.append((v, ws))
    order = rng.permutation(len(lines))
    out = [f"{g.n} {g.m}"]
    for i in order:
        v, ws = lines[i]
        sep = "\t" if rng.random() < 0.1 else " "
        end = "\r" if rng.random() < 0.1 else ""
        out.append(sep.join(str(x) for x in [v] + ws) + end)
        if rng.random() < 0.05:
            out.append("")
    return "\n".join(out) + "\n", [lines[i] for i in order]


@pytest.mark.parametrize("threads", ["1", "3", "16"])
def test_parallel_loader_matches_reference_csr(tmp_path, monkeypatch, threads):
    rng = np.random.default_rng(5)
    g = cds_amd.Graph.rmat(14, 16, seed=9)  # ~3 MB of text: several chunks
    text, lines = messy_text(g, rng)
    path = tmp_path / "g.txt"
    path.write_text(text)
    monkeypatch.setenv("CDS_IO_THREADS", threads)
    h = cds_amd.Graph.load(str(path))
    off, nb = reference_csr(g.n, lines)
    a, b = h.csr()
    assert np.array_equal(a, off) and np.array_equal(b, nb)
    assert h.m == g.m


def test_first_error_is_reported_for_any_thread_count(tmp_path, monkeypatch):
    g = cds_amd.Graph.rmat(14, 16, seed=9)
    good = tmp_path / "g.txt"
    g.write(str(good))
    lines = good.read_text().split("\n")
    # two faults; the earlier one (line numbers are 1-based, header = 1) wins
    early, late = len(lines) // 2, len(lines) - 10
    lines[early] = lines[early] + " x"
    lines[late] = lines[late] + f" {g.n + 7}"
    bad = tmp_path / "bad.txt"
    bad.write_text("\n".join(lines))
    msgs = set()
    for t in ("1", "4", "16"):
        monkeypatch.setenv("CDS_IO_THREADS", t)
        with pytest.raises(cds_amd.CdsError) as ei:
            cds_amd.Graph.load(str(bad))
        msgs.add(str(ei.value))
    assert len(msgs) == 1
    msg = msgs.pop()
    assert f":{early + 1}: bad neighbor token" in msg


def test_asymmetry_reports_the_smallest_pair(tmp_path, monkeypatch):
    g = cds_amd.Graph.rmat(13, 8, seed=4)
    off, nb = g.csr()
    text = [f"{g.n} {g.m}"]
    dropped = []
    for v in range(g.n):
        ws = [int(w) for w in nb[off[v]:off[v + 1]]]
        if v in (300, 5000) and ws:
            dropped.append((ws[-1], v))  # w lists v, v no longer lists w
            ws = ws[:-1]
        text.append(" ".join(str(x) for x in [v] + ws))
    path = tmp_path / "asym.txt"
    path.write_text("\n".join(text) + "\n")
    w, v = min(dropped)
    for t in ("1", "8"):
        monkeypatch.setenv("CDS_IO_THREADS", t)
        with pytest.raises(cds_amd.CdsError) as ei:
            cds_amd.Graph.load(str(path))
        assert f"{w} lists {v} but not back" in str(ei.value)


def test_binary_csr_roundtrip(tmp_path):
    g = cds_amd.Graph.chung_lu(3000, 2.3, 8.0, 30, 0.9, seed=12)
    p = str(tmp_path / "g.csr")
    g.save_csr(p)
    h = cds_amd.Graph.load_csr(p)
    assert (h.n, h.m, h._g.max_degree) == (g.n, g.m, g._g.max_degree)
    a, b = g.csr(), h.csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    e = cds_amd.Graph.from_csr(np.zeros(1, np.uint32), np.zeros(0, np.uint32))
    e.save_csr(p)
    assert cds_amd.Graph.load_csr(p).n == 0


def damaged(tmp_path, g, how):
    p = tmp_path / f"{how}.csr"
    g.save_csr(str(p))
    raw = bytearray(p.read_bytes())
    hdr = 32
    n = g.n
    nb0 = hdr + 4 * (n + 1)
    if how == "magic":
        raw[0] ^= 1
    elif how == "truncated":
        raw = raw[:-4]
    elif how == "range":
        raw[nb0:nb0 + 4] = (n + 5).to_bytes(4, "little")
    elif how == "asymmetric":
        # last list entry of v replaced by n-1, larger than all of v's
        # neighbours (the list stays sorted) and so not adjacent to v
        off, nb = g.csr()
        v = next(v for v in range(n - 1)
                 if off[v + 1] > off[v] and nb[off[v + 1] - 1] < n - 1)
        a = nb0 + 4 * (int(off[v + 1]) - 1)
        raw[a:a + 4] = (n - 1).to_bytes(4, "little")
    elif how == "unsorted":
        off, nb = g.csr()
        v = int(np.argmax(np.diff(off.astype(np.int64))))
        a = nb0 + 4 * int(off[v])
        raw[a:a + 4], raw[a + 4:a + 8] = raw[a + 4:a + 8], raw[a:a + 4]
    p.write_bytes(bytes(raw))
    return str(p)


@pytest.mark.parametrize("how", ["magic", "truncated", "range", "asymmetric", "unsorted"])
def test_binary_csr_rejects_damage(tmp_path, how):
    g = cds_amd.Graph.chung_lu(500, 2.5, 6.0, 10, 0.9, seed=3)
    with pytest.raises(cds_amd.CdsError) as ei:
        cds_amd.Graph.load_csr(damaged(tmp_path, g, how))
    assert ei.value.status == 2


def test_result_json_renders_every_section():
    r = cds_amd.CResult()
    r.k, r.n, r.m, r.total_cliques = 3, 10, 20, 7
    r.density, r.size, r.set_cliques, r.set_size = 1.25, 4, 5, 4
    verts = (ctypes.c_uint32 * 4)(1, 3, 5, 9)
    r.vertices = ctypes.cast(verts, ctypes.POINTER(ctypes.c_uint32))
    lb = (ctypes.c_float * 2)(0.5, 1.0)
    ub = (ctypes.c_float * 2)(2.0, 1.5)
    al = (ctypes.c_float * 2)(1.25, 1.25)
    comp = (ctypes.c_uint32 * 2)(0, 0)
    mar = (ctypes.c_double * 2)(3.0, float("nan"))
    r.trace_len = 2
    r.trace_lb = ctypes.cast(lb, ctypes.POINTER(ctypes.c_float))
    r.trace_ub = ctypes.cast(ub, ctypes.POINTER(ctypes.c_float))
    r.trace_alpha = ctypes.cast(al, ctypes.POINTER(ctypes.c_float))
    r.trace_component = ctypes.cast(comp, ctypes.POINTER(ctypes.c_uint32))
    r.trace_margin = ctypes.cast(mar, ctypes.POINTER(ctypes.c_double))
    frag = (ctypes.c_uint8 * 2)(0, 1)
    drift = (ctypes.c_double * 2)(0.5, 0.25)
    r.trace_fragile = ctypes.cast(frag, ctypes.POINTER(ctypes.c_uint8))
    r.trace_drift = ctypes.cast(drift, ctypes.POINTER(ctypes.c_double))
    r.fragile_decisions = 1
    r.peel_grid, r.flow_grid, r.replay_grid, r.coop_shrinks = 256, 128, 64, 1
    r.times.total = 12.5
    r.ub1_val = float("inf")
    text = cds_amd.result_json(r)
    d = json.loads(text)
    assert d["answer"] == {"density": 1.25, "size": 4, "set_cliques": 5, "set_size": 4,
                           "density_exact": "5/4", "vertices": [1, 3, 5, 9]}
    assert d["k"] == 3 and d["total_cliques"] == 7
    assert d["search"]["ub1"] is None
    assert d["search"]["trace"][1] == {"component": 0, "lb": 1.0, "ub": 1.5,
                                       "alpha": 1.25, "margin": None,
                                       "drift_bound": 0.25, "fragile": 1}
    assert d["search"]["fragile_decisions"] == 1
    assert d["cooperative"] == {"peel_workgroups": 256, "flow_workgroups": 128,
                                "replay_workgroups": 64, "shrinks": 1}
    assert d["times_ms"]["total"] == 12.5
    # snprintf convention: a short buffer is cut and NUL-terminated
    buf = ctypes.create_string_buffer(10)
    need = cds_amd.lib().cds_result_json(ctypes.byref(r), buf, 10)
    assert need == len(text) and buf.value.decode() == text[:9]


def test_result_json_density_keeps_the_float():
    r = cds_amd.CResult()
    r.density = 57145.640625
    d = json.loads(cds_amd.result_json(r))
    assert np.float32(d["answer"]["density"]) == np.float32(57145.640625)
    r.density = 1 / 3
    d = json.loads(cds_amd.result_json(r))
    assert np.float32(d["answer"]["density"]) == np.float32(1 / 3)
