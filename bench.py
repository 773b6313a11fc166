"""Benchmark of the plain exact DSD path on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3]

A step is one exact k-clique densest-subgraph solve (DAG, listing, peel,
densest core, prune, components, exact search) of one synthetic graph that
is already resident in HBM.  The default workload is the plain-method GPU
configuration: R-MAT scale 20 (a=.57, b=.19, c=.19, d=.05), edge factor 16,
symmetrised, de-duplicated, self loops removed, isolated vertices kept,
k = 3.  Prints one JSON line (metric/value/unit/... + roofline +
cpu_baseline).
"""
import argparse
import json
import math
import os
import re
import struct
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact DSD solves/s + p50/p99 latency, served vs plain; k-cliques/s (k=3)"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c1": dict(kind="chung_lu", n=10000, exponent=2.5, mean_degree=10.0,
               planted=40, p=0.9, k=3, seed=1,
               text="plain method, Chung-Lu n=10000 exponent 2.5 mean degree 10 "
                    "+ planted G(40,0.9), k=3"),
    "c2": dict(kind="rmat", scale=20, edge_factor=16, k=3, seed=2,
               text="plain method, R-MAT scale 20 (a=.57 b=.19 c=.19 d=.05) "
                    "edge factor 16, symmetrised, deduplicated, k=3"),
    "c3": dict(kind="rmat", scale=18, edge_factor=24, planted=32, pmin=30,
               pmax=120, p=0.85, k=4, seed=3,
               text="plain method, R-MAT scale 18 edge factor 24 + 32 planted "
                    "G(s,0.85) s in [30,120], k=4"),
}


def make_graph(w):
    import cds_amd

    if w["kind"] == "chung_lu":
        return cds_amd.Graph.chung_lu(w["n"], w["exponent"], w["mean_degree"],
                                      w["planted"], w["p"], seed=w["seed"])
    return cds_amd.Graph.rmat(w["scale"], w["edge_factor"], 0.57, 0.19, 0.19,
                              w.get("planted", 0), w.get("pmin", 0),
                              w.get("pmax", 0), w.get("p", 0.0), seed=w["seed"])


def list3_bytes(n, m, W):
    """Algorithmic HBM bytes of one k=3 listing count launch: CSR offsets and
    out-lists read once, per DAG arc (u,v) its two offsets and N+(v), per
    vertex one clique-degree write."""
    return 4 * (n + 1) + 4 * m + 8 * m + 4 * W + 8 * n


BATCH_FIELDS = ("total_cliques", "k_prime", "core_size", "core_cliques",
                "num_components", "size", "set_cliques", "set_size")


def mixed_graph(rng, seed):
    """One graph of the batched workload: Chung-Lu with n log-uniform in
    [200, 20000], exponent in [2.1, 3.0], mean degree in [4, 24] and one
    planted G(s, 0.9), s in [8, 60]."""
    import cds_amd

    n = int(math.exp(rng.uniform(math.log(200), math.log(20000))))
    return cds_amd.Graph.chung_lu(n, float(rng.uniform(2.1, 3.0)),
                                  float(rng.uniform(4, 24)),
                                  int(rng.integers(8, 61)), 0.9, seed=seed)


def same_answer(a, b):
    return (all(a[f] == b[f] for f in BATCH_FIELDS)
            and struct.pack("f", a["density"]) == struct.pack("f", b["density"])
            and list(a["vertices"]) == list(b["vertices"]))


def bench_batched(args):
    """4096 independent graphs (k = 3 w.p. 0.7, else 4) solved by
    cds_solve_batch, every answer compared with the plain per-graph path
    (cds_solve one graph at a time) on the same device-resident graphs."""
    import numpy as np
    import cds_amd

    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    graphs = [mixed_graph(rng, 1_000_000 + i) for i in range(args.graphs)]
    ks = [3 if rng.random() < 0.7 else 4 for _ in graphs]
    print(f"[bench] generated {len(graphs)} graphs in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    eng = cds_amd.Engine(device=0, tune_cache="" if args.untuned else None)
    dgs = [eng.upload(g) for g in graphs]
    eng.solve_batch(dgs[:64], ks[:64])  # warm-up
    t0 = time.perf_counter()
    plain = []
    for i, (d, k) in enumerate(zip(dgs, ks)):
        plain.append(eng.solve(d, k))
        if (i + 1) % 512 == 0:
            print(f"[bench] plain {i + 1}: {time.perf_counter() - t0:.2f}s",
                  file=sys.stderr, flush=True)
    t_plain = time.perf_counter() - t0
    print(f"[bench] plain: {t_plain:.2f}s", file=sys.stderr, flush=True)
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        got = []
        for c in range(0, len(dgs), args.chunk):
            got += eng.solve_batch(dgs[c:c + args.chunk], ks[c:c + args.chunk])
            print(f"[bench] batched step {i}: {c + args.chunk} graphs "
                  f"{time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
        times.append(time.perf_counter() - t0)
        print(f"[bench] batched step {i}: {times[-1]:.2f}s", file=sys.stderr, flush=True)
    mism = sum(1 for a, b in zip(got, plain) if not same_answer(a, b))
    t_batch = min(times)
    n_total = sum(g.n for g in graphs)
    return {
        "metric": METRIC, "value": len(graphs) / t_batch, "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": t_batch * 1000, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"batched form: {len(graphs)} Chung-Lu graphs, n "
                               "log-uniform [200, 20000], exponent [2.1, 3.0], mean "
                               "degree [4, 24], planted G(s, 0.9) s in [8, 60], k = 3 "
                               "(70%) / 4 (30%)",
                   "graphs": len(graphs), "graphs_per_call": args.chunk,
                   "vertices": n_total, "seed": args.seed},
        "plain_solves_per_s": len(graphs) / t_plain,
        "batched_vs_plain": t_plain / t_batch,
        "mismatches_vs_plain": mism,
        "plain_p50_ms": percentile([r["times"]["total"] for r in plain], 50),
        "plain_p99_ms": percentile([r["times"]["total"] for r in plain], 99),
    }


def serving_stream(rng, rate, duration, seed0):
    """Open-loop Poisson arrivals at `rate` per second for `duration` s; 95%
    batched-workload graphs, 5% R-MAT at scale 16..18 (edge factor 16); k in
    {2, 3, 4} w.p. 0.3 / 0.5 / 0.2.  R-MAT requests reuse a pool of six
    seeded graphs so the stream can be generated ahead of time."""
    import cds_amd

    pool = {}
    t, out, i = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        if rng.random() < 0.05:
            scale = int(rng.integers(16, 19))
            key = (scale, int(rng.integers(0, 2)))
            if key not in pool:
                pool[key] = cds_amd.Graph.rmat(scale, 16, seed=seed0 + 97 * scale + key[1])
            g = pool[key]
        else:
            g = mixed_graph(rng, seed0 + i)
        k = int(rng.choice([2, 3, 4], p=[0.3, 0.5, 0.2]))
        out.append((t, g, k))
        i += 1
    return out


def run_stream(stream, submit_fn, wait_all):
    """Replays the arrival times from a producer thread; returns per-request
    latency (completion - arrival) in ms and the wall time."""
    import threading

    t0 = time.perf_counter()
    done_at = [None] * len(stream)

    def producer():
        for i, (t, g, k) in enumerate(stream):
            delay = t - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            submit_fn(i, g, k)

    th = threading.Thread(target=producer)
    th.start()
    wait_all(done_at, t0)
    th.join()
    wall = time.perf_counter() - t0
    lat = [(done_at[i] - stream[i][0]) * 1000 for i in range(len(stream))]
    return lat, wall


def bench_serving(args):
    """Served form: an open-loop Poisson stream per offered load, through
    cds_server (continuous batching, HBM budget 200 GB) and through the
    plain one-graph-at-a-time path on the same stream."""
    import queue
    import numpy as np
    import cds_amd

    loads = [float(x) for x in args.loads.split(",")]
    eng = cds_amd.Engine(device=0, tune_cache="" if args.untuned else None)
    rows = []
    for li, rate in enumerate(loads):
        rng = np.random.default_rng(args.seed + li)
        t0 = time.time()
        stream = serving_stream(rng, rate, args.duration, 2_000_000 + 10_000 * li)
        print(f"[bench] load {rate}/s: {len(stream)} requests generated in "
              f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        # served
        srv = cds_amd.Server(eng, hbm_budget_bytes=200 << 30)
        tickets = [None] * len(stream)

        def submit_served(i, g, k):
            tickets[i] = srv.submit(g, k)

        served = [None] * len(stream)

        def progress(path, i):
            if (i + 1) % 250 == 0:
                print(f"[bench] load {rate}/s {path}: {i + 1}/{len(stream)} done",
                      file=sys.stderr, flush=True)

        def wait_served(done_at, t0):
            for i in range(len(stream)):
                while tickets[i] is None:
                    time.sleep(0.0005)
                try:
                    served[i] = srv.wait(tickets[i])
                except cds_amd.CdsError as ex:  # refused: the request's answer is its status
                    served[i] = ex.status
                done_at[i] = time.perf_counter() - t0
                progress("served", i)

        lat_s, wall_s = run_stream(stream, submit_served, wait_served)
        stats = srv.stats()
        srv.close()
        # plain: one worker, FIFO, one graph per solve
        q = queue.Queue()

        def submit_plain(i, g, k):
            q.put((i, g, k))

        plain = [None] * len(stream)
        n_done = [0]

        def wait_plain(done_at, t0):
            for _ in range(len(stream)):
                i, g, k = q.get()
                s0 = time.perf_counter()
                try:
                    plain[i] = eng.solve(eng.upload(g), k)  # device graph freed here
                except cds_amd.CdsError as ex:
                    plain[i] = ex.status
                done_at[i] = time.perf_counter() - t0
                if done_at[i] - (s0 - t0) > 2.0:  # a slow solve is named as it ends
                    print(f"[bench] load {rate}/s plain: request {i} (n={g.n}, k={k}) "
                          f"took {done_at[i] - (s0 - t0):.1f}s", file=sys.stderr, flush=True)
                progress("plain", n_done[0])
                n_done[0] += 1

        lat_p, wall_p = run_stream(stream, submit_plain, wait_plain)
        refused = sum(isinstance(x, int) for x in served)
        # a request refused by both paths (the server by its HBM budget or
        # range, the plain path by range) counts as the same outcome
        mismatch = sum(1 for a, b in zip(served, plain)
                       if isinstance(a, int) != isinstance(b, int)
                       or (not isinstance(a, int) and not same_answer(a, b)))
        row = {"offered_per_s": rate, "requests": len(stream),
               "refused": refused, "mismatches_served_vs_plain": mismatch,
               "served_solves_per_s": len(stream) / wall_s,
               "served_p50_ms": percentile(lat_s, 50),
               "served_p99_ms": percentile(lat_s, 99),
               "plain_solves_per_s": len(stream) / wall_p,
               "plain_p50_ms": percentile(lat_p, 50),
               "plain_p99_ms": percentile(lat_p, 99),
               "batches": stats["batches"], "max_batch": stats["max_batch_seen"]}
        print(f"[bench] load {rate}/s: {json.dumps(row)}", file=sys.stderr, flush=True)
        rows.append(row)
    best = max(rows, key=lambda r: r["served_solves_per_s"])
    return {
        "metric": METRIC, "value": best["served_solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": len(rows), "warmup": 0,
        "ms_per_step": 1000.0 / best["served_solves_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": "served form: open-loop Poisson stream, 95% batched-"
                               "workload Chung-Lu graphs, 5% R-MAT scale 16-18 edge "
                               "factor 16, k in {2,3,4} (30/50/20%), HBM budget 200 GB",
                   "duration_s": args.duration, "loads_per_s": loads,
                   "seed": args.seed},
        "p50_ms": best["served_p50_ms"], "p99_ms": best["served_p99_ms"],
        "loads": rows,
    }


def peel_bytes(n, T, k):
    """Algorithmic HBM bytes of one persistent peel launch: every clique's
    incidence entries read once (4kT), members read, kill word set and each
    clique killed once ((4k+8)T), the k-1 degree decrements per killed clique
    as read-modify-writes (8(k-1)T), and per vertex its degree read, core
    write and one frontier slot (12n)."""
    return 4 * k * T + (4 * k + 8) * T + 8 * (k - 1) * T + 12 * n


def flow_bytes_wide(r, k):
    """flow_bytes for the wide kernel (networks of 2^32 clique-member slots
    or more): per clique visit its excess and height (8 B) and per arc the
    member id, the 4-byte residual and the member height (12 B); per vertex
    arc the clique id (4 B), the clique's k member ids to find the slot
    (4k B), the residual and the clique height (8 B); per global relabel
    every arc once (12 B); per round the vertex state (12 B)."""
    arcs = k * r["core_cliques"]
    return (r["flow_clique_visits"] * (8 + 12 * k) + r["flow_arc_visits"] * (12 + 4 * k)
            + r["flow_relabels"] * 12 * arcs + r["flow_sweeps"] * 12 * r["core_size"])


def flow_bytes(r, k):
    """Algorithmic HBM bytes of the whole-network max flows of one solve, from
    the work the rounds did: per clique work-list visit its excess and
    height (12 B) and its k arcs (member id 4 B, residual 8 B, member height
    4 B); per vertex incidence arc posted for discharge its clique slot
    (4 B), residual (8 B) and clique height (4 B); per global relabel pass
    every arc once (16 B, the per-pass figure of a BFS over the network);
    per round the vertex state (excess 8 B, height 4 B).  The per-sweep
    figure 16|E| + 12|V| would count every arc in every round, which the
    work lists avoid."""
    arcs = k * r["core_cliques"]
    return (r["flow_clique_visits"] * (12 + 16 * k) + r["flow_arc_visits"] * 16
            + r["flow_relabels"] * 16 * arcs + r["flow_sweeps"] * 12 * r["core_size"])


def flow_bytes_imp(r, k):
    """Algorithmic HBM bytes of the implicit-clique max flows of one solve:
    per vertex incidence arc posted for discharge its clique slot (4 B) and
    the clique's record (member ids and residuals, stride words of 4 B, one
    line); per global relabel pass every arc once (the same 4 B + record);
    per round the vertex state (excess, inflow 8 B each, height 4 B)."""
    stride = 4 if 2 * k <= 4 else 8 if 2 * k <= 8 else 16
    arc = 4 + 4 * stride
    arcs = k * r["core_cliques"]
    return (r["flow_arc_visits"] * arc + r["flow_relabels"] * arc * arcs
            + r["flow_sweeps"] * 20 * r["core_size"])


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summaries (profiles/*.txt lines
    `traffic_bytes_per_launch <workload> <kernel> N`)."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".txt"):
            continue
        for line in open(os.path.join(pdir, name)):
            m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+(\S+)\s+(\d+)", line)
            if m and m.group(1) == workload and m.group(2) == kernel:
                return int(m.group(3))
    return None


def host_cpu():
    """The host's CPU model (lscpu) and core counts."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        model = m.group(1).strip() if m else None
    except Exception:
        pass
    return dict(model=model, nproc=os.cpu_count(),
                affinity=len(os.sched_getaffinity(0)))


def _threads_available():
    """Cores this process may use: OMP_NUM_THREADS when set (the GPU host
    sets it to the share of one GPU), else the affinity mask."""
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return len(os.sched_getaffinity(0))


def _run_reference(cmd, timeout):
    """Runs a reference binary; returns (wall s, stdout, status) with status
    'ok', 'timeout' or 'exit <code>'."""
    t0 = time.time()
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return time.time() - t0, "", "timeout"
    return time.time() - t0, r.stdout, "ok" if r.returncode == 0 else f"exit {r.returncode}"


def _printed_answer(out):
    dens = re.search(r"^Density:\s*(\S+)", out, re.M)
    size = re.search(r"^Size:\s*(\d+)", out, re.M)
    verts = re.search(r"^densest subgraph:\s*(.*)$", out, re.M)
    return (dens.group(1) if dens else None, int(size.group(1)) if size else None,
            [int(x) for x in verts.group(1).split()] if verts else None)


C1_REPS = 3          # serial reference solves of C1 (about 1.7 s each)
C2_CAP_S = 30.0      # wall cap for the serial reference on the scale-20 graph


def cpu_baseline(tmpdir, eng):
    """The reference's CPU paths timed on this host, in this run.

    * cpu_baseline proper: cpuCode/c++ (oracle/_ref/dsd, built from the
      reference's sources), one core, C1_REPS full solves of C1 (Chung-Lu
      n=10,000, the reference-runnable configuration), with its printed
      answer compared to the engine's solve of the same graph.
    * cpuCode/parallel (oracle/_ref/dsd_parallel, its sources with a main
      that takes the thread count) on C1 at 1 thread, at its built-in 2
      threads and at every available core; a run that aborts is reported
      with its exit status (its connectedComponentDecompose overflows a
      heap buffer on C1 with more than one thread, helpers.cpp:1396).
    * the serial reference on the scale-20 graph under a C2_CAP_S wall cap:
      it does not finish (O(n^2) peel), so the speedup there is a bound."""
    dsd = os.path.join(ROOT, "oracle", "_ref", "dsd")
    dsdp = os.path.join(ROOT, "oracle", "_ref", "dsd_parallel")
    if not os.path.exists(dsd):
        return None
    import cds_amd

    motif = os.path.join(tmpdir, "motif3.txt")
    with open(motif, "w") as f:
        f.write("3 1 3\n0 1\n0 2\n1 2\n")
    c1 = make_graph(WORKLOADS["c1"])
    path = os.path.join(tmpdir, "c1.txt")
    c1.write(path)
    times, answers = [], set()
    out = ""
    for _ in range(C1_REPS):
        wall, out, st = _run_reference([dsd, path, motif, "1", "1"], 120)
        if st != "ok":
            return dict(error=f"cpuCode/c++ on C1: {st}")
        m = re.search(r"Execution_time:\s*([0-9.eE+-]+)", out)
        times.append(float(m.group(1)) / 1000.0 if m else wall)
        answers.add(repr(_printed_answer(out)))
    cpu_s = sorted(times)[len(times) // 2]
    dens, size, verts = _printed_answer(out)
    dg = eng.upload(c1)
    gr = eng.solve(dg, 3)
    reps, t1 = 20, time.perf_counter()
    for _ in range(reps):
        gr = eng.solve(dg, 3)
    gpu_s = (time.perf_counter() - t1) / reps
    dg.free()
    gpu_answer = (f"{float(gr['density']):g}", gr["size"], list(gr["vertices"]))
    res = dict(value=1.0 / cpu_s, unit="solves/s", cores=1, kind="reference",
               sample=f"C1 (Chung-Lu n=10000, exponent 2.5, mean degree 10, planted "
                      f"G(40,0.9), seed 1, k=3): median of {C1_REPS} full solves by "
                      f"cpuCode/c++ built from source (-O3 -march=x86-64-v3 "
                      f"-ffast-math), {cpu_s * 1000:.0f} ms each",
               host=host_cpu(),
               gpu_same_graph_solves_per_s=1.0 / gpu_s,
               speedup_same_graph=cpu_s / gpu_s,
               same_answer=(dens, size, verts) == gpu_answer and len(answers) == 1,
               reference_answer={"density": dens, "size": size},
               gpu_answer={"density": gpu_answer[0], "size": gpu_answer[1],
                           "exact": f"{gr['set_cliques']}/{gr['set_size']}"})
    if os.path.exists(dsdp):
        par = []
        for t in sorted({1, 2, _threads_available()}):
            wall, pout, st = _run_reference([dsdp, path, motif, "1", "1", str(t)], 120)
            m = re.search(r"Execution_time:\s*([0-9.eE+-]+)", pout)
            entry = dict(threads=t, status=st)
            if st == "ok":
                entry["seconds"] = float(m.group(1)) / 1000.0 if m else wall
                entry["solves_per_s"] = 1.0 / entry["seconds"]
            par.append(entry)
        res["cpuCode_parallel_C1"] = par
    return res


def cpu_c2_bound(tmpdir, graph, gpu_ms):
    """The serial reference on the scale-20 graph under a wall cap: a lower
    bound on the plain path's speedup there."""
    dsd = os.path.join(ROOT, "oracle", "_ref", "dsd")
    if not os.path.exists(dsd):
        return None
    motif = os.path.join(tmpdir, "motif3.txt")
    with open(motif, "w") as f:
        f.write("3 1 3\n0 1\n0 2\n1 2\n")
    path = os.path.join(tmpdir, "c2.txt")
    graph.write(path)
    wall, out, st = _run_reference([dsd, path, motif, "1", "1"], C2_CAP_S)
    res = dict(cores=1, kind="reference", cap_s=C2_CAP_S, status=st)
    if st == "timeout":
        res["seconds"] = f"> {C2_CAP_S:g}"
        res["speedup_lower_bound"] = C2_CAP_S * 1000.0 / gpu_ms
    elif st == "ok":
        res["seconds"] = wall
        res["speedup"] = wall * 1000.0 / gpu_ms
    return res


def cpu_listing_rate(graph, k):
    """k-cliques/s of the oracle restatement's listing (kClist recursion of
    cpuCode/c++/src/helpers.cpp:373-511), one core, on the full graph."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from oracle_ref import oracle_count
    except Exception:
        return None
    t0 = time.time()
    T, _ = oracle_count(graph, k)
    dt = time.time() - t0
    return dict(value=T / dt, unit="k-cliques/s", cores=1, kind="port",
                sample=f"listing only, full graph, {dt:.1f} s", total_cliques=int(T))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2",
                    choices=sorted(WORKLOADS) + ["c4", "c5"],
                    help="c4: batched form; c5: served form")
    ap.add_argument("--graphs", type=int, default=4096, help="c4: graphs")
    ap.add_argument("--chunk", type=int, default=4096, help="c4: graphs per batched call")
    ap.add_argument("--duration", type=float, default=20.0, help="c5: seconds per load")
    ap.add_argument("--loads", default="20,60,180", help="c5: offered requests/s")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--untuned", action="store_true",
                    help="built-in launch-shape rules instead of cds_amd/tune_cache.txt")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (the "
                         "reference point that scales rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.ones(args.calibrate_copy // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)  # one elementwise kernel: reads and writes BYTES
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; multi-GPU runs are not "
              "part of it (each GPU would solve its own graphs independently)",
              file=sys.stderr)
        sys.exit(2)
    import cds_amd

    if args.workload == "c4":
        print(json.dumps(bench_batched(args)))
        return
    if args.workload == "c5":
        print(json.dumps(bench_serving(args)))
        return
    w = WORKLOADS[args.workload]
    t0 = time.time()
    g = make_graph(w)
    print(f"[bench] generated n={g.n} m={g.m} in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    eng = cds_amd.Engine(device=0, tune_cache="" if args.untuned else None)
    dg = eng.upload(g)
    for i in range(args.warmup):
        r = eng.solve(dg, w["k"])
        print(f"[bench] warmup {i}: {r['times']['total']:.1f} ms", file=sys.stderr,
              flush=True)
    lat, results = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        r = eng.solve(dg, w["k"])
        lat.append((time.perf_counter() - s0) * 1000)
        results.append(r)
        print(f"[bench] step {i}: {lat[-1]:.1f} ms", file=sys.stderr, flush=True)
    total_s = time.perf_counter() - t_start
    r = results[-1]
    T = r["total_cliques"]

    def avg(key):
        return sum(x[key] for x in results) / len(results)

    # per-kernel HIP-event times measured inside the engine on its stream;
    # the peel runs as one first pass plus one launch per re-peeled coarse
    # step, so its per-launch figures are the solve's totals over launches
    cnt_ms = avg("t_list_count_ms")
    peel_launches = avg("peel_launches") or 1
    list_ms = sum(x["times"]["dag"] + x["times"]["listing"] for x in results) / len(results)
    write_ms = avg("t_list_write_ms")
    # networks of 2^32 clique-member slots or more run the wide flow kernel
    wide = r["core_cliques"] * w["k"] >= 2 ** 32
    kernels = {}
    if w["k"] == 3 and r["list_staged"]:
        # one listing launch that also stages every triangle (12 bytes
        # each), then a copy of the staged runs into place (read + write)
        kernels["k_list3"] = {
            "launches_per_solve": 1, "total_ms_per_solve": cnt_ms,
            "algorithmic_bytes_per_solve":
                list3_bytes(g.n, r["dag_arcs"], r["list_work"]) + 4 * w["k"] * T}
        kernels["k_list_unstage"] = {
            "launches_per_solve": 1, "total_ms_per_solve": write_ms,
            "algorithmic_bytes_per_solve": 2 * 4 * w["k"] * T}
    elif w["k"] == 3:
        # two launches per solve: the count pass, then the write pass that
        # also stores every clique (12 bytes per triangle more)
        kernels["k_list3"] = {
            "launches_per_solve": 2 if write_ms else 1,
            "total_ms_per_solve": cnt_ms + write_ms,
            "algorithmic_bytes_per_solve":
                list3_bytes(g.n, r["dag_arcs"], r["list_work"]) * (2 if write_ms else 1)
                + (4 * w["k"] * T if write_ms else 0)}
    else:
        kernels["k_list4 (count)"] = {"launches_per_solve": 1, "total_ms_per_solve": cnt_ms,
                                      "algorithmic_bytes_per_solve": None}
    if r["dense_peel"]:
        # the bitset peel of the pruned core: two launches per removal
        kernels["k_dense_argmin + k_dense_dec"] = {
            "launches_per_solve": 2 * r["prune_vertices"],
            "total_ms_per_solve": avg("t_peel_kernel_ms"),
            "algorithmic_bytes_per_solve": None}
    else:
        kernels["k_peel_persistent"] = {
            "launches_per_solve": peel_launches,
            "total_ms_per_solve": avg("t_peel_kernel_ms"),
            "algorithmic_bytes_per_solve": peel_bytes(g.n, T, w["k"])}
        kernels["k_build_inc"] = {"launches_per_solve": 1,
                                  "total_ms_per_solve": avg("t_incidence_ms"),
                                  # read the cliques, write k entries of k words each
                                  "algorithmic_bytes_per_solve":
                                      4 * w["k"] * T + 4 * w["k"] * w["k"] * T}
    if r["flow_implicit"] and r["flow_implicit"] == r["flow_launches"]:
        kernels["k_grid_flow_imp_wide" if wide else "k_grid_flow_imp"] = {
            "launches_per_solve": avg("flow_launches") or 1,
            "total_ms_per_solve": avg("t_flow_kernel_ms"),
            "algorithmic_bytes_per_solve": flow_bytes_imp(r, w["k"])}
    else:
        kernels["k_grid_flow_wide" if wide else "k_grid_flow"] = {
            "launches_per_solve": avg("flow_launches") or 1,
            "total_ms_per_solve": avg("t_flow_kernel_ms"),
            "algorithmic_bytes_per_solve": (flow_bytes_wide if wide else flow_bytes)(r, w["k"])}
    if avg("t_replay_ms"):
        # exact replay of the levels the answer depends on; its bytes depend
        # on the phase (a pass over the phase's incidence, or over bitset
        # rows, per removal), not counted here
        bits = avg("replay_bits_phases") + avg("refine_bits_steps")
        grid = avg("replay_grid_phases")
        name = ("k_replay_bits" if bits and not grid else
                "k_replay_grid" if grid and not bits else
                "k_replay_bits + k_replay_grid" if bits else "k_replay")
        kernels[name] = {"launches_per_solve": max(1, avg("replay_phases")),
                         "total_ms_per_solve": avg("t_replay_ms"),
                         "algorithmic_bytes_per_solve": None}
    for kv in kernels.values():
        t = kv["total_ms_per_solve"]
        if kv["algorithmic_bytes_per_solve"] is None:
            kv["avg_launch_ms"] = t / kv["launches_per_solve"]
            kv["algorithmic_bytes"] = kv["achieved_GBps"] = None
            continue
        kv["avg_launch_ms"] = t / kv["launches_per_solve"]
        kv["algorithmic_bytes"] = kv["algorithmic_bytes_per_solve"] / kv["launches_per_solve"]
        kv["achieved_GBps"] = (kv["algorithmic_bytes_per_solve"] / (t / 1000.0) / 1e9
                               if t > 0 else None)
    dom = max((name for name in kernels if kernels[name]["algorithmic_bytes"]),
              key=lambda name: kernels[name]["total_ms_per_solve"])
    alg = kernels[dom]["algorithmic_bytes"]
    dom_ms = kernels[dom]["avg_launch_ms"]
    achieved = kernels[dom]["achieved_GBps"]
    out = {
        "metric": METRIC,
        "value": args.steps / total_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1000 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": w["text"], "n": g.n, "m": g.m, "k": w["k"],
                   "seed": w["seed"], "total_cliques": T,
                   "params": "pSize=8000 cpSize=16000 glBufferSize=60000 "
                             "partitionSize=30000 earlyStop=-1"},
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "k_cliques_per_s": T / (cnt_ms / 1000.0) if cnt_ms else None,
        "k_cliques_per_s_materialised": T / (list_ms / 1000.0) if list_ms else None,
        "stage_ms": r["times"],
        "answer": {"density": r["density"], "size": r["size"],
                   "exact": f"{r['set_cliques']}/{r['set_size']}",
                   "k_prime": r["k_prime"], "core_size": r["core_size"],
                   "components": r["num_components"],
                   "flow_networks": r["flow_networks_solved"],
                   "alpha_steps": len(r["trace_alpha"])},
        "roofline": {"kernel": dom, "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": read_traffic(dom, args.workload),
                     "algorithmic_bytes": alg, "avg_launch_ms": dom_ms},
        "kernels": kernels,
        "tuning": dict(eng.get_tuning(w["k"], g.m),
                       cache=os.path.relpath(cds_amd.TUNE_CACHE, ROOT)
                       if os.path.exists(cds_amd.TUNE_CACHE) else None),
        "peel": {"level_table_entries": r["peel_distinct_levels"],
                 "sub_rounds": r["peel_levels"], "workgroups": r["peel_grid"],
                 "coarse_shift": r["peel_coarse_shift"],
                 "refined_steps": r["peel_refined_steps"],
                 "replay_phases": r["replay_phases"], "replay_ms": avg("t_replay_ms"),
                 "replay_bits_phases": r["replay_bits_phases"],
                 "refine_bits_steps": r["refine_bits_steps"],
                 # one thread's wall-clock split of the persistent launches
                 "split_ms": r["peel_split_ms"],
                 # graphs too large to materialise: solved on the prune_tau-core
                 "prune_tau": r["prune_tau"], "prune_vertices": r["prune_vertices"],
                 "prune_passes": r["prune_passes"], "prune_unlinks": r["prune_unlinks"],
                 "dense_peel": r["dense_peel"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        with tempfile.TemporaryDirectory() as td:
            out["cpu_baseline"] = cpu_baseline(td, eng)
            if args.workload == "c2":
                out["cpu_c2_bound"] = cpu_c2_bound(td, g, out["ms_per_step"])
        if w["k"] == 3:
            out["cpu_listing"] = cpu_listing_rate(g, w["k"])
            # the oracle's count of the same graph must equal the engine's
            if out["cpu_listing"] and out["cpu_listing"]["total_cliques"] != T:
                raise SystemExit(f"bench.py: engine T={T} but the oracle counts "
                                 f"{out['cpu_listing']['total_cliques']}")
    dg.free()
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
