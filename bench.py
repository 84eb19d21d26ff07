"""Benchmark: maximal biclique enumeration on one MI355X.

Workload (a step): the power-law bipartite graph at full size (Chung-Lu
gamma=2.1 on both sides, |U|=100,000, |V|=200,000, 1.5M distinct edges, degree
cap 20,000, seed 2; loaded as -t 1 -o 1 -f; bound_height=20, bound_size=1500),
enumerated through IterFinderGpu6::Execute over the top-level vertices of
processing rank >= 1024, i.e. the 98,976 lightest of the 100,000 task roots
(about 1.9e8 maximal bicliques).  The 1,024 heaviest roots hold the graph's
hub core: they yield more than 4e8 maximal bicliques in the first 40 s of each
of their rank ranges and do not fit a benchmark step; they are measured
beside this one, in rank ranges with the finder's deadline knob.

The graph is resident in HBM before the timed region.  The JSON line reports
maximal bicliques per second over the timed steps, the roofline of the
enumeration kernel (algorithmic bytes per launch / kernel time from HIP events
on the finder's stream, against the 8 TB/s HBM peak) and the CPU baseline:
the restatement of the reference's CPU finder (MbeaFinder with MbeaAdvFinder's
2-hop top level) timed on this host's cores over a bounded sample of the same
top-level vertices.  Beside it, cpu_baseline_imbea and cpu_baseline_pmbe time
the restatements of the authors' faster CPU finders (ImbeaAdvFinder and
PmbeAdvFinder of baselines/MBE.zip) on samples of the same graph.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --workload batch     # 4,096 small graphs: batched vs looped launches
    python bench.py --workload serve     # Poisson requests at 2,000/s: p50/p99 latency
    python bench.py --copy-calibration   # plain 1 GiB device copy, for HBM counters
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("MBE wall time + maximal bicliques/s on 1 MI355X; p50/p99 latency of served "
          "queries")
HBM_PEAK = 8.0e12  # bytes/s, MI355X
ROOT_BEGIN = 1024  # first processing rank of the benchmark step
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "mbe_persistent_config2_hbm_traffic.json")


def workload_name(config: int, scale: float, root_begin: int) -> str:
    base = ("Chung-Lu gamma=2.1 |U|=100,000 |V|=200,000 1.5M edges cap 20,000 seed 2"
            if config == 2 else f"config {config} graph of gmbe_amd/workloads.py")
    if scale != 1.0:
        base += f", sizes x{scale:g}"
    return (f"MBE {base}; -t 1 -o 1 -f; bound_height=20 bound_size=1500; "
            f"top-level vertices of processing rank >= {root_begin}")


def cpu_baseline(graph, root_begin: int, seconds: float, algo: str = "mbea"):
    """The oracle's MbeaFinder restatement on a sample of the step's roots;
    algo="imbea" or "pmbe" runs the authors' ImbeaAdvFinder or PmbeAdvFinder
    restatement instead.  PMBE numbers V by its containment order, so its
    ranks select other vertices than the step's; the order is built once per
    call, before the timed search (its time is reported apart)."""
    import oracle
    run, name = {
        "mbea": (oracle.mbea, "MbeaFinder restatement (oracle/oracle.c orc_mbea"),
        "imbea": (oracle.imbea, "ImbeaAdvFinder restatement (oracle/oracle.c orc_imbea"),
        "pmbe": (oracle.pmbe, "PmbeAdvFinder restatement (oracle/oracle.c orc_pmbe"),
    }[algo]
    threads = min(16, os.cpu_count() or 1)
    stride = 97
    og = graph.oracle_graph()
    maximal, secs, wall, tasks, phases = 0, 0.0, 0.0, 0, 0
    for j in range(1 if algo == "pmbe" else stride):
        t0 = time.perf_counter()
        r = run(og, threads=threads, stride=stride, phase=root_begin + j,
                time_limit=max(1.0, seconds - secs))
        wall += time.perf_counter() - t0
        maximal += r.maximal
        secs += r.seconds
        tasks += r.tasks
        phases += 1
        if secs >= seconds:
            break
    order = (f"; ranks in the containment order, built in {wall - secs:.1f} s before the search"
             if algo == "pmbe" else "")
    return {
        "value": maximal / secs if secs > 0 else 0.0, "unit": "bicliques/s", "cores": threads,
        "kind": "port",
        "sample": (f"{name}, {threads} threads) over "
                   f"{tasks} top-level vertices of the same {'graph' if order else 'step'} "
                   f"(ranks {root_begin}+j+97t, "
                   f"j < {phases}{order}): {maximal} maximal bicliques in {secs:.1f} s"),
    }


def copy_calibration():
    """A plain device copy of 1 GiB (read 1 GiB, write 1 GiB), three times,
    so that the HBM counters of a profiler can be scaled to bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))


def measured_traffic(workload: str):
    """HBM bytes per launch of the enumeration kernel from the committed
    counter summary, when it was taken on this workload."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    return t.get("hbm_bytes_per_launch")


def bench_batch(args):
    """The batched workload: graphs of the small-graph distribution (per-side
    sizes log-uniform in [50, 5000], Chung-Lu gamma=2.2, mean degree uniform
    in [3, 20], seeds 4<<32 | i), split into consecutive groups of
    --group graphs; each group is one batch (one captured persistent launch
    per step).  The autotuner runs on the first group and its choice is used
    for all.  Baseline: the same kernel launched once per graph with a host
    wait each (the looped plain kernel), on the same groups."""
    import numpy as np
    from gmbe_amd import Batch, workloads
    t0 = time.time()
    gs = workloads.batch_graphs(args.n_graphs)
    print(f"# {len(gs)} graphs built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    groups = [gs[i:i + args.group] for i in range(0, len(gs), args.group)]
    batches = [Batch(g, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE) for g in groups]
    t0 = time.time()
    b0 = batches[0]
    if not args.no_autotune:
        b0.set("verbose", 1)
        b0.set("autotune", 1)
    choice = {"waves_per_block": b0.get("waves_per_block"), "lds_kb": b0.get("lds_kb"),
              "bounds_by_class": [[b0.get(f"bound_height.{c}"), b0.get(f"bound_size.{c}")]
                                  for c in range(b0.get("n_classes"))]}
    for b in batches[1:]:
        b.set("waves_per_block", choice["waves_per_block"])
        b.set("lds_kb", choice["lds_kb"])
        for c, (bh, bs) in enumerate(choice["bounds_by_class"]):
            b.set(f"bound_height.{c}", bh)
            b.set(f"bound_size.{c}", bs)
    print(f"# autotune {time.time() - t0:.1f} s: {choice}", file=sys.stderr, flush=True)

    def one_pass(tag):
        per_all, tot_maximal, kern, balg = [], 0, 0.0, 0
        t1 = time.perf_counter()
        for k, b in enumerate(batches):
            per, tot = b.execute()
            per_all += [r.digest() for r in per]
            tot_maximal += tot.maximal
            kern += tot.kernel_time_s
            balg += tot.alg_bytes
            print(f"# {tag} group {k}: {tot.maximal} bicliques, wall {tot.exe_time_s:.3f} s",
                  file=sys.stderr, flush=True)
        return per_all, tot_maximal, time.perf_counter() - t1, kern, balg

    for i in range(args.warmup):
        one_pass(f"warmup {i}")
    runs = [one_pass(f"step {i}") for i in range(args.steps)]
    digests = runs[0][0]
    for r in runs[1:]:
        if r[0] != digests:
            raise SystemExit("bench: batched steps disagree")
    for b in batches:
        b.set("per_graph_launch", 1)
    per_l, max_l, looped, _, _ = one_pass("looped")
    if per_l != digests:
        raise SystemExit("bench: looped per-graph results differ from the batched ones")
    for b in batches:
        b.close()
    total = sum(r[1] for r in runs)
    elapsed = sum(r[2] for r in runs)
    value = total / elapsed
    looped_value = max_l / looped
    kernel_avg = sum(r[3] for r in runs) / len(runs)
    balg = runs[0][4]
    achieved = balg / kernel_avg if kernel_avg > 0 else 0.0
    sizes = np.array([g.stats()["edge_size"] for g in gs])
    line = {
        "metric": METRIC, "value": value, "unit": "bicliques/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded generator, built in-process)",
        "config": {
            "workload": (f"MBE batched: {len(gs)} graphs, per-side sizes log-uniform in [50,5000], "
                         "Chung-Lu gamma=2.2, mean degree U[3,20], seeds 4<<32|i; -t 1 -o 1 -f; "
                         f"{len(batches)} batches of {args.group} graphs, one captured "
                         "persistent launch each"),
            "edges_total": int(sizes.sum()), "edges_max": int(sizes.max()),
            "maximal_bicliques": int(runs[0][1]), "autotune": choice,
            "looped_bicliques_per_s": looped_value, "looped_wall_s": looped,
            "speedup_vs_looped": value / looped_value if looped_value > 0 else None,
            "kernel_time_s": kernel_avg,
        },
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "alg_bytes_per_step": int(balg)},
    }
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        m, secs, k = 0, 0.0, 0
        for g in gs:
            r = oracle.mbea(g.oracle_graph(), threads=threads,
                            time_limit=max(1.0, args.cpu_seconds - secs))
            m += r.maximal
            secs += r.seconds
            k += 1
            if secs >= args.cpu_seconds:
                break
        line["cpu_baseline"] = {
            "value": m / secs if secs > 0 else 0.0, "unit": "bicliques/s", "cores": threads,
            "kind": "port",
            "sample": (f"MbeaFinder restatement (oracle/oracle.c orc_mbea, {threads} threads) over "
                       f"the first {k} graphs of the batch (time-limited): {m} maximal bicliques "
                       f"in {secs:.1f} s")}
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """The served workload: Poisson arrivals at --rate requests/s for
    --duration s (seed 5); half are 2-hop ego subgraphs (at most 20,000 edges)
    of a uniformly random U vertex of the power-law graph at full size, half
    are small graphs of the batched distribution (a pool of --pool seeded
    graphs, seeds 5<<32 | j, drawn uniformly).  Latency = completion on the
    host minus scheduled arrival."""
    import threading
    import numpy as np
    from gmbe_amd import Server, enumerate_graph, workloads
    rng = np.random.default_rng(5)
    t0 = time.time()
    resident = workloads.raw_graph(2, 1.0)
    nU = resident.stats()["U_size"]
    pool = [workloads.small_graph(workloads.serve_small_seed(j)) for j in range(args.pool)]
    n = int(args.rate * args.duration * 1.2) + 16
    gaps = rng.exponential(1.0 / args.rate, n)
    arrive = np.cumsum(gaps)
    arrive = arrive[arrive < args.duration]
    kinds = rng.random(len(arrive)) < 0.5
    us = rng.integers(0, nU, len(arrive))
    js = rng.integers(0, len(pool), len(arrive))
    print(f"# {len(arrive)} requests prepared in {time.time() - t0:.1f} s", file=sys.stderr,
          flush=True)
    s = Server(resident, slots=args.slots, mem_budget=args.mem_budget)
    s.set("request_ms", args.request_ms)
    tickets = [None] * len(arrive)
    errors = []
    base = time.monotonic() + 0.5
    n_sub = 4

    def submitter(k):
        for i in range(k, len(arrive), n_sub):
            delay = base + arrive[i] - time.monotonic()
            if delay > 0:
                time.sleep(delay)
            try:
                tickets[i] = (s.submit_ego(int(us[i]), 20000) if kinds[i]
                              else s.submit_graph(pool[int(js[i])]))
            except Exception as e:  # rejected requests are reported, not retried
                errors.append((i, str(e)))

    ths = [threading.Thread(target=submitter, args=(k,)) for k in range(n_sub)]
    for t in ths:
        t.start()
    done = {}
    last = time.monotonic()
    while any(t.is_alive() for t in ths):
        for d in s.poll():
            done[d.ticket] = d
        if time.monotonic() - last > 10:
            print(f"# {len(done)} requests done", file=sys.stderr, flush=True)
            last = time.monotonic()
        time.sleep(0.002)
    for t in ths:
        t.join()
    want = sum(1 for t in tickets if t is not None)
    while len(done) < want:
        for d in s.poll(wait=True):
            done[d.ticket] = d
        if time.monotonic() - last > 10:
            print(f"# {len(done)} of {want} requests done", file=sys.stderr, flush=True)
            last = time.monotonic()
    sstats = s.stats()
    s.close()
    lat, lat_all, total, t_end, cut = [], [], 0, 0.0, 0
    for i, t in enumerate(tickets):
        if t is None:
            continue
        d = done[t]
        lat_all.append(d.t_done_s - (base + arrive[i]))
        t_end = max(t_end, d.t_done_s)
        if d.status == -10:
            cut += 1
            continue
        if d.status != 0:
            errors.append((i, f"status {d.status}"))
            continue
        lat.append(d.t_done_s - (base + arrive[i]))
        total += d.maximal
    lat = np.array(lat)
    lat_all = np.array(lat_all)
    # spot check: a sample of the requests reported exact equal their
    # stand-alone runs (requests cut by the deadline can take minutes alone)
    exact = [i for i, t in enumerate(tickets) if t is not None and done[t].status == 0]
    check = rng.choice(exact, size=min(args.check, len(exact)), replace=False) if exact else []
    mism = 0
    for k, i in enumerate(check):
        g = resident.ego(int(us[i]), 20000) if kinds[i] else pool[int(js[i])]
        r = enumerate_graph(g.derive(**workloads.RUN_OPTION))
        d = done[tickets[i]]
        mism += (d.maximal, d.hash_sum, d.hash_xor) != r.digest()
        if k % 25 == 24:
            print(f"# checked {k + 1} requests alone", file=sys.stderr, flush=True)
    span = t_end - base
    line = {
        "metric": METRIC, "value": total / span if span > 0 else 0.0, "unit": "bicliques/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": span * 1e3,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded generators, built in-process)",
        "config": {
            "workload": (f"MBE served: Poisson {args.rate:g} req/s for {args.duration:g} s (seed 5); "
                         "50% 2-hop ego subgraphs (<=20,000 edges) of the full-size power-law "
                         f"graph, 50% small graphs from a pool of {len(pool)}; -t 1 -o 1 -f"),
            "requests": int(len(arrive)), "completed_exact": int(len(lat)),
            "cut_by_deadline": int(cut), "errors": len(errors), "request_ms": args.request_ms,
            "latency_p50_ms": float(np.percentile(lat_all, 50) * 1e3) if len(lat_all) else None,
            "latency_p99_ms": float(np.percentile(lat_all, 99) * 1e3) if len(lat_all) else None,
            "latency_max_ms": float(lat_all.max() * 1e3) if len(lat_all) else None,
            "exact_latency_p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "exact_latency_p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "achieved_rate_per_s": len(lat) / span if span > 0 else 0.0,
            "maximal_bicliques": int(total), "checked_alone": int(len(check)),
            "mismatches": int(mism), "slots": args.slots, "mem_budget": args.mem_budget,
            "server": sstats,
        },
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--root-begin", type=int, default=ROOT_BEGIN)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", choices=["plain", "batch", "serve"], default="plain")
    ap.add_argument("--n-graphs", type=int, default=4096)
    ap.add_argument("--group", type=int, default=512)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--pool", type=int, default=1024)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--mem-budget", type=int, default=3 << 30)
    ap.add_argument("--check", type=int, default=200)
    ap.add_argument("--request-ms", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU (asked for --gpus {args.gpus}); "
              "sharding over several GPUs is designed but not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return

    from gmbe_amd import IterFinderGpu6, workloads

    t0 = time.time()
    graph = workloads.run_graph(args.config, args.scale)
    st = graph.stats()
    print(f"# graph: |U|={st['U_size']} |V|={st['V_size']} |E|={st['edge_size']} "
          f"(built in {time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    f = IterFinderGpu6(graph, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE,
                       root_begin=args.root_begin)
    for i in range(args.warmup):
        r = f.execute()
        print(f"# warmup {i}: {r.maximal} bicliques in {r.exe_time_s:.3f} s", file=sys.stderr,
              flush=True)
    results = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        r = f.execute()  # synchronous: ends with the counts on the host
        results.append(r)
        print(f"# step {i}: {r.maximal} bicliques, exe {r.exe_time_s:.3f} s, kernel "
              f"{r.kernel_time_s:.3f} s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    f.close()
    digests = {r.digest() for r in results}
    if len(digests) != 1:
        raise SystemExit(f"bench: steps disagree: {digests}")
    r0 = results[0]
    total = sum(r.maximal for r in results)
    kernel_avg = sum(r.kernel_time_s for r in results) / len(results)
    achieved = r0.alg_bytes / kernel_avg if kernel_avg > 0 else 0.0
    workload = workload_name(args.config, args.scale, args.root_begin)
    line = {
        "metric": METRIC,
        "value": total / elapsed,
        "unit": "bicliques/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, built in-process)",
        "config": {
            "workload": workload,
            "U_size": st["U_size"], "V_size": st["V_size"], "edge_size": st["edge_size"],
            "task_roots": st["V_size"] - min(args.root_begin, st["V_size"]),
            "maximal_bicliques": int(r0.maximal),
            "hash_sum": f"{r0.hash_sum:#018x}", "hash_xor": f"{r0.hash_xor:#018x}",
            "non_maximal": int(r0.non_maximal), "nodes": int(r0.nodes),
            "tasks": int(r0.tasks), "splits": int(r0.splits),
            "exe_time_s": r0.exe_time_s, "kernel_time_s": kernel_avg,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": measured_traffic(workload),
            "alg_bytes_per_launch": int(r0.alg_bytes),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(graph, args.root_begin, args.cpu_seconds)
        # the authors' improved CPU finder on the same sample, for comparison
        line["cpu_baseline_imbea"] = cpu_baseline(graph, args.root_begin, args.cpu_seconds / 3,
                                                  algo="imbea")
        line["cpu_baseline_pmbe"] = cpu_baseline(graph, args.root_begin, args.cpu_seconds / 3,
                                                 algo="pmbe")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
