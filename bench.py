"""Benchmark: maximal biclique enumeration on one MI355X.

Workload: the power-law bipartite graph at full size (Chung-Lu gamma=2.1 on
both sides, |U|=100,000, |V|=200,000, 1.5M distinct edges, degree cap 20,000,
seed 2; loaded as -t 1 -o 1 -f; bound_height=20, bound_size=1500), enumerated
whole through IterFinderGpu6::Execute.  The K timed steps partition its
100,000 top-level vertices: step i runs the vertices of processing rank
r = i (mod K).  So the timed region is exactly one enumeration of the whole
graph, and the line's count and hash are those of the whole graph (checked
against the recorded value in profiles/ when one is there).

The graph is resident in HBM before the timed region.  Untimed warmup: the
first warmup step counts bytes per node (count_bytes) on the vertices of
processing rank 0 mod 4K; further ones run a light sample.  The JSON line
reports maximal bicliques per second over the timed steps, the roofline of
the dominant kernel (the one with the most time: the wide-bitset pass on this
graph): its algorithmic bytes per launch -- the bytes of the bitset lists its
pushes read and write plus the CSR rows its task roots read, per node from
the counted warmup part, times the launch's nodes -- over its average launch
time from HIP events on the finder's stream, the measured HBM traffic per
launch from the committed counter summary, and per kernel the same rate
under the reference's id-list byte model (4 (deg(x) + |L| + |C|) bytes per
push), which this kernel's bitset lists undercut by far; and the CPU
baseline: the restatement of the reference's CPU finder
(MbeaFinder with MbeaAdvFinder's 2-hop top level) on all of this host's
hardware threads over a bounded root sample, with the GPU timed on the same
sample.  cpu_baseline_imbea and cpu_baseline_pmbe time the restatements of
the authors' faster CPU finders (ImbeaAdvFinder, PmbeAdvFinder of
baselines/MBE.zip) on the same sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --only 1             # time only the first of the K parts
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "mbe_persistent_config2_hbm_traffic.json")
DIGEST_FILE = os.path.join(ROOT, "profiles", "config2_full_digest.json")
WORKLOAD = ("MBE Chung-Lu gamma=2.1 |U|=100,000 |V|=200,000 1.5M edges cap 20,000 seed 2; "
            "-t 1 -o 1 -f; bound_height=20 bound_size=1500; whole graph, all 100,000 "
            "top-level vertices, split into K launches by processing rank mod K")
# the two enumeration kernels of a launch: the main pass and the wide-bitset pass
KERNELS = {"main": "mbe_persistent<4, 7168, 4, 4>", "wide": "mbe_persistent<4, 8192, 16, 2>"}
# CPU baseline sample: light top-level vertices the CPU finishes in seconds
CPU_STRIDE, CPU_FIRST = 499, 8192


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(graph, seconds: float, algo: str = "mbea"):
    """The oracle's MbeaFinder restatement over the ranks CPU_FIRST,
    CPU_FIRST + CPU_STRIDE, ... (algo="imbea" or "pmbe": the authors'
    ImbeaAdvFinder or PmbeAdvFinder restatement; PMBE numbers V by its
    containment order, so its ranks select other vertices, and its order is
    built before the timed search) on every hardware thread of this host,
    cut at `seconds`."""
    import oracle
    run, name = {
        "mbea": (oracle.mbea, "MbeaFinder restatement (oracle/oracle.c orc_mbea"),
        "imbea": (oracle.imbea, "ImbeaAdvFinder restatement (oracle/oracle.c orc_imbea"),
        "pmbe": (oracle.pmbe, "PmbeAdvFinder restatement (oracle/oracle.c orc_pmbe"),
    }[algo]
    threads = os.cpu_count() or 1
    og = graph.oracle_graph()
    n_roots = max(0, (graph.stats()["V_size"] - CPU_FIRST + CPU_STRIDE - 1) // CPU_STRIDE)
    t0 = time.perf_counter()
    r = run(og, threads=threads, stride=CPU_STRIDE, phase=CPU_FIRST, time_limit=seconds)
    wall = time.perf_counter() - t0
    complete = r.tasks >= n_roots and r.cut == 0
    order = (f"; ranks in the containment order, built in {wall - r.seconds:.1f} s before the "
             "search" if algo == "pmbe" else "")
    return {
        "value": r.maximal / r.seconds if r.seconds > 0 else 0.0, "unit": "bicliques/s",
        "cores": threads, "kind": "port", "cpu": cpu_model(),
        "sample": (f"{name}, {threads} threads) over top-level vertices of processing rank "
                   f"{CPU_FIRST} + {CPU_STRIDE} t ({n_roots} vertices{order}): "
                   + (f"all done, {r.maximal} maximal bicliques in {r.seconds:.1f} s" if complete
                      else f"cut by the time limit after {r.seconds:.1f} s with {r.maximal} "
                           f"maximal bicliques ({r.tasks} vertices started)")),
        "complete": bool(complete), "digest": [int(r.maximal), int(r.hash_sum), int(r.hash_xor)],
    }


def gpu_on_cpu_sample(graph):
    """The GPU on the CPU baseline's root sample (same vertices, same digest)."""
    from gmbe_amd import IterFinderGpu6, workloads
    f = IterFinderGpu6(graph, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE,
                       root_begin=CPU_FIRST, root_stride=CPU_STRIDE,
                       root_phase=CPU_FIRST % CPU_STRIDE)
    f.execute()
    r = f.execute()
    f.close()
    return r


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


def measured_traffic(kernel: str):
    """HBM bytes per launch of one enumeration kernel ("main" or "wide") from
    the committed counter summary, when it was taken on this workload."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("workload") != WORKLOAD:
        return None
    return t.get("hbm_bytes_per_launch", {}).get(kernel)


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
    ncls = b0.get("n_classes")
    choice = {"variant_by_class": [[b0.get(f"waves_per_block.{c}"), b0.get(f"lds_kb.{c}")]
                                   for c in range(ncls)],
              "bounds_by_class": [[b0.get(f"bound_height.{c}"), b0.get(f"bound_size.{c}")]
                                  for c in range(ncls)]}
    for b in batches[1:]:
        for c, (w, l) in enumerate(choice["variant_by_class"]):
            b.set(f"waves_per_block.{c}", w)
            b.set(f"lds_kb.{c}", l)
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
                         f"{len(batches)} batches of {args.group} graphs, one captured run "
                         "each (a persistent launch pair per size class)"),
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
        threads = os.cpu_count() or 1
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--only", type=int, default=0, help="time only the first N of the K parts")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
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
    graph = workloads.run_graph(2)
    st = graph.stats()
    print(f"# graph: |U|={st['U_size']} |V|={st['V_size']} |E|={st['edge_size']} "
          f"(built in {time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    K = max(1, args.steps)
    f = IterFinderGpu6(graph, workloads.BOUND_HEIGHT, workloads.BOUND_SIZE)
    # warmup: the byte model of each kernel on a 1/(4K) part of the graph
    # (processing ranks 0 mod 4K, counted with count_bytes), then light samples
    model = None
    for i in range(args.warmup):
        if i == 0:
            f.set("root_stride", 4 * K)
            f.set("root_phase", 0)
            f.set("root_begin", 0)
            f.set("count_bytes", 1)
        else:
            f.set("root_stride", 97)
            f.set("root_phase", 1024 % 97)
            f.set("root_begin", 1024)
            f.set("count_bytes", 0)
        r = f.execute()
        if i == 0:
            sw = f.stats()
            mn, mb, ml = sw["main_nodes"], sw["main_alg_bytes"], sw["main_list_bytes"]
            wn = max(1, r.nodes - mn)
            model = {"main": {"list": ml / max(1, mn), "id_list": mb / max(1, mn)},
                     "wide": {"list": (sw["list_bytes"] - ml) / wn,
                              "id_list": (r.alg_bytes - mb) / wn},
                     "sample": f"processing ranks 0 mod {4 * K}: {r.nodes} nodes, "
                               f"{sw['list_bytes']} list bytes, {r.alg_bytes} bytes of the "
                               "id-list model"}
        print(f"# warmup {i}: {r.maximal} bicliques in {r.exe_time_s:.3f} s", file=sys.stderr,
              flush=True)
    f.set("count_bytes", 0)
    f.set("root_begin", 0)
    f.set("root_stride", K)
    steps = []
    parts = K if args.only <= 0 else min(K, args.only)
    t_start = time.perf_counter()
    for i in range(parts):
        f.set("root_phase", i)
        r = f.execute()  # synchronous: ends with the counts on the host
        sts = f.stats()
        steps.append({"digest": r.digest(), "nodes": r.nodes, "main_nodes": sts["main_nodes"],
                      "main_s": sts["main_kernel_ns"] * 1e-9,
                      "wide_s": sts["wide_kernel_ns"] * 1e-9, "exe_s": r.exe_time_s})
        print(f"# step {i}: {r.maximal} bicliques, exe {r.exe_time_s:.3f} s, main kernel "
              f"{sts['main_kernel_ns'] * 1e-9:.3f} s, wide pass {sts['wide_kernel_ns'] * 1e-9:.3f} s",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    f.close()
    total = sum(x["digest"][0] for x in steps)
    hsum = sum(x["digest"][1] for x in steps) % (1 << 64)
    hxor = 0
    for x in steps:
        hxor ^= x["digest"][2]
    nodes = sum(x["nodes"] for x in steps)
    recorded = None
    if parts == K:
        try:
            with open(DIGEST_FILE) as fh:
                d = json.load(fh)
            recorded = [d["maximal"], int(d["hash_sum"], 16), int(d["hash_xor"], 16)] == \
                [total, hsum, hxor]
        except (OSError, ValueError, KeyError):
            recorded = None
    # per kernel: average launch time (HIP events) and nodes per launch; the
    # dominant kernel (most time) carries the roofline
    kern = {
        "main": {"s": sum(x["main_s"] for x in steps),
                 "nodes": sum(x["main_nodes"] for x in steps), "launches": len(steps)},
        "wide": {"s": sum(x["wide_s"] for x in steps),
                 "nodes": sum(x["nodes"] - x["main_nodes"] for x in steps),
                 "launches": sum(1 for x in steps if x["wide_s"] > 0)},
    }
    for name, kd in kern.items():
        n = max(1, kd["launches"])
        kd["avg_launch_s"] = kd["s"] / n
        kd["nodes_per_launch"] = kd["nodes"] / n
        for b in ("list", "id_list"):
            bpn = model[name][b] if model else None
            per_launch = bpn * kd["nodes_per_launch"] if bpn is not None else None
            kd[f"{b}_bytes_per_node"] = bpn
            kd[f"{b}_bytes_per_launch"] = per_launch
            kd[f"{b}_GBs"] = (per_launch / kd["avg_launch_s"] / 1e9
                              if per_launch is not None and kd["avg_launch_s"] > 0 else None)
    dom = max(kern, key=lambda k: kern[k]["s"])
    kd = kern[dom]
    achieved = kd["list_GBs"] * 1e9 if kd["list_GBs"] else 0.0
    traffic = measured_traffic(dom)
    line = {
        "metric": METRIC,
        "value": total / elapsed,
        "unit": "bicliques/s",
        "n_gpus": 1,
        "steps": parts,
        "warmup": args.warmup,
        "ms_per_step": elapsed / parts * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, built in-process)",
        "config": {
            "workload": WORKLOAD if parts == K else WORKLOAD + f"; only the first {parts} of {K}",
            "U_size": st["U_size"], "V_size": st["V_size"], "edge_size": st["edge_size"],
            "parts": K, "parts_timed": parts,
            "maximal_bicliques": int(total), "hash_sum": f"{hsum:#018x}",
            "hash_xor": f"{hxor:#018x}", "matches_recorded_digest": recorded,
            "nodes": int(nodes), "nodes_per_s": nodes / elapsed,
            "exe_time_s_total": sum(x["exe_s"] for x in steps),
            "kernels": kern,
        },
        "roofline": {
            "bound": "hbm", "kernel": KERNELS[dom], "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_bandwidth_GBs": (traffic / kd["avg_launch_s"] / 1e9)
            if traffic and kd["avg_launch_s"] else None,
            "traffic_frac": (traffic / kd["avg_launch_s"] / HBM_PEAK)
            if traffic and kd["avg_launch_s"] else None,
            "byte_model": model,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(graph, args.cpu_seconds)
        g = gpu_on_cpu_sample(graph)
        line["gpu_on_cpu_sample"] = {
            "value": g.maximal / g.kernel_time_s if g.kernel_time_s > 0 else 0.0,
            "unit": "bicliques/s", "kernel_time_s": g.kernel_time_s,
            "digest_matches_cpu": line["cpu_baseline"]["complete"] and
            [int(g.maximal), int(g.hash_sum), int(g.hash_xor)] == line["cpu_baseline"]["digest"],
        }
        line["cpu_baseline_imbea"] = cpu_baseline(graph, args.cpu_seconds / 3, algo="imbea")
        line["cpu_baseline_pmbe"] = cpu_baseline(graph, args.cpu_seconds / 3, algo="pmbe")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
