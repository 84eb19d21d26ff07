"""The benchmark workloads: seeded synthetic bipartite graphs.

Config numbering follows the five benchmark configurations:
  1  uniform |U|=2,000 |V|=1,500, 12,000 distinct edges, seed 1
  2  Chung-Lu gamma=2.1, |U|=100,000 |V|=200,000, 1.5M edges, cap 20,000, seed 2
  3  Chung-Lu gamma=2.0, |U|=50,000 |V|=1,000,000, 8M edges, cap 20,000,
     + 200 planted a x b blocks (a, b in [10, 40]) with 5% of block edges removed,
     seed 3
  4  4,096 small Chung-Lu graphs (gamma=2.2), per-side sizes log-uniform in
     [50, 5,000], mean degree uniform in [3, 20]; graph i uses seed (4 << 32) | i
  5  serving stream: Poisson arrivals at 2,000 req/s for 60 s (seed 5); half
     config-4-style small graphs (seed (5 << 32) | j), half 2-hop ego subgraphs
     of a random U vertex of config 2's graph, capped at 20,000 edges
All enumerations run with uvtrans=1, order=1, fast_mode=1 (the reference
scripts' `-t 1 -o 1 -f`), bound_height=20, bound_size=1500.
"""
from __future__ import annotations

from .graph import BiGraph, GenParams

RUN_OPTION = dict(uvtrans=1, order=1, fast_mode=1)
BOUND_HEIGHT = 20
BOUND_SIZE = 1500


def gen_params(config: int, scale: float = 1.0) -> GenParams:
    """Generator parameters of config 1..3 (scale < 1 shrinks sizes for tests)."""
    p = GenParams()
    if config == 1:
        p.kind, p.nU, p.nV, p.m, p.seed = 1, 2000, 1500, 12000, 1
    elif config == 2:
        p.kind, p.nU, p.nV, p.m = 2, 100_000, 200_000, 1_500_000
        p.gamma_u = p.gamma_v = 2.1
        p.cap, p.seed = 20_000, 2
    elif config == 3:
        p.kind, p.nU, p.nV, p.m = 3, 50_000, 1_000_000, 8_000_000
        p.gamma_u = p.gamma_v = 2.0
        p.cap, p.seed = 20_000, 3
        p.blocks, p.block_min, p.block_max, p.drop = 200, 10, 40, 0.05
    else:
        raise ValueError(f"config {config} is not a single generated graph")
    if scale != 1.0:
        p.nU = max(64, int(p.nU * scale))
        p.nV = max(64, int(p.nV * scale))
        p.m = max(max(p.nU, p.nV) + 1, int(p.m * scale))
        if p.cap:
            # keep the cap feasible: the smaller side must be able to hold m edges
            p.cap = max(16, int(p.cap * scale), -(-4 * p.m // min(p.nU, p.nV)))
        if p.blocks:
            p.blocks = max(1, int(p.blocks * scale))
            p.block_max = min(p.block_max, p.nU, p.nV)
            p.block_min = min(p.block_min, p.block_max)
    return p


def raw_graph(config: int, scale: float = 1.0) -> BiGraph:
    """Config graph in file orientation (rows = V side of the .adj file)."""
    return BiGraph.generate(gen_params(config, scale))


def run_graph(config: int, scale: float = 1.0) -> BiGraph:
    """Config graph after `-t 1 -o 1 -f`, ready for enumeration."""
    return raw_graph(config, scale).derive(**RUN_OPTION)


def small_graph(seed: int) -> BiGraph:
    """One graph of the batched workload (config 4 distribution)."""
    return BiGraph.generate_small(seed, 50, 5000, 3.0, 20.0, 2.2)


def batch_seed(i: int) -> int:
    return (4 << 32) | i


def batch_graphs(n: int = 4096, first: int = 0) -> list:
    """Graphs first..first+n-1 of the batched workload, in working form."""
    return [small_graph(batch_seed(i)).derive(**RUN_OPTION) for i in range(first, first + n)]


def serve_small_seed(j: int) -> int:
    return (5 << 32) | j
