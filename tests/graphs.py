"""Graphs used by the tests: closed-form families and seeded random graphs.

Each family comes with its known maximal-biclique count (and, for the crown
graph, the full list).  Graphs are given in file orientation: rows = V.
"""
from __future__ import annotations

import itertools

import numpy as np

from gmbe_amd import BiGraph


def complete(a, b):
    """K_{a,b}: exactly one maximal biclique."""
    return a, [list(range(a)) for _ in range(b)], 1


def matching(n):
    """n disjoint edges: n bicliques."""
    return n, [[i] for i in range(n)], n


def crown(n):
    """K_{n,n} minus a perfect matching: 2^n - 2 bicliques (n >= 2)."""
    return n, [[u for u in range(n) if u != v] for v in range(n)], (2 ** n - 2 if n >= 2 else 0)


def crown_list(n):
    """Closed form: for every non-empty proper subset S of [n],
    (L, R) = ({u_i : i in S}, {v_j : j not in S})."""
    out = set()
    for r in range(1, n):
        for S in itertools.combinations(range(n), r):
            rest = tuple(j for j in range(n) if j not in S)
            out.add((tuple(S), rest))
    return out


def chain(k):
    """Nested neighbourhoods N(v_0) ⊂ ... ⊂ N(v_{k-1}): k bicliques."""
    return k, [list(range(i + 1)) for i in range(k)], k


def star_v(n):
    """One V vertex adjacent to n U vertices: 1 biclique."""
    return n, [list(range(n))], 1


def star_u(n):
    """One U vertex adjacent to n V vertices: 1 biclique."""
    return 1, [[0] for _ in range(n)], 1


def disjoint(*parts):
    """Disjoint union of (nU, rows, count) families: counts add."""
    nU, rows, cnt = 0, [], 0
    for pu, prow, pc in parts:
        rows.extend([[u + nU for u in r] for r in prow])
        nU += pu
        cnt += pc
    return nU, rows, cnt


def random_rows(nU, nV, p, seed):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(nV):
        r = sorted(int(u) for u in np.nonzero(rng.random(nU) < p)[0])
        if not r:
            r = [int(rng.integers(nU))]
        rows.append(r)
    return rows


KATS = {
    "K_3x4": complete(3, 4),
    "K_1x1": complete(1, 1),
    "matching_7": matching(7),
    "crown_2": crown(2),
    "crown_5": crown(5),
    "crown_9": crown(9),
    "chain_6": chain(6),
    "star_v_10": star_v(10),
    "star_u_10": star_u(10),
    "union": disjoint(complete(3, 2), crown(4), matching(3), chain(4)),
}


def kat_graph(name, **opt):
    nU, rows, cnt = KATS[name]
    return BiGraph.from_rows(nU, rows, **opt), cnt


def nested_rows(nU, nV, seed):
    """Rows of which most are subsets of an earlier row (sometimes with one
    extra vertex), shuffled: long chains of neighbourhood containment, which
    PMBE's containment order and its skip rule act on."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(nV):
        if rows and rng.random() < 0.7:
            par = rows[int(rng.integers(len(rows)))]
            keep = rng.random(len(par)) < rng.uniform(0.3, 0.95)
            r = {u for u, k in zip(par, keep) if k}
            if rng.random() < 0.2:
                r.add(int(rng.integers(nU)))
        else:
            r = {int(u) for u in rng.choice(nU, int(rng.integers(1, nU)), replace=False)}
        rows.append(sorted(r))
    return [rows[i] for i in rng.permutation(nV)]


def pmbe_recipe_rows(rec):
    """(nU, rows) of a small seeded graph, at most 16 vertices a side."""
    rng = np.random.default_rng(rec["seed"])
    nU, nV = int(rng.integers(3, 17)), int(rng.integers(3, 17))
    if rec["kind"] == "random":
        return nU, random_rows(nU, nV, float(rng.uniform(0.1, 0.7)), rec["seed"])
    return nU, nested_rows(nU, nV, rec["seed"])


def mbe_orient_rows(nU, rows):
    """The graph as the loader of baselines/MBE.zip sees it (MBE/BiGraph.cpp:
    38-79): each non-empty row is an L vertex (empty rows dropped), the items
    0..max are R vertices, and the sides swap when the largest L degree
    exceeds the largest R degree.  Returned as (|L|, R rows): R is the side
    its finders take their roots from, the V side here."""
    rows_nz = [r for r in rows if r]
    nR = max((max(r) for r in rows_nz), default=-1) + 1
    cols = [[] for _ in range(nR)]
    for i, r in enumerate(rows_nz):
        for u in r:
            cols[u].append(i)
    if max(map(len, rows_nz), default=0) > max(map(len, cols), default=0):
        return nR, rows_nz
    return len(rows_nz), cols


def mbe_orient(nU, rows):
    """mbe_orient_rows as a graph (V = the loader's R side)."""
    return BiGraph.from_rows(*mbe_orient_rows(nU, rows))


def mbe_reorder2(nU, rows):
    """mbe_orient, then R renumbered as BiGraph::Reorder(2) of
    baselines/MBE.zip numbers it before ImbeaAdvFinder runs
    (MBE/BiGraph.cpp:103-140): degree ascending, then the neighbour list
    lexicographically; equal lists keep their order (std::sort sorts up to
    16 elements by insertion in libstdc++)."""
    nL, vrows = mbe_orient_rows(nU, rows)
    order = sorted(range(len(vrows)), key=lambda i: (len(vrows[i]), vrows[i], i))
    return BiGraph.from_rows(nL, [vrows[i] for i in order])


def random_graph(nU, nV, p, seed, **opt):
    return BiGraph.from_rows(nU, random_rows(nU, nV, p, seed), **opt)
