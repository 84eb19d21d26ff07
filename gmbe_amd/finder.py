"""Python face of the finder API (IterFinderGpu6, src/IterFinderGpu.h:77-90).

    g = BiGraph.read_adj("x.adj", uvtrans=1, order=1, fast_mode=1)
    f = IterFinderGpu6(g, bound_height=20, bound_size=1500)
    r = f.execute()          # Result: maximal, non_maximal, hash_sum, hash_xor, ...
    f.print_result()         # the reference's stdout lines

The enumeration always runs in the HIP kernels of libgmbe.so; there is no CPU
fallback, and creating a finder without a GPU raises GmbeError(GMBE_ERR_HIP).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Result, check, lib
from .graph import BiGraph


class IterFinderGpu6:
    def __init__(self, graph: BiGraph, bound_height: int = 20, bound_size: int = 1500,
                 sel: int = 2, ngpus: int = 1, **knobs):
        self.graph = graph  # borrowed: keep the graph alive as long as the finder
        out = C.c_void_p()
        check(lib().gmbe_finder_create(graph.handle, sel, bound_height, bound_size, ngpus,
                                       C.byref(out)))
        self._h = out
        self.result = None
        for k, v in knobs.items():
            self.set(k, v)

    def set(self, key: str, value: int):
        check(lib().gmbe_finder_set(self._h, key.encode(), int(value)))

    def execute(self) -> Result:
        r = Result()
        rc = lib().gmbe_finder_execute(self._h, C.byref(r))
        self.result = r  # partial counts stay readable after GMBE_ERR_DEADLINE
        check(rc)
        return r

    def bicliques(self):
        """List-mode records as a set of (A, B) tuples of sorted original ids."""
        used = C.c_size_t()
        check(lib().gmbe_finder_list(self._h, None, 0, C.byref(used)))
        n = used.value
        buf = np.zeros(max(n, 1), np.uint32)
        check(lib().gmbe_finder_list(self._h, buf.ctypes.data, n, C.byref(used)))
        return records_to_set(buf[:n])

    def exit_times(self) -> list:
        """Per-block exit times (s) of the last run with exit_stamps=1."""
        n = C.c_size_t()
        check(lib().gmbe_finder_exit_times(self._h, None, 0, C.byref(n)))
        buf = (C.c_double * max(n.value, 1))()
        check(lib().gmbe_finder_exit_times(self._h, buf, n.value, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    STAT_NAMES = ["cyc_wait", "cyc_prefix", "cyc_expand", "cyc_extract", "cyc_build", "cyc_dfs",
                  "cyc_split", "edges", "entries", "ovf_arena", "ovf_wide", "ovf_defer",
                  "entries_w1", "nodes_w1", "cyc_tail", "sub_pass", "sub_merge",
                  "sub_compress", "sub_publish", "publishes", "merges", "abandoned", "late",
                  "sub_reduce", "dropped", "skipped", "list_bytes", "main_kernel_ns",
                  "wide_kernel_ns", "wide_tasks", "main_nodes", "main_maximal",
                  "main_alg_bytes", "main_list_bytes"]

    def stats(self) -> dict:
        out = (C.c_uint64 * 64)()
        n = lib().gmbe_finder_stats(self._h, out, 64)
        return {k: int(out[i]) for i, k in enumerate(self.STAT_NAMES[:n])}

    def print_result(self):
        r = self.result
        print(f"non_maximal: {r.non_maximal}")
        print(f"Total processing time: {r.exe_time_s:f} seconds")
        print(f"maximal nodes:{r.maximal}")
        print()

    def close(self):
        if self._h and self._h.value:
            lib().gmbe_finder_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def records_to_set(buf) -> set:
    """[nA, nB, A..., B...]* -> {(tuple(sorted A), tuple(sorted B))}."""
    out = []
    i, n = 0, len(buf)
    while i < n:
        a, b = int(buf[i]), int(buf[i + 1])
        A = tuple(sorted(int(x) for x in buf[i + 2:i + 2 + a]))
        B = tuple(sorted(int(x) for x in buf[i + 2 + a:i + 2 + a + b]))
        out.append((A, B))
        i += 2 + a + b
    s = set(out)
    if len(s) != len(out):
        raise AssertionError(f"duplicate bicliques in list: {len(out) - len(s)}")
    return s


def enumerate_graph(graph: BiGraph, bound_height=20, bound_size=1500, sel=2, **knobs) -> Result:
    f = IterFinderGpu6(graph, bound_height, bound_size, sel=sel, **knobs)
    try:
        return f.execute()
    finally:
        f.close()


def execute_static(graph: BiGraph, bound_height=20, bound_size=1500) -> Result:
    """-s 2 enumeration through gmbe_finder_execute_static: no device memory
    is allocated; every buffer comes out of the library's static 128 MiB
    array.  Calls that overlap on one GPU race (the array is shared)."""
    r = Result()
    check(lib().gmbe_finder_execute_static(graph.handle, bound_height, bound_size,
                                           C.byref(r)))
    return r
