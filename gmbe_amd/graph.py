"""Host-side graph handle: the Python face of CSRBiGraph (src/BiGraph.h:23-38)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import GenParams, Option, check, lib


def make_option(uvtrans: int = 0, order: int = 1, fast_mode: int = 0) -> Option:
    """== Option (src/BiGraph.h:13-22); defaults as the reference."""
    return Option(uvtrans, order, fast_mode)


class BiGraph:
    """Owns a gmbe_graph*.  Internal orientation: V = enumerated side."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    # -- construction -------------------------------------------------------
    @classmethod
    def read_adj(cls, path: str, uvtrans=0, order=1, fast_mode=0) -> "BiGraph":
        """CSRBiGraph::Read (src/BiGraph.cu:13-278)."""
        out = C.c_void_p()
        opt = make_option(uvtrans, order, fast_mode)
        check(lib().gmbe_graph_read_adj(path.encode(), C.byref(opt), C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_csr(cls, nU: int, v_off, v_adj, uvtrans=0, order=0, fast_mode=1) -> "BiGraph":
        v_off = np.ascontiguousarray(v_off, dtype=np.int64)
        v_adj = np.ascontiguousarray(v_adj, dtype=np.uint32)
        nV = v_off.size - 1
        out = C.c_void_p()
        opt = make_option(uvtrans, order, fast_mode)
        check(lib().gmbe_graph_from_csr(nU, nV, v_off.ctypes.data,
                                        v_adj.ctypes.data if v_adj.size else None,
                                        C.byref(opt), C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_rows(cls, nU: int, rows, **kw) -> "BiGraph":
        off = [0]
        adj = []
        for r in rows:
            adj.extend(r)
            off.append(len(adj))
        return cls.from_csr(nU, np.array(off, np.int64), np.array(adj, np.uint32), **kw)

    @classmethod
    def generate(cls, p: GenParams) -> "BiGraph":
        out = C.c_void_p()
        check(lib().gmbe_generate(C.byref(p), C.byref(out)))
        return cls(out.value)

    @classmethod
    def generate_small(cls, seed, lo, hi, dlo, dhi, gamma) -> "BiGraph":
        out = C.c_void_p()
        check(lib().gmbe_generate_small(seed, lo, hi, dlo, dhi, gamma, C.byref(out)))
        return cls(out.value)

    def derive(self, uvtrans=0, order=1, fast_mode=0) -> "BiGraph":
        out = C.c_void_p()
        opt = make_option(uvtrans, order, fast_mode)
        check(lib().gmbe_graph_derive(self._h, C.byref(opt), C.byref(out)))
        return BiGraph(out.value)

    def ego(self, u: int, max_edges: int) -> "BiGraph":
        out = C.c_void_p()
        check(lib().gmbe_ego_subgraph(self._h, u, max_edges, C.byref(out)))
        return BiGraph(out.value)

    def write_adj(self, path: str):
        check(lib().gmbe_graph_write_adj(self._h, path.encode()))

    def write_edges(self, path: str):
        """Edge-list file ("rows cols" header, one "v u" line per edge) in the
        original orientation and ids (preprocess/adj_to_graph.cpp:31-39)."""
        check(lib().gmbe_graph_write_edges(self._h, path.encode()))

    # -- access ---------------------------------------------------------------
    @property
    def handle(self):
        return self._h

    def stats(self) -> dict:
        a = (C.c_int64 * 8)()
        check(lib().gmbe_graph_stats(self._h, a))
        keys = ["U_size", "V_size", "edge_size", "U_degree", "V_degree", "U_2degree",
                "V_2degree", "swapped"]
        return dict(zip(keys, [int(x) for x in a]))

    def csr(self):
        """Copies of (v_off, v_adj, u_off, u_adj, u_orig, v_orig)."""
        s = self.stats()
        nU, nV, E = s["U_size"], s["V_size"], s["edge_size"]
        ptrs = [C.c_void_p() for _ in range(6)]
        check(lib().gmbe_graph_csr(self._h, *[C.byref(p) for p in ptrs]))

        def arr(p, n, ct, dt):
            if n == 0 or not p.value:
                return np.zeros(n, dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()

        return (arr(ptrs[0], nV + 1, C.c_int64, np.int64),
                arr(ptrs[1], E, C.c_uint32, np.uint32),
                arr(ptrs[2], nU + 1, C.c_int64, np.int64),
                arr(ptrs[3], E, C.c_uint32, np.uint32),
                arr(ptrs[4], nU, C.c_uint32, np.uint32),
                arr(ptrs[5], nV, C.c_uint32, np.uint32))

    def oracle_graph(self):
        """Tuple consumed by the CPU oracle: (nU, nV, v_off, v_adj, u_orig, v_orig, swapped)."""
        s = self.stats()
        v_off, v_adj, _, _, u_orig, v_orig = self.csr()
        return (s["U_size"], s["V_size"], v_off, v_adj, u_orig, v_orig, s["swapped"])

    def close(self):
        if self._h and self._h.value:
            lib().gmbe_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
