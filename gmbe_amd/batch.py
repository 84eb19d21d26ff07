"""Python face of the batched form and the serving loop (include/gmbe.h).

    b = Batch([g0, g1, ...], bound_height=20, bound_size=1500)
    b.set("autotune", 1)                 # optional: measured choice of variant/thresholds
    per_graph, total = b.execute()

    s = Server(resident_graph, slots=256, mem_budget=1 << 30)
    t = s.submit_ego(u, 20000); t2 = s.submit_graph(g)
    done = s.poll(wait=True)             # [ServeDone, ...]

Both run only in the HIP kernels of libgmbe.so; without a GPU the
constructors raise GmbeError(GMBE_ERR_HIP).
"""
from __future__ import annotations

import ctypes as C

from ._lib import Result, ServeDone, check, lib


class Batch:
    def __init__(self, graphs, bound_height: int = 20, bound_size: int = 1500, **knobs):
        self.graphs = list(graphs)  # borrowed by the batch: keep them alive
        arr = (C.c_void_p * len(self.graphs))(*[g.handle.value for g in self.graphs])
        out = C.c_void_p()
        check(lib().gmbe_batch_create(arr, len(self.graphs), bound_height, bound_size,
                                      C.byref(out)))
        self._h = out
        for k, v in knobs.items():
            self.set(k, v)

    def set(self, key: str, value: int):
        check(lib().gmbe_batch_set(self._h, key.encode(), int(value)))

    def get(self, key: str) -> int:
        v = C.c_int64()
        check(lib().gmbe_batch_get(self._h, key.encode(), C.byref(v)))
        return int(v.value)

    def execute(self):
        per = (Result * len(self.graphs))()
        tot = Result()
        check(lib().gmbe_batch_execute(self._h, per, C.byref(tot)))
        return list(per), tot

    def close(self):
        if self._h and self._h.value:
            lib().gmbe_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    def __init__(self, resident=None, slots: int = 256, mem_budget: int = 1 << 30,
                 bound_height: int = 20, bound_size: int = 1500):
        self.resident = resident
        out = C.c_void_p()
        check(lib().gmbe_serve_create(resident.handle if resident is not None else None, slots,
                                      mem_budget, bound_height, bound_size, C.byref(out)))
        self._h = out

    def set(self, key: str, value: int):
        check(lib().gmbe_serve_set(self._h, key.encode(), int(value)))

    def stats(self) -> dict:
        out = (C.c_uint64 * 8)()
        n = lib().gmbe_serve_stats(self._h, out, 8)
        names = ["launches", "launched", "host_us", "gpu_us", "waiting"]
        return {k: int(out[i]) for i, k in enumerate(names[:n])}

    def submit_graph(self, g) -> int:
        t = C.c_uint64()
        check(lib().gmbe_serve_submit_graph(self._h, g.handle, C.byref(t)))
        return int(t.value)

    def submit_ego(self, u: int, max_edges: int = 20000) -> int:
        t = C.c_uint64()
        check(lib().gmbe_serve_submit_ego(self._h, u, max_edges, C.byref(t)))
        return int(t.value)

    def poll(self, max_n: int = 4096, wait: bool = False):
        buf = (ServeDone * max_n)()
        n = C.c_int()
        check(lib().gmbe_serve_poll(self._h, buf, max_n, 1 if wait else 0, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def close(self):
        if self._h and self._h.value:
            lib().gmbe_serve_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
