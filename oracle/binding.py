"""ctypes binding of oracle/liboracle.so (built from oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: the checker, never the thing measured or shipped.
Every function takes a graph in internal orientation as numpy arrays:
``v_off`` (int64, nV+1), ``v_adj`` (uint32, ascending rows), ``u_orig`` /
``v_orig`` (uint32 internal->original ids) and ``swapped``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OracleResult(C.Structure):
    _fields_ = [
        ("maximal", C.c_uint64), ("non_maximal", C.c_uint64),
        ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64),
        ("nodes", C.c_uint64), ("b_alg", C.c_uint64), ("tasks", C.c_uint64),
        ("splits", C.c_uint64), ("skipped", C.c_uint64), ("cut", C.c_uint64),
        ("seconds", C.c_double),
    ]

    def digest(self):
        return (int(self.maximal), int(self.hash_sum), int(self.hash_xor))


_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def _l():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        l = C.CDLL(LIB_PATH)
        P = C.c_void_p
        graph = [C.c_uint32, C.c_uint32, P, P, P, P, C.c_int]
        l.orc_gmbe.restype = C.c_int
        l.orc_gmbe.argtypes = graph + [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_uint32, C.c_uint32, C.c_double,
                                       C.POINTER(OracleResult), P]
        l.orc_mbea.restype = C.c_int
        l.orc_mbea.argtypes = graph + [C.c_int, C.c_uint32, C.c_uint32, C.c_double,
                                       C.POINTER(OracleResult), P]
        l.orc_imbea.restype = C.c_int
        l.orc_imbea.argtypes = l.orc_mbea.argtypes
        l.orc_pmbe.restype = C.c_int
        l.orc_pmbe.argtypes = l.orc_mbea.argtypes
        l.orc_brute.restype = C.c_int
        l.orc_brute.argtypes = graph + [C.POINTER(OracleResult), P]
        l.orc_list_new.restype = P
        l.orc_list_free.argtypes = [P]
        l.orc_list_size.restype = C.c_size_t
        l.orc_list_size.argtypes = [P]
        l.orc_list_data.restype = C.POINTER(C.c_uint32)
        l.orc_list_data.argtypes = [P]
        l.orc_hash_pair.restype = C.c_uint64
        l.orc_hash_pair.argtypes = [P, C.c_size_t, P, C.c_size_t]
        _lib = l
    return _lib


def _graph_args(g):
    nU, nV, v_off, v_adj, u_orig, v_orig, swapped = g
    v_off = np.ascontiguousarray(v_off, dtype=np.int64)
    v_adj = np.ascontiguousarray(v_adj, dtype=np.uint32)
    u_orig = np.ascontiguousarray(u_orig, dtype=np.uint32)
    v_orig = np.ascontiguousarray(v_orig, dtype=np.uint32)
    keep = (v_off, v_adj, u_orig, v_orig)
    args = [nU, nV, v_off.ctypes.data, v_adj.ctypes.data if v_adj.size else None,
            u_orig.ctypes.data if u_orig.size else None,
            v_orig.ctypes.data if v_orig.size else None, int(swapped)]
    return args, keep


def _take_list(handle):
    l = _l()
    n = l.orc_list_size(handle)
    data = np.ctypeslib.as_array(l.orc_list_data(handle), shape=(n,)).copy() if n else \
        np.zeros(0, np.uint32)
    l.orc_list_free(handle)
    return data


def gmbe(g, bound_height=-1, bound_size=0, prune=1, max_prefix=4, threads=1, stride=1,
         phase=0, time_limit=0.0, want_list=False):
    """GMBE-order restatement; bound_height < 0 = unsplit traversal (B_alg)."""
    l = _l()
    args, keep = _graph_args(g)
    res = OracleResult()
    h = l.orc_list_new() if want_list else None
    rc = l.orc_gmbe(*args, bound_height, bound_size, prune, max_prefix, threads, stride, phase,
                    time_limit, C.byref(res), h)
    lst = _take_list(h) if want_list else None
    if rc != 0:
        raise RuntimeError(f"orc_gmbe failed: {rc}")
    del keep
    return (res, lst) if want_list else res


def mbea(g, threads=1, stride=1, phase=0, time_limit=0.0, want_list=False):
    """MbeaFinder restatement (the reference's CPU path)."""
    l = _l()
    args, keep = _graph_args(g)
    res = OracleResult()
    h = l.orc_list_new() if want_list else None
    rc = l.orc_mbea(*args, threads, stride, phase, time_limit, C.byref(res), h)
    lst = _take_list(h) if want_list else None
    if rc != 0:
        raise RuntimeError(f"orc_mbea failed: {rc}")
    del keep
    return (res, lst) if want_list else res


def imbea(g, threads=1, stride=1, phase=0, time_limit=0.0, want_list=False):
    """iMBEA restatement (ImbeaAdvFinder of baselines/MBE.zip): the authors'
    improved CPU finder, a further CPU baseline."""
    l = _l()
    args, keep = _graph_args(g)
    res = OracleResult()
    h = l.orc_list_new() if want_list else None
    rc = l.orc_imbea(*args, threads, stride, phase, time_limit, C.byref(res), h)
    lst = _take_list(h) if want_list else None
    if rc != 0:
        raise RuntimeError(f"orc_imbea failed: {rc}")
    del keep
    return (res, lst) if want_list else res


def pmbe(g, threads=1, stride=1, phase=0, time_limit=0.0, want_list=False):
    """PMBE restatement (PmbeAdvFinder of baselines/MBE.zip): the authors' CPU
    finder over the neighbourhood-containment order, a further CPU baseline.
    Processing ranks follow that order, not the other finders' ids."""
    l = _l()
    args, keep = _graph_args(g)
    res = OracleResult()
    h = l.orc_list_new() if want_list else None
    rc = l.orc_pmbe(*args, threads, stride, phase, time_limit, C.byref(res), h)
    lst = _take_list(h) if want_list else None
    if rc != 0:
        raise RuntimeError(f"orc_pmbe failed: {rc}")
    del keep
    return (res, lst) if want_list else res


def brute(g, want_list=False):
    """Definition check over all subsets of V (|V| <= 24, |U| <= 64)."""
    l = _l()
    args, keep = _graph_args(g)
    res = OracleResult()
    h = l.orc_list_new() if want_list else None
    rc = l.orc_brute(*args, C.byref(res), h)
    lst = _take_list(h) if want_list else None
    if rc != 0:
        raise RuntimeError(f"orc_brute failed: {rc}")
    del keep
    return (res, lst) if want_list else res


def hash_pair(A, B) -> int:
    A = np.ascontiguousarray(A, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    return int(_l().orc_hash_pair(A.ctypes.data, A.size, B.ctypes.data, B.size))
