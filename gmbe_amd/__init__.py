"""gmbe_amd — maximal biclique enumeration on one AMD Instinct MI355X (gfx950).

Host C++ + hand-written HIP kernels behind the C ABI in include/gmbe.h; this
package is the Python face of that ABI (ctypes), mirroring the reference's
CSRBiGraph / IterFinderGpu6 surface.
"""
from ._lib import GmbeError, GenParams, Option, Result, lib  # noqa: F401
from .graph import BiGraph, make_option  # noqa: F401
from .finder import IterFinderGpu6, enumerate_graph, records_to_set  # noqa: F401
from .batch import Batch, Server  # noqa: F401


def hash_biclique(A, B) -> int:
    """Per-biclique hash (A on the original U side, B on the original V side)."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.uint32)
    B = np.ascontiguousarray(B, dtype=np.uint32)
    return int(lib().gmbe_hash_biclique(A.ctypes.data, A.size, B.ctypes.data, B.size))


def digest_of(bicliques) -> tuple:
    """(count, sum h mod 2^64, xor h) of a set of (A, B) bicliques."""
    s = x = 0
    for A, B in bicliques:
        h = hash_biclique(A, B)
        s = (s + h) & 0xFFFFFFFFFFFFFFFF
        x ^= h
    return (len(bicliques), s, x)
