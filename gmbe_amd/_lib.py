"""ctypes binding of libgmbe.so (the C ABI declared in include/gmbe.h).

The shared library is built in-tree by ``gmbe_amd/csrc/Makefile`` (see
``__graft_entry__.build``).  Loading fails loudly when it is missing: there is
no CPU fallback for the enumeration.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GMBE_LIB", os.path.join(_HERE, "libgmbe.so"))

GMBE_OK = 0
ERRORS = {
    -1: "GMBE_ERR_ARG", -2: "GMBE_ERR_IO", -3: "GMBE_ERR_FORMAT",
    -4: "GMBE_ERR_CAPACITY", -5: "GMBE_ERR_HIP", -6: "GMBE_ERR_NOMEM",
    -7: "GMBE_ERR_REJECTED", -8: "GMBE_ERR_UNSUPPORTED", -9: "GMBE_ERR_STATE", -10: "GMBE_ERR_DEADLINE",
    -11: "GMBE_ERR_INTERNAL",
}


class GmbeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Option(C.Structure):
    _fields_ = [("uvtrans", C.c_int32), ("order", C.c_int32), ("fast_mode", C.c_int32)]


class Result(C.Structure):
    _fields_ = [
        ("maximal", C.c_uint64), ("non_maximal", C.c_uint64),
        ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64),
        ("nodes", C.c_uint64), ("tasks", C.c_uint64), ("splits", C.c_uint64),
        ("overflow", C.c_uint64), ("alg_bytes", C.c_uint64),
        ("exe_time_s", C.c_double), ("kernel_time_s", C.c_double),
    ]

    def digest(self):
        return (int(self.maximal), int(self.hash_sum), int(self.hash_xor))


class GenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nU", C.c_uint32), ("nV", C.c_uint32), ("m", C.c_uint64),
        ("gamma_u", C.c_double), ("gamma_v", C.c_double), ("cap", C.c_uint32),
        ("blocks", C.c_uint32), ("block_min", C.c_uint32), ("block_max", C.c_uint32),
        ("drop", C.c_double), ("seed", C.c_uint64),
    ]


class ServeDone(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64), ("status", C.c_int32),
        ("maximal", C.c_uint64), ("hash_sum", C.c_uint64), ("hash_xor", C.c_uint64),
        ("t_submit_s", C.c_double), ("t_done_s", C.c_double),
    ]


# name -> (restype, argtypes); every symbol include/gmbe.h declares
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
SIGNATURES = {
    "gmbe_option_default": (None, [C.POINTER(Option)]),
    "gmbe_graph_read_adj": (C.c_int, [C.c_char_p, C.POINTER(Option), _PP]),
    "gmbe_graph_from_csr": (C.c_int, [C.c_uint32, C.c_uint32, _P, _P, C.POINTER(Option), _PP]),
    "gmbe_graph_derive": (C.c_int, [_P, C.POINTER(Option), _PP]),
    "gmbe_graph_stats": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gmbe_graph_csr": (C.c_int, [_P, _PP, _PP, _PP, _PP, _PP, _PP]),
    "gmbe_graph_write_adj": (C.c_int, [_P, C.c_char_p]),
    "gmbe_graph_write_edges": (C.c_int, [_P, C.c_char_p]),
    "gmbe_graph_destroy": (None, [_P]),
    "gmbe_generate": (C.c_int, [C.POINTER(GenParams), _PP]),
    "gmbe_generate_small": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                      C.c_double, C.c_double, _PP]),
    "gmbe_ego_subgraph": (C.c_int, [_P, C.c_uint32, C.c_uint64, _PP]),
    "gmbe_hash_biclique": (C.c_uint64, [_P, C.c_size_t, _P, C.c_size_t]),
    "gmbe_finder_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _PP]),
    "gmbe_finder_set": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "gmbe_finder_execute": (C.c_int, [_P, C.POINTER(Result)]),
    "gmbe_finder_execute_static": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Result)]),
    "gmbe_finder_exit_times": (C.c_int, [_P, C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_size_t)]),
    "gmbe_finder_list": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "gmbe_finder_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_size_t]),
    "gmbe_finder_destroy": (None, [_P]),
    "gmbe_batch_create": (C.c_int, [_PP, C.c_int, C.c_int, C.c_int, _PP]),
    "gmbe_batch_set": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "gmbe_batch_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64)]),
    "gmbe_batch_execute": (C.c_int, [_P, C.POINTER(Result), C.POINTER(Result)]),
    "gmbe_batch_destroy": (None, [_P]),
    "gmbe_serve_create": (C.c_int, [_P, C.c_int, C.c_uint64, C.c_int, C.c_int, _PP]),
    "gmbe_serve_set": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "gmbe_serve_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int]),
    "gmbe_serve_submit_graph": (C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    "gmbe_serve_submit_ego": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "gmbe_serve_poll": (C.c_int, [_P, C.POINTER(ServeDone), C.c_int, C.c_int,
                                  C.POINTER(C.c_int)]),
    "gmbe_serve_destroy": (None, [_P]),
    "gmbe_last_error": (C.c_char_p, []),
    "gmbe_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
}

_lib = None


def lib():
    """Load libgmbe.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int):
    if rc != GMBE_OK:
        msg = lib().gmbe_last_error()
        raise GmbeError(rc, msg.decode() if msg else "")
    return rc
