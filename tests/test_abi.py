"""The C ABI library loads and exports every symbol include/gmbe.h declares."""
import ctypes as C
import os
import re

import pytest

from gmbe_amd import _lib
from gmbe_amd._lib import GmbeError

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gmbe.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmbe_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_match_binding():
    assert set(declared_symbols()) == set(_lib.SIGNATURES), \
        "ctypes binding and include/gmbe.h disagree"


def test_error_codes_match_header():
    """Every GMBE_ERR_* the header defines has its name in the binding's map,
    with the same value, and the map holds no other code."""
    text = open(os.path.join(ROOT, "include", "gmbe.h")).read()
    codes = {int(v): n for n, v in re.findall(r"#define (GMBE_ERR_\w+) \((-\d+)\)", text)}
    assert codes and codes == _lib.ERRORS


def test_library_exports_every_declared_symbol():
    l = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(l, s)]
    assert not missing, missing


def test_errors_are_codes_not_exits(tmp_path):
    from gmbe_amd import BiGraph
    with pytest.raises(GmbeError) as e:
        BiGraph.read_adj(str(tmp_path / "missing.adj"))
    assert e.value.code == -2 and "missing.adj" in str(e.value)
    bad = tmp_path / "bad.adj"
    bad.write_text("0 1 x\n")
    with pytest.raises(GmbeError) as e:
        BiGraph.read_adj(str(bad))
    assert e.value.code == -3


def test_finder_without_gpu_fails_loudly():
    """Without a GPU the finder refuses to run (no CPU fallback)."""
    from gmbe_amd import BiGraph, IterFinderGpu6
    n = C.c_int()
    if _lib.lib().gmbe_device_info(None, 0, C.byref(n)) == 0:
        pytest.skip("a GPU is visible")
    g = BiGraph.from_rows(2, [[0, 1], [1]])
    with pytest.raises(GmbeError) as e:
        IterFinderGpu6(g)
    assert e.value.code == -5


def test_multi_gpu_request_is_rejected():
    from gmbe_amd import BiGraph
    g = BiGraph.from_rows(2, [[0, 1], [1]])
    out = C.c_void_p()
    rc = _lib.lib().gmbe_finder_create(g.handle, 3, 20, 1500, 8, C.byref(out))
    assert rc in (-8, -5)


def test_batch_and_server_without_gpu_fail_loudly():
    """The batched form and the serving loop refuse to run without a GPU."""
    from gmbe_amd import Batch, BiGraph, Server
    n = C.c_int()
    if _lib.lib().gmbe_device_info(None, 0, C.byref(n)) == 0:
        pytest.skip("a GPU is visible")
    g = BiGraph.from_rows(2, [[0, 1], [1]], uvtrans=1, order=1, fast_mode=1)
    with pytest.raises(GmbeError) as e:
        Batch([g])
    assert e.value.code == -5
    with pytest.raises(GmbeError) as e:
        Server(None, slots=4, mem_budget=1 << 20)
    assert e.value.code == -5


def test_batch_and_server_argument_errors():
    """Argument checks come before any device work."""
    lib = _lib.lib()
    out = C.c_void_p()
    assert lib.gmbe_batch_create(None, 0, 20, 1500, C.byref(out)) == -1
    arr = (C.c_void_p * 1)(None)
    assert lib.gmbe_batch_create(arr, 1, 20, 1500, C.byref(out)) == -1
    assert lib.gmbe_serve_create(None, 0, 1 << 20, 20, 1500, C.byref(out)) == -1
    assert lib.gmbe_serve_create(None, 4, 1 << 20, -1, 1500, C.byref(out)) == -1
    assert lib.gmbe_serve_poll(None, None, 0, 0, None) == -1
    assert lib.gmbe_serve_set(None, b"request_ms", 1) == -1
    assert lib.gmbe_batch_get(None, b"lds_kb", None) == -1
