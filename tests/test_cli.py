"""The drop-in CLI (gmbe_amd/mbe_gpu, the reference's src/main.cu:76-149):
flags, the stdout lines the reference's scripts grep, and exit codes."""
import ctypes as C
import os
import subprocess

import pytest

from gmbe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gmbe_amd", "mbe_gpu")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


def gpu_visible():
    n = C.c_int()
    return _lib.lib().gmbe_device_info(None, 0, C.byref(n)) == 0


@pytest.fixture
def adj(tmp_path):
    p = tmp_path / "t.adj"
    p.write_text("0 1 2\n1 2\n0\n")
    return str(p)


def test_cli_prints_the_reference_lines(adj):
    r = run("-i", adj, "-s", "2", "-t", "1", "-o", "1", "-f", "-m", "20", "-n", "1500")
    assert "Graph:" in r.stdout and "|U|:" in r.stdout and "|E|:" in r.stdout
    assert "trans:1 sel:2 order:1 fast_mode:1" in r.stdout
    assert "Graph load time:" in r.stdout
    if gpu_visible():
        assert r.returncode == 0
        assert "maximal nodes:" in r.stdout and "Total processing time" in r.stdout
    else:
        # no GPU: the finder fails loudly, as gpuAssert does
        assert r.returncode == 5 and "GPUassert" in r.stderr


def test_cli_non_fast_mode_prints_two_hop_degrees(adj):
    r = run("-i", adj, "-s", "2")
    assert "d2(U):" in r.stdout and "d2(V):" in r.stdout


def test_cli_refuses_several_gpus(adj):
    r = run("-i", adj, "-s", "3", "-x", "4")
    assert r.returncode == 8 and "one GPU" in r.stderr


def test_cli_missing_file_is_an_error(tmp_path):
    r = run("-i", str(tmp_path / "none.adj"), "-s", "2")
    assert r.returncode != 0 and "can't open" in r.stderr


@pytest.mark.gpu
def test_cli_count_on_the_gpu(gpu, tmp_path):
    """The drop-in CLI on the uniform 2,000 x 1,500 graph: the reference's
    stdout lines, in the order its scripts read them, with the oracle's count."""
    import oracle
    from gmbe_amd import workloads
    raw = workloads.raw_graph(1)
    path = str(tmp_path / "c1.adj")
    raw.write_adj(path)
    expect = oracle.gmbe(raw.derive(**workloads.RUN_OPTION).oracle_graph()).maximal
    r = run("-i", path, "-s", "2", "-t", "1", "-o", "1", "-f")
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert f"maximal nodes:{expect}" in out
    assert out.index("non_maximal:") < out.index("Total processing time") < \
        out.index("maximal nodes:")
