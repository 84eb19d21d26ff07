"""The host-side loader, writers and generators under AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only): malformed and edge-case
.adj files through every option combination, derive, both writers, the
small-graph generator with ego cuts, the planted-block generator, and the
guard that turns an exception inside a C entry point into an error code."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "gmbe_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_loader_and_generators_are_sanitizer_clean(tmp_path):
    exe = tmp_path / "drive"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), "-o", str(exe),
           os.path.join(HERE, "host", "sanitize_drive.cpp"),
           os.path.join(CSRC, "graph.cpp"), os.path.join(CSRC, "gen.cpp")]
    subprocess.run(cmd, check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_oracle_is_sanitizer_clean(tmp_path):
    """The CPU oracle itself: the five enumerators agree on random small
    graphs, lists on, worker threads on, under ASan and UBSan."""
    exe = tmp_path / "drive_oracle"
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-pthread",
           "-o", str(exe), os.path.join(HERE, "host", "sanitize_oracle.c"),
           os.path.join(ROOT, "oracle", "oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
