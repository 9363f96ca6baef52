"""The warm round's target order (hbt-herons_amd/csrc/hilbert.h): the Hilbert
index of an octree cell, compiled for the host and checked over the whole
1024^3 grid by tests/cpp/hilbert_test.cpp (a bijection, consecutive indices on
face-adjacent cells)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = shutil.which("g++")


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_hilbert_index_is_a_hilbert_curve(tmp_path):
    exe = tmp_path / "hilbert_test"
    src = os.path.join(ROOT, "tests", "cpp", "hilbert_test.cpp")
    inc = os.path.join(ROOT, "hbt-herons_amd", "csrc")
    r = subprocess.run([CXX, "-O2", "-std=c++11", "-pthread", "-Wall", "-Werror", "-I", inc, "-o", str(exe), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
