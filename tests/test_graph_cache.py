"""The graph cache's policy (giant-impact_amd/csrc/gi_graph_cache.h) on the
host alone: built with g++ and run against a simulated device that keeps
launches in flight until a synchronisation (tests/cpp/graph_cache_test.cpp)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_graph_cache_never_destroys_held_or_running_graphs(tmp_path):
    exe = tmp_path / "graph_cache_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", str(REPO / "giant-impact_amd" / "csrc"),
                    str(REPO / "tests" / "cpp" / "graph_cache_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout
