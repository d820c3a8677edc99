"""Host logic of the request server (arena allocator, batch planner), built
from visondf_amd/csrc/serve_core.h with g++ and run on the CPU."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_serve_core(tmp_path):
    exe = str(tmp_path / "serve_core_test")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    "-o", exe, os.path.join(HERE, "serve_core_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "serve_core ok" in r.stdout
