"""The header-only C++ façade (include/vison_gpu.hpp) driven by a compiled
program: reference-shaped Dataframe calls (add_col, sort_by<ASC, Soft>,
materialize, group_by, transform_group_by) checked against host computations
(tests/facade_test.cpp, built by visondf_amd/csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "bin", "facade_test")


def test_facade_program():
    assert os.path.exists(EXE), f"{EXE} is missing: build with __graft_entry__.build()"
    print("running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade ok" in r.stdout
