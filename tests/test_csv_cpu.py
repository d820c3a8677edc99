"""CPU unit test of the CSV field conversions (visondf_amd/csrc/csv_parse.h,
built with g++): doubles bit-exact against glibc strtod, integers against
strtoll / strtoull with std::from_chars' range rules."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_csv_parse(tmp_path):
    exe = str(tmp_path / "csv_parse_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", exe,
                    os.path.join(HERE, "csv_parse_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "csv_parse ok" in r.stdout
