"""Host logic of the request server (arena allocator, batch planner), built
from visondf_amd/csrc/serve_core.h with g++ and run on the CPU, and a CPU
replay of the served benchmark's request trace through the planner."""
from __future__ import annotations

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "visondf_amd")

KEY_U64, KEY_CHAR16 = 2, 4
# device seconds per row of one group-by (8 radix passes for uint64 keys over
# all 64 bits, 16 for 16-byte ASCII keys), the scale of the measured rates
COST = {"u64": 1.0 / 10e9, "char16": 1.0 / 4e9}


def test_serve_core(tmp_path):
    exe = str(tmp_path / "serve_core_test")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    "-o", exe, os.path.join(HERE, "serve_core_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "serve_core ok" in r.stdout


def _replay(tmp_path, exe, arena, workers, max_tables, max_rows, by_head, rate=2000.0, requests=1000, cost=None):
    from visondf_amd import datagen

    cost = cost or COST
    sizes, kinds, arrivals, picks = datagen.served_trace(32, requests, rate)
    lines = [f"{arena} {workers} {max_tables} {max_rows} {int(by_head)}", str(len(sizes))]
    for n, kind in zip(sizes, kinds):
        lines.append(f"{KEY_U64 if kind == 'u64' else KEY_CHAR16} {int(n)} {int(n) // 16 + 1} {cost[kind]!r}")
    lines.append(str(len(arrivals)))
    lines += [f"{float(a)!r} {int(t)}" for a, t in zip(arrivals, picks)]
    path = tmp_path / f"trace_{workers}_{max_rows}_{int(by_head)}.txt"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode in (0, 1), r.stdout + r.stderr
    return r.stdout.strip()


@pytest.fixture(scope="module")
def replay_exe(tmp_path_factory):
    lib = os.path.join(LIBDIR, "libvdf_hip.so")
    if not os.path.exists(lib):
        pytest.fail("libvdf_hip.so is not built (run __graft_entry__.build())")
    exe = str(tmp_path_factory.mktemp("replay") / "serve_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-o", exe, os.path.join(HERE, "serve_replay.cpp"), "-L" + LIBDIR, "-lvdf_hip",
                    "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


def test_served_trace_replay(tmp_path, replay_exe):
    """The benchmark's trace (32 tables of 10K-50M rows, uint64 and 16-byte
    keys, 1000 Poisson requests, 48 GiB arena, 3 workers, up to 64 requests
    per batch, the default row limit) finishes every request with a caller
    that takes results in submission order."""
    arena = 48 << 30
    default_rows = max(arena // (96 * 2 * 3), 1 << 20)
    out = _replay(tmp_path, replay_exe, arena, 3, 64, default_rows, False)
    assert out.startswith("OK consumed=1000/1000"), out
    # also with no row limit, one and three workers, and twice the load
    for workers, rows, rate in ((3, 1 << 62, 2000.0), (1, default_rows, 2000.0), (3, default_rows, 4000.0)):
        out = _replay(tmp_path, replay_exe, arena, workers, 64, rows, False, rate=rate)
        assert out.startswith("OK consumed=1000/1000"), out


def test_served_trace_replay_timings(tmp_path, replay_exe):
    """The same trace under other device speeds for the two key types (the
    model's only timing input), 2 to 4 workers, with and without the row
    limit.  Before holds were limited to requests a batch actually passes,
    most of these stalled with an empty arena: holds for younger requests of
    the other key type filled it, so the oldest request's scratch never fit."""
    arena = 48 << 30
    default_rows = max(arena // (96 * 2 * 3), 1 << 20)
    for cu in (30e9, 10e9, 6e9):
        for cc in (8e9, 4e9, 2e9):
            for workers in (2, 3, 4):
                for rows in (default_rows, 1 << 62):
                    out = _replay(tmp_path, replay_exe, arena, workers, 64, rows, False,
                                  cost={"u64": 1.0 / cu, "char16": 1.0 / cc})
                    assert out.startswith("OK consumed=1000/1000"), (cu, cc, workers, rows, out)
