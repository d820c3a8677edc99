"""Benchmark of the radix-sort group-by path on one MI355X.

Workloads (inputs resident in HBM before timing; synthetic seeded data):
  argsort_u32   headline: 100M rows, uint32 keys uniform over 2^32; stable LSD
                radix argsort (8-bit digits, 4 passes) + gather of one float64
                column through the permutation.  One step = argsort + gather.
  zipf_groupby  200M rows, int32 keys Zipf(1.1) over 2^24 values, float64
                N(0,1) and int64 uniform [0, 2^40) values; group-by
                sum/count/mean/min/max.  One step = the whole group-by.
  batched       4096 tables, rows log-uniform in [1K, 4M], int32 keys uniform
                in [0, rows/8), float64 values; every table sorted and grouped
                by one segmented launch sequence, against one call per table.
  served        Poisson arrivals (2000 requests/s) of group-bys over 32
                device-resident tables of 10K-50M rows (uint64 or 16-byte
                string keys, float64 values), continuous batching over several
                streams against one launch per request; p50/p99 latency.
                Switched off (see DEFAULT_WORKLOADS).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]          # all workloads
    python bench.py --workload argsort_u32 [--steps ...]          # one, in-process

Without --workload the parent process touches no GPU: it runs each workload
in a child process with its own time limit and prints one JSON line whose
top-level fields are the headline workload's, with the others under
"workloads".
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "group-by rows/sec + HBM GB/s (% peak) at 100M rows; served p50/p99 latency"
KERNELS = ("radix_histogram", "radix_plan", "radix_pass", "radix_local_plan", "radix_local", "gather", "fill",
           "groupby_empty", "groupby_tile",
           "groupby_fixup", "csv_newlines", "csv_lines", "csv_measure", "csv_convert")
WORKLOADS = ("argsort_u32", "groupby_c1", "groupby_i32", "groupby_i32_full", "zipf_groupby", "batched", "csv_ingest",
             "served")
# served is left out of the default run and refused by --workload: its
# continuous-batching replay (three workers) stopped making progress with
# batches in flight and the arena far from full (cause not found)
DEFAULT_WORKLOADS = ("argsort_u32", "groupby_c1", "groupby_i32", "groupby_i32_full", "zipf_groupby", "batched",
                     "csv_ingest")
CHILD_TIMEOUT = {"argsort_u32": 600, "groupby_c1": 300, "groupby_i32": 400, "groupby_i32_full": 400, "zipf_groupby": 400, "batched": 500, "csv_ingest": 300, "served": 500}


def _log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _host() -> dict:
    """The host CPUs this process may use: model, logical CPUs in its affinity
    mask, distinct physical cores among them, and the cgroup CPU quota."""
    allowed = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in allowed:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add((str(c), ""))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    return {"model": model, "logical_cpus": len(allowed), "physical_cores": len(cores), "cgroup_cpus": quota,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def _threads(args) -> int:
    """CPU-baseline threads: min(physical cores available, 32), the
    reference's cap (MAX_THREADS, details/constants.inl:9), and within the
    cgroup's CPU quota."""
    if args.cpu_threads:
        return args.cpu_threads
    h = _host()
    t = min(h["physical_cores"] or 1, 32)
    if h["cgroup_cpus"]:
        t = min(t, max(int(h["cgroup_cpus"]), 1))
    return max(t, 1)


TRAFFIC_SOURCES = ("radix.hip", "groupby.hip", "gather.hip")


def _sources_hash() -> str:
    """sha256 prefix over the sources of the kernels whose traffic
    profiles/traffic.json records (TRAFFIC_SOURCES in visondf_amd/csrc), the
    stamp that file keeps for the build it measured."""
    import hashlib

    h = hashlib.sha256()
    for name in TRAFFIC_SOURCES:
        h.update(name.encode())
        with open(os.path.join(ROOT, "visondf_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _load_traffic(workload: str, name: str):
    """(HBM bytes per launch of a kernel in a workload, whether the kernel
    sources are the ones that were measured) from the committed rocprofv3
    counter summaries (profiles/traffic.json); (None, None) when absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    t = d.get(workload, {}).get(name)
    if t is None:
        return None, None
    return t, d.get("_sources") == _sources_hash()


def _kernel_table(v) -> dict:
    kernels = {}
    for name in KERNELS:
        ms, launches, nbytes = v.profile_read(name)
        if launches:
            kernels[name] = {"launches": launches, "total_ms": ms, "avg_us": 1e3 * ms / launches,
                             "bytes_per_launch": nbytes / launches,
                             "gbps": (nbytes / (ms * 1e-3)) / 1e9 if ms > 0 and nbytes else None}
    return kernels


def _roofline(workload: str, kernels: dict, name: str | None = None) -> dict:
    """HBM roofline of one kernel (default: the one with the most time that
    moves algorithmic bytes)."""
    if name is None:
        cands = [(k["total_ms"], n) for n, k in kernels.items() if k["gbps"]]
        name = max(cands)[1]
    k = kernels[name]
    traffic, current = _load_traffic(workload, name)
    return {"bound": "hbm", "kernel": name, "achieved": k["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": k["gbps"] / HBM_PEAK_GBPS, "traffic": traffic, "traffic_sources_current": current}


def _handle(args):
    """A library handle on device 0 with the --tune settings applied."""
    from visondf_amd.device import Vdf

    v = Vdf(0)
    for item in args.tune:
        name, _, val = item.partition("=")
        v.tune(name, int(val))
    return v


def _time_steps(step, steps: int, warmup: int, v=None):
    """Times `steps` steps between torch.cuda.synchronize() calls (no per-launch
    events in the timed region); with a handle, then runs the same number of
    steps again with per-launch HIP events for the kernel table."""
    import torch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        step()
    t1.record()
    torch.cuda.synchronize()
    kernels = None
    if v is not None:
        v.profile(True)
        v.profile_reset()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        kernels = _kernel_table(v)
        v.profile(False)
    return t0.elapsed_time(t1), kernels


def _base_record(args, value, unit, ms_step, dtype, config, higher=True) -> dict:
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


# ---------------------------------------------------------------------------
# headline: 100M uint32 argsort + f64 gather
# ---------------------------------------------------------------------------
def cpu_baseline_argsort(keys, val, threads: int, reps: int = 3, port: bool = False) -> dict:
    """The reference's own radix_sort_uint32_mt (compiled from its .inl files
    into oracle/_ref, 16 threads) when that build is present, else the oracle's
    restatement of it, plus the gather (permute_block_mt restatement), timed on
    the host cores: median of reps after one warm-up."""
    import ctypes as C

    import numpy as np

    import oracle

    n = keys.shape[0]
    ref = None if port else oracle.ref_lib()
    use_ref = ref is not None and threads >= 8
    # radix_sort_uint32_mt<CORES>: CORES is a compile-time constant; the
    # largest instantiation not above the threads available
    cores = 32 if threads >= 32 else (16 if threads >= 16 else 8)
    ref_fn = None if not use_ref else getattr(ref, f"vref_radix_sort_u32_mt{cores}")
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        if use_ref:
            # sort_uintegers copies the key column and starts from iota (sort_by_mt.hpp:21-24)
            perm = np.arange(n, dtype=np.uint64)
            k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
            ref_fn(C.c_void_p(k.ctypes.data), C.c_void_p(perm.ctypes.data), n)
        else:
            perm = oracle.argsort(keys, 0, threads)
        oracle.gather(val, perm, threads)
        dt = time.perf_counter() - t0
        if r:
            times.append(dt)
        _log(f"cpu baseline rep {r}: {dt:.3f} s")
    t = float(np.median(times))
    if use_ref:
        exp = oracle.argsort(keys, 0, threads)
        assert np.array_equal(perm, exp), "reference radix and oracle disagree"
        return {"value": n / t, "unit": "rows/s", "cores": cores, "kind": "reference", "host": _host(),
                "sample": f"full workload ({n} rows): the reference's radix_sort_uint32_mt<{cores}, Simd=false> "
                          f"(its .inl files compiled in place, oracle/_ref) + the f64 gather on {threads} threads, "
                          f"median of {reps} after 1 warm-up"}
    return {"value": n / t, "unit": "rows/s", "cores": threads, "kind": "port", "host": _host(),
            "sample": f"full workload ({n} rows) argsort + f64 gather, median of {reps} after 1 warm-up; "
                      f"oracle radix_sort_uint32_mt restatement, 16-bit digits, {threads} OpenMP threads"}


def run_argsort_u32(args) -> dict:
    import numpy as np
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.device import Vdf

    n = args.rows or 100_000_000
    _log(f"argsort_u32: generating {n} rows")
    keys, val = datagen.argsort_u32(n)
    d_keys = torch.from_numpy(keys.view(np.int32)).cuda()
    d_val = torch.from_numpy(val).cuda()
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    v = _handle(args)
    tune = None
    if not args.no_tune:
        _log("argsort_u32: autotuning the launch shape")
        tune = v.autotune_sort_by(d_keys, d_val, perm, out, key_type=N.KEY_U32, reps=3)
        _log(f"argsort_u32: autotune {tune}")

    # One step: the stable argsort of the keys with the float64 column carried
    # through the radix passes (vdf_sort_by: permutation + column in key order).
    def step():
        v.sort_by(d_keys, d_val, key_type=N.KEY_U32, perm_out=perm, vals_out=out)

    ms_total, kernels = _time_steps(step, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    rows_per_s = n * args.steps / (ms_total * 1e-3)
    step_bytes = sum(k["bytes_per_launch"] * k["launches"] for k in kernels.values()) / args.steps
    # spot check of the last step: the permutation sorts the keys (stably) and
    # the column follows it
    ks = d_keys.to(torch.int64) & 0xFFFFFFFF
    p64 = perm.to(torch.int64)
    sk = ks[p64]
    assert bool((sk[1:] >= sk[:-1]).all()), "permutation does not sort the keys"
    assert bool((p64[1:][sk[1:] == sk[:-1]] > p64[:-1][sk[1:] == sk[:-1]]).all()), "sort is not stable"
    assert torch.equal(out, d_val[p64]), "value column does not follow the permutation"

    # the same result through two calls (argsort, then a random-access gather)
    two = None
    if not args.skip_compare:
        perm2 = torch.empty_like(perm)
        out2 = torch.empty_like(out)

        def step_two_calls():
            v.argsort(d_keys, key_type=N.KEY_U32, perm_out=perm2)
            v.gather(d_val, perm2, out=out2)

        ms2, kernels2 = _time_steps(step_two_calls, args.steps, args.warmup, v)
        assert torch.equal(perm2, perm) and torch.equal(out2, out), "argsort + gather differs from sort_by"
        two = {"method": "vdf_argsort + vdf_gather", "ms_per_step": ms2 / args.steps,
               "rows_per_s": n * args.steps / (ms2 * 1e-3), "kernels": kernels2}

    # the plan is chosen on the device: global passes and the LDS local finish per step
    passes = kernels.get("radix_pass", {}).get("launches", 0) / args.steps
    local = kernels.get("radix_local", {}).get("launches", 0) / args.steps
    rec = _base_record(args, rows_per_s, "rows/s", ms_step, "u32",
                       {"workload": "argsort_u32_gather_f64", "rows": n, "key": "uint32 uniform over 2^32",
                        "value_column": "float64", "digit_bits": 8, "radix_passes": passes,
                        "local_finish_16bit": bool(local),
                        "method": "vdf_sort_by (column carried through the passes)",
                        "parallelism": "single GPU"})
    rec["hbm_gbps_step"] = step_bytes / (ms_step * 1e-3) / 1e9
    rec["hbm_pct_peak_step"] = 100.0 * rec["hbm_gbps_step"] / HBM_PEAK_GBPS
    rec["roofline"] = _roofline("argsort_u32", kernels, "radix_pass")
    rec["kernels"] = kernels
    rec["two_calls"] = two
    rec["autotune"] = tune
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline_argsort(keys, val, _threads(args))
        rec["speedup_vs_cpu"] = rows_per_s / rec["cpu_baseline"]["value"]
        if rec["cpu_baseline"]["kind"] == "reference":
            # the oracle's restatement as well (faster than the reference's own loops)
            rec["cpu_port"] = cpu_baseline_argsort(keys, val, _threads(args), port=True)
    else:
        rec["cpu_baseline"] = None
    v.close()
    return rec


# ---------------------------------------------------------------------------
# 200M-row Zipf group-by
# ---------------------------------------------------------------------------
def _cpu_ref_groupby(tables, threads: int, what: str, refapi: bool = False, reps: int = 1) -> dict:
    """The reference's hash group-by (oracle/_ref: its ankerl map and
    simd_hash under loops restated from onecol_functions.inl /
    merge_functions.inl) over a list of (keys, f64 column, second column or
    None) tables, one call per table, on `threads` OpenMP threads: median of
    `reps` timed runs after a warm-up on the first table's first rows."""
    import numpy as np

    import oracle

    rows = sum(k.shape[0] for k, _, _ in tables)
    k, a, b = tables[0]
    m = min(k.shape[0], 100_000)
    oracle.ref_hash_groupby(k[:m], a[:m], None if b is None else b[:m], threads=threads, refapi=refapi,
                            outputs=False)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for k, a, b in tables:
            oracle.ref_hash_groupby(k, a, b, threads=threads, refapi=refapi, outputs=False)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    form = ("reference API: one hash pass per aggregate per column, Gather + functor for min and max"
            if refapi else "fused: one hash pass keeping every aggregate")
    _log(f"cpu baseline ({what}, {'api' if refapi else 'fused'}): {rows} rows in {dt:.3f} s on {threads} threads")
    return {"value": rows / dt, "unit": "rows/s", "cores": threads, "kind": "reference", "host": _host(),
            "sample": f"{what}; the reference's hash group-by ({form}): its ankerl::unordered_dense map and "
                      f"simd_hash compiled from its headers in oracle/_ref, loops restated from "
                      f"onecol_functions.inl and merge_functions.inl; {threads} OpenMP threads; "
                      f"median of {reps} after a warm-up"}


def _cpu_groupby_legs(rec: dict, rows_per_s: float, full, sample_api, threads: int, what_full: str,
                      what_api: str, reps: int = 1):
    """cpu_baseline = the fused hash group-by on `full`; beside it the
    reference API form on `sample_api`, and (when more cores are available
    than the reference's cap allows) the fused form on every physical core."""
    rec["cpu_baseline"] = _cpu_ref_groupby(full, threads, what_full, refapi=False, reps=reps)
    rec["speedup_vs_cpu"] = rows_per_s / rec["cpu_baseline"]["value"]
    rec["cpu_reference_api"] = _cpu_ref_groupby(sample_api, threads, what_api, refapi=True)
    rec["speedup_vs_cpu_reference_api"] = rows_per_s / rec["cpu_reference_api"]["value"]
    allc = _host()["physical_cores"]
    if allc and allc > threads and (_host()["cgroup_cpus"] or allc) >= allc:
        rec["cpu_all_cores"] = _cpu_ref_groupby(full, allc, what_full + ", all physical cores", refapi=False)


def run_groupby_i32(args, full_range: bool = False) -> dict:
    """100M rows, int32 key uniform in [0, 1000) (or, full_range, over all
    2^32 values: about 98M groups), float64 and int32 values, every
    aggregate (the CPU reference configuration's group-by at 100M rows)."""
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.device import Vdf

    n = args.rows or 100_000_000
    name = "groupby_i32_full" if full_range else "groupby_i32"
    _log(f"{name}: generating {n} rows")
    key, vf, vi = datagen.groupby_i32_full(n) if full_range else datagen.groupby_i32(n)
    d_key = torch.from_numpy(key).cuda()
    d_vf = torch.from_numpy(vf).cuda()
    d_vi = torch.from_numpy(vi).cuda()
    v = _handle(args)
    res = {}

    def step():
        res["r"] = v.groupby(d_key, [d_vf, d_vi], aggs=N.AGG_ALL, want_perm=False, trim=False)

    ms_total, kernels = _time_steps(step, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    rows_per_s = n * args.steps / (ms_total * 1e-3)
    g = int(res["r"].n_groups.item())
    cnt = res["r"].counts[:g].to(torch.int64).sum().item()
    assert cnt == n, "group counts do not add up to the rows"
    rec = _base_record(args, rows_per_s, "rows/s", ms_step, "f64+i32",
                       {"workload": name, "rows": n,
                        "key": "int32 uniform over 2^32" if full_range else "int32 uniform [0, 1000)",
                        "values": ["float64 uniform [0,1)", "int32 uniform [-1e6,1e6]"],
                        "aggregates": "count,sum,mean,min,max", "groups": g, "parallelism": "single GPU"})
    rec["roofline"] = _roofline(name, kernels)
    rec["kernels"] = kernels
    if not args.no_cpu:
        if full_range:  # ~one group per row: bounded samples
            m, ma = min(n, 20_000_000), min(n, 2_000_000)
            _cpu_groupby_legs(rec, rows_per_s, [(key[:m], vf[:m], vi[:m])], [(key[:ma], vf[:ma], vi[:ma])],
                              _threads(args), f"first {m} rows of the workload", f"first {ma} rows of the workload")
        else:
            m = min(n, 20_000_000)
            _cpu_groupby_legs(rec, rows_per_s, [(key, vf, vi)], [(key[:m], vf[:m], vi[:m])], _threads(args),
                              f"full workload ({n} rows)", f"first {m} rows of the workload", reps=3)
    v.close()
    return rec


def run_groupby_i32_full(args) -> dict:
    return run_groupby_i32(args, full_range=True)


def run_csv_ingest(args) -> dict:
    """CSV text resident in HBM (airline delay-cause shape: 21 columns,
    quoted names holding the delimiter, two-decimal numbers; 5M rows, about
    1 GB) parsed into typed device columns by vdf_read_csv; MB/s as the
    reference's CSV benchmark reports it (benchmarks/readf.cpp:36-45)."""
    import ctypes as C

    import numpy as np
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen

    nrows = args.rows or 5_000_000
    _log(f"csv_ingest: generating {nrows} rows")
    text = datagen.csv_airline(nrows)
    nbytes = int(text.size)
    pad = torch.zeros(nbytes + 16, dtype=torch.uint8)
    pad[:nbytes] = torch.from_numpy(text)
    d_text = pad.cuda()[:nbytes]
    v = _handle(args)
    opt = N.VdfCsvOptions(delim=b",", quote=b'"', trailing=b"0", reserved0=b"\0", header=1, int_bytes=8,
                          char_width=0)
    tables = []

    def step():
        t = C.c_void_p()
        N.check(v.lib.vdf_read_csv(v.h, C.c_void_p(d_text.data_ptr()), nbytes, C.byref(opt), C.byref(t)))
        tables.append(t)

    ms_total, kernels = _time_steps(step, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    mbps = nbytes * args.steps / (ms_total * 1e-3) / 2**20
    # spot check of the last table against the oracle on the first 20000 rows
    nrow, ncol = C.c_uint64(), C.c_int()
    N.check(v.lib.vdf_csv_info(tables[-1], C.byref(nrow), C.byref(ncol)))
    assert nrow.value == nrows and ncol.value == 21, (nrow.value, ncol.value)
    import oracle

    cut = int(np.nonzero(text == 10)[0][20_000]) + 1
    o = oracle.read_csv(text[:cut], quote='"', threads=4)
    for c in (0, 1, 6, 13, 20):  # numeric columns (text widths depend on all rows)
        col = N.VdfCsvColumn()
        N.check(v.lib.vdf_csv_column_get(tables[-1], c, C.byref(col)))
        exp = o["cols"][c]
        got = np.empty(exp.shape[0] * col.width, dtype=np.uint8)
        from visondf_amd.server import _hip_lib

        assert _hip_lib().hipMemcpy(got.ctypes.data, col.data, got.size, 2) == 0
        assert np.array_equal(got, exp.view(np.uint8).reshape(-1) if exp.ndim == 1 else exp.reshape(-1)), c
    for t in tables:
        N.check(v.lib.vdf_csv_free(t))
    rec = _base_record(args, mbps, "MB/s", ms_step, "u8",
                       {"workload": "csv_ingest", "rows": nrows, "bytes": nbytes, "columns": 21,
                        "text": "airline delay-cause shape: year, month, 2-char carrier, carrier name, 3-char "
                                "airport, quoted airport name with a comma, 15 two-decimal numbers",
                        "types": "2 uint64, 4 string, 15 float64 (inferred from the first 10 rows)",
                        "parallelism": "single GPU"})
    rec["kernels"] = kernels
    rec["roofline"] = _roofline("csv_ingest", kernels)
    if not args.no_cpu:
        threads = _threads(args)
        times = []
        for r in range(3):
            t0 = time.perf_counter()
            oracle.read_csv(text, quote='"', threads=threads)
            times.append(time.perf_counter() - t0)
            _log(f"csv_ingest: cpu rep {r}: {times[-1]:.3f} s")
        tc = float(np.median(times))
        rec["cpu_baseline"] = {"value": nbytes / tc / 2**20, "unit": "MB/s", "cores": threads, "kind": "port",
                               "host": _host(),
                               "sample": f"full workload ({nbytes} bytes): the oracle's CSV ingest (newline "
                                         f"index, field split, from_chars-prefix conversions through strtod), "
                                         f"{threads} OpenMP threads, median of 3"}
        rec["speedup_vs_cpu"] = mbps / rec["cpu_baseline"]["value"]
    v.close()
    return rec


def run_groupby_c1(args) -> dict:
    """The CPU reference configuration: 1M rows, int64 key uniform in
    [0, 1000), float64 uniform [0, 1) and int32 uniform [-1e6, 1e6] values,
    every aggregate; the reference's hash group-by timed on the host (fused
    and API forms, the reference's thread count and one thread) beside the
    GPU group-by of the same rows."""
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.device import Vdf

    n = args.rows or 1_000_000
    key, vf, vi = datagen.groupby_small(n, seed=1)
    d_key = torch.from_numpy(key).cuda()
    d_vf = torch.from_numpy(vf).cuda()
    d_vi = torch.from_numpy(vi).cuda()
    v = _handle(args)
    res = {}

    def step():
        res["r"] = v.groupby(d_key, [d_vf, d_vi], aggs=N.AGG_ALL, want_perm=False, trim=False)

    ms_total, kernels = _time_steps(step, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    rows_per_s = n * args.steps / (ms_total * 1e-3)
    g = int(res["r"].n_groups.item())
    assert int(res["r"].counts[:g].to(torch.int64).sum().item()) == n, "group counts do not add up to the rows"
    rec = _base_record(args, rows_per_s, "rows/s", ms_step, "f64+i32",
                       {"workload": "groupby_c1_cpu_reference_config", "rows": n, "key": "int64 uniform [0, 1000)",
                        "values": ["float64 uniform [0,1)", "int32 uniform [-1e6,1e6]"],
                        "aggregates": "count,sum,mean,min,max", "groups": g, "parallelism": "single GPU"})
    rec["kernels"] = kernels
    # the same group-by as a captured HIP-graph plan (a batch of this one
    # table): a replay launches the whole sequence at once, which at this
    # size is mostly launch latency
    pb = v.prepare_groupby_batch([(d_key, [d_vf, d_vi])], aggs=N.AGG_ALL, want_perm=False).capture()
    ms_graph, _ = _time_steps(pb.replay, args.steps, args.warmup)
    rg = pb.result(0)
    assert rg.n_groups == g and torch.equal(rg.counts, res["r"].counts[:g]), "graph replay differs"
    for c in range(2):
        assert torch.equal(rg.aggs[c]["sum"], res["r"].aggs[c]["sum"][:g]), "graph replay differs"
    pb.close()
    rec["graph_replay"] = {"ms_per_step": ms_graph / args.steps, "rows_per_s": n * args.steps / (ms_graph * 1e-3)}
    if not args.no_cpu:
        T = _threads(args)
        _cpu_groupby_legs(rec, rows_per_s, [(key, vf, vi)], [(key, vf, vi)], T, f"full workload ({n} rows)",
                          f"full workload ({n} rows)", reps=5)
        rec["cpu_one_thread"] = _cpu_ref_groupby([(key, vf, vi)], 1, f"full workload ({n} rows), one thread",
                                                 reps=3)
        rec["cpu_reference_api_one_thread"] = _cpu_ref_groupby([(key, vf, vi)], 1,
                                                               f"full workload ({n} rows), one thread",
                                                               refapi=True, reps=3)
    v.close()
    return rec


def run_zipf_groupby(args) -> dict:
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.device import Vdf

    n = args.rows or 200_000_000
    _log(f"zipf_groupby: generating {n} rows")
    key, vf, vi = datagen.zipf_groupby(n)
    d_key = torch.from_numpy(key).cuda()
    d_vf = torch.from_numpy(vf).cuda()
    d_vi = torch.from_numpy(vi).cuda()
    v = _handle(args)
    res = {}

    def step():
        res["r"] = v.groupby(d_key, [d_vf, d_vi], aggs=N.AGG_ALL, want_perm=False, trim=False)

    ms_total, kernels = _time_steps(step, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    rows_per_s = n * args.steps / (ms_total * 1e-3)
    g = int(res["r"].n_groups.item())
    cnt = res["r"].counts[:g].to(torch.int64).sum().item()
    assert cnt == n, "group counts do not add up to the rows"
    rec = _base_record(args, rows_per_s, "rows/s", ms_step, "f64+i64",
                       {"workload": "zipf_groupby", "rows": n, "key": "int32 Zipf(1.1) over 2^24 values",
                        "values": ["float64 N(0,1)", "int64 uniform [0,2^40)"],
                        "aggregates": "count,sum,mean,min,max", "groups": g, "parallelism": "single GPU"})
    rec["roofline"] = _roofline("zipf_groupby", kernels)
    rec["kernels"] = kernels
    if not args.no_cpu:
        m = min(n, 5_000_000)
        _cpu_groupby_legs(rec, rows_per_s, [(key, vf, vi)], [(key[:m], vf[:m], vi[:m])], _threads(args),
                          f"full workload ({n} rows)", f"first {m} rows of the workload")
    v.close()
    return rec


# ---------------------------------------------------------------------------
# batched: 4096 tables in one segmented launch sequence
# ---------------------------------------------------------------------------
def run_batched(args) -> dict:
    import ctypes as C

    import numpy as np
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.device import Vdf, _stream_ptr

    ntab = args.tables or 4096
    sizes = datagen.batched_table_sizes(ntab, 1_000, 4_000_000, seed=4)
    total = int(sizes.sum())
    _log(f"batched: generating {ntab} tables, {total} rows")
    keys = np.empty(total, dtype=np.int32)
    vals = np.empty(total, dtype=np.float64)
    offs = np.zeros(ntab + 1, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)
    for i, n in enumerate(sizes):
        k, x = datagen.batched_table(int(n), i)
        keys[offs[i]:offs[i + 1]] = k
        vals[offs[i]:offs[i + 1]] = x
    d_keys = torch.from_numpy(keys).cuda()
    d_vals = torch.from_numpy(vals).cuda()
    v = _handle(args)
    caps = [max(int(n) // 8, 1) for n in sizes]  # keys are drawn from [0, rows/8)
    tables = [(d_keys[offs[i]:offs[i + 1]], [d_vals[offs[i]:offs[i + 1]]]) for i in range(ntab)]
    # descriptors and outputs built once: the timed region is the library call
    pb = v.prepare_groupby_batch(tables, aggs=N.AGG_ALL, want_perm=False, key_type=N.KEY_I32, capacities=caps)
    tune = None
    if not args.no_tune:
        _log("batched: autotuning the launch shape")
        tune = pb.autotune(reps=1)
        _log(f"batched: autotune {tune}")
    ms_total, kernels = _time_steps(pb.run, args.steps, args.warmup, v)
    ms_step = ms_total / args.steps
    rows_per_s = total * args.steps / (ms_total * 1e-3)
    ng = torch.cat([t["ng"] for t in pb.keep]).cpu().numpy()
    assert np.all(ng <= np.array(caps)), "a table has more groups than its capacity"
    cnt = sum(int(t["counts"][:g].to(torch.int64).sum().item()) for t, g in zip(pb.keep[:64], ng[:64]))
    assert cnt == int(sizes[:64].sum()), "group counts do not add up to the rows"

    if args.skip_compare:  # profiling runs: the batched launches only
        rec = _base_record(args, rows_per_s, "rows/s", ms_step, "i32+f64",
                           {"workload": "batched_groupby", "tables": ntab, "rows": total})
        rec["kernels"] = kernels
        v.close()
        return rec

    # the same batch as a captured HIP-graph plan
    pb.capture()
    ms_graph, _ = _time_steps(pb.replay, args.steps, 1)
    pb.release_plan()

    # one vdf_groupby call per table, same outputs, descriptors also built beforehand
    stream = C.c_void_p(_stream_ptr(None))
    cols = []
    for i, (k, vs) in enumerate(tables):
        c = (N.VdfValueCol * 1)()
        c[0].type = N.VAL_F64
        c[0].data = vs[0].data_ptr()
        cols.append(c)

    def per_table():
        for i, (k, _) in enumerate(tables):
            N.check(v.lib.vdf_groupby(v.h, N.KEY_I32, k.data_ptr(), k.shape[0], cols[i], 1, N.AGG_ALL,
                                      C.byref(pb.outs[i]), stream))

    per_table()
    torch.cuda.synchronize()
    ms_single, _ = _time_steps(per_table, 1, 0)
    rec = _base_record(args, rows_per_s, "rows/s", ms_step, "i32+f64",
                       {"workload": "batched_groupby", "tables": ntab, "rows": total,
                        "rows_per_table": "log-uniform [1K, 4M]", "key": "int32 uniform [0, rows/8)",
                        "values": ["float64 uniform [0,1)"], "aggregates": "count,sum,mean,min,max",
                        "parallelism": "single GPU"})
    rec["one_call_per_table"] = {"ms": ms_single, "rows_per_s": total / (ms_single * 1e-3),
                                 "batched_speedup": ms_single / ms_step}
    rec["graph_replay"] = {"ms_per_step": ms_graph / args.steps,
                           "rows_per_s": total * args.steps / (ms_graph * 1e-3)}
    rec["autotune"] = tune
    rec["roofline"] = _roofline("batched", kernels)
    rec["kernels"] = kernels
    if not args.no_cpu:
        # bounded samples: every 16th table (fused), every 256th (reference API)
        sample = [(keys[offs[i]:offs[i + 1]], vals[offs[i]:offs[i + 1]], None) for i in range(0, ntab, 16)]
        api = [(keys[offs[i]:offs[i + 1]], vals[offs[i]:offs[i + 1]], None) for i in range(0, ntab, 256)]
        _cpu_groupby_legs(rec, rows_per_s, sample, api, _threads(args),
                          f"every 16th table ({len(sample)} tables, {sum(t[0].shape[0] for t in sample)} rows), "
                          f"one call each",
                          f"every 256th table ({len(api)} tables, {sum(t[0].shape[0] for t in api)} rows)")
    v.close()
    return rec


# ---------------------------------------------------------------------------
# served: Poisson arrivals, continuous batching vs one launch per request
# ---------------------------------------------------------------------------
def _replay(srv, ids, rows, arrivals, picks):
    """Submits requests at their arrival times (busy-wait pacing) while a
    second thread waits for each result in order and releases it.  Returns
    per-request (latency s, queue s, batch size) and the makespan."""
    import numpy as np

    R = len(arrivals)
    rids = [None] * R
    ready = [threading.Event() for _ in range(R)]
    out = [None] * R

    def collector():
        for i in range(R):
            ready[i].wait()
            s = srv.wait(rids[i], timeout_s=60.0)
            if s is None:
                # a stall: report it and end the process at once (closing the
                # server would wait for the stalled work)
                _log(f"served: request {i} not finished after 60 s; stats {srv.stats()}")
                sys.stderr.flush()
                os._exit(3)
            out[i] = s
            srv.release(rids[i])
            if (i + 1) % 250 == 0:
                _log(f"served: {i + 1} of {R} results collected")

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    for i in range(R):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        rids[i] = srv.submit(ids[picks[i]], want_perm=False)
        ready[i].set()
    th.join()
    if any(s is None for s in out):
        raise RuntimeError("collector failed")
    lat = np.array([(s.done_ns - s.submit_ns) * 1e-9 for s in out])
    qs = np.array([(s.start_ns - s.submit_ns) * 1e-9 for s in out])
    bt = np.array([s.batch_tables for s in out])
    bad = [s.status for s in out if s.status != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} served requests failed (status {bad[0]})")
    t_first = min(s.submit_ns for s in out)
    span = (max(s.done_ns for s in out) - t_first) * 1e-9
    # queued requests (submitted, not yet started) over time, at 20 points of the makespan
    sub = np.sort(np.array([s.submit_ns - t_first for s in out], dtype=np.int64))
    sta = np.sort(np.array([s.start_ns - t_first for s in out], dtype=np.int64))
    pts = np.linspace(0, span * 1e9, 21)[1:]
    depth = (np.searchsorted(sub, pts, side="right") - np.searchsorted(sta, pts, side="right")).tolist()
    return lat, qs, bt, span, depth


def run_served(args) -> dict:
    import numpy as np
    import torch

    from visondf_amd import _native as N
    from visondf_amd import datagen
    from visondf_amd.server import Server

    T = args.tables or 32
    R = args.requests
    rate = args.rate
    sizes, kinds, arrivals, picks = datagen.served_trace(T, R, rate)
    _log(f"served: generating {T} tables, {int(sizes.sum())} rows")
    tables = []
    for i, (n, kind) in enumerate(zip(sizes, kinds)):
        k, x = datagen.served_table(int(n), i, kind)
        dk = torch.from_numpy(k.view(np.int64) if kind == "u64" else k).cuda()
        tables.append((kind, int(n), dk, torch.from_numpy(x).cuda(), k, x))
    torch.cuda.synchronize()
    rows = np.array([t[1] for t in tables])
    req_rows = int(rows[picks].sum())

    def serve(workers, max_batch):
        with Server(0, arena_bytes=args.arena_gb << 30, workers=workers, max_batch_tables=max_batch) as srv:
            ids = [srv.add_table(dk, [dx], key_type=N.KEY_U64 if kind == "u64" else N.KEY_CHAR16,
                                 max_groups=n // 16 + 1) for kind, n, dk, dx, _, _ in tables]
            for tid in ids:  # warm-up: every table once
                r = srv.submit(tid)
                s = srv.wait(r, timeout_s=120.0)
                if s is None or s.status != 0:
                    raise RuntimeError("warm-up request failed")
                srv.release(r)
            lat, qs, bt, span, depth = _replay(srv, ids, rows, arrivals, picks)
            st = srv.stats()
        return {"workers": workers, "max_batch_tables": max_batch,
                "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                "mean_ms": 1e3 * float(lat.mean()), "queue_p50_ms": 1e3 * float(np.percentile(qs, 50)),
                "rows_per_s": req_rows / span, "requests_per_s": R / span, "makespan_s": span,
                "queue_depth_over_makespan": depth, "max_queue_depth": int(max(depth)),
                "mean_batch": float(bt.mean()), "max_batch": int(bt.max()), "launches": st["launches"],
                "arena_peak_gb": st["arena_peak"] / 2**30}

    _log("served: continuous batching")
    batched = serve(args.workers, 64)
    _log(f"served: batched p50 {batched['p50_ms']:.2f} ms p99 {batched['p99_ms']:.2f} ms")
    _log("served: one launch per request")
    single = serve(1, 1)
    _log(f"served: single p50 {single['p50_ms']:.2f} ms p99 {single['p99_ms']:.2f} ms")
    rec = _base_record(args, batched["rows_per_s"], "rows/s", 1e3 * batched["makespan_s"], "u64/char16+f64",
                       {"workload": "served_groupby", "tables": T, "rows_per_table": "log-uniform [10K, 50M]",
                        "keys": "uint64 (even tables) / 16-byte ASCII (odd tables), rows/16 distinct",
                        "values": ["float64 uniform [0,1)"], "arrivals": f"Poisson {rate:g}/s",
                        "requests": R, "request_rows": req_rows, "parallelism": "single GPU"})
    rec["steps"] = R
    rec["warmup"] = T
    rec["served"] = {"continuous_batching": batched, "one_launch_per_request": single,
                     "p99_speedup": single["p99_ms"] / batched["p99_ms"],
                     "p50_speedup": single["p50_ms"] / batched["p50_ms"]}
    if not args.no_cpu:
        # every served table once (the requests' tables), one call each
        full = [(t[4], t[5], None) for t in tables]
        api = [(t[4], t[5], None) for t in tables if t[1] <= 2_000_000]
        _cpu_groupby_legs(rec, batched["rows_per_s"], full, api, _threads(args),
                          f"each of the {len(full)} served tables once ({sum(t[1] for t in tables)} rows)",
                          f"the {len(api)} served tables of <= 2M rows")
    return rec


RUNNERS = {"argsort_u32": run_argsort_u32, "groupby_c1": run_groupby_c1, "groupby_i32": run_groupby_i32,
           "groupby_i32_full": run_groupby_i32_full, "zipf_groupby": run_zipf_groupby, "batched": run_batched,
           "csv_ingest": run_csv_ingest, "served": run_served}


def _child(name: str, args) -> dict:
    steps, warmup = args.steps, args.warmup
    if name in ("groupby_i32", "groupby_i32_full", "zipf_groupby", "batched"):  # long steps: fewer of them
        steps, warmup = max(3, args.steps // 4), min(args.warmup, 2)
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", name, "--steps", str(steps),
           "--warmup", str(warmup)]
    if args.no_cpu:
        cmd.append("--no-cpu")
    if args.no_tune:
        cmd.append("--no-tune")
    if args.calibrate and name == "argsort_u32":
        cmd.append("--calibrate")
    _log(f"running {name}: {' '.join(cmd[1:])}")
    t0 = time.time()
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=CHILD_TIMEOUT[name])
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if p.returncode != 0 or not lines:
        raise RuntimeError(f"{name} exited with status {p.returncode}")
    rec = json.loads(lines[-1])
    rec["wall_s"] = time.time() - t0
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=WORKLOADS, default=None,
                    help="run one workload in this process (default: all, each in a child process)")
    ap.add_argument("--workloads", default=",".join(DEFAULT_WORKLOADS), help="workloads of the default run")
    ap.add_argument("--rows", type=int, default=0, help="rows of argsort_u32 / groupby_i32 / zipf_groupby (0 = the workload's)")
    ap.add_argument("--tables", type=int, default=0, help="tables of batched / served (0 = the workload's)")
    ap.add_argument("--requests", type=int, default=1000, help="served: requests replayed")
    ap.add_argument("--rate", type=float, default=2000.0, help="served: Poisson arrival rate, requests/s")
    ap.add_argument("--workers", type=int, default=3, help="served: launch sequences in flight")
    ap.add_argument("--arena-gb", type=int, default=48, help="served: server arena (admission budget), GiB")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner (argsort_u32, batched)")
    ap.add_argument("--skip-compare", action="store_true",
                    help="time the workload's own step only: argsort_u32 without the two-call (argsort + "
                         "gather) comparison, batched without per-table calls and graph replay")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tune", action="append", default=[], metavar="KNOB=VALUE",
                    help="set a launch-shape knob of the handle (vdf_tune_set), e.g. sort_items_32=12; "
                         "with --no-tune the setting is what is timed")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run one plain 1 GiB device copy (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is a design "
              f"note, not part of this build", file=sys.stderr)
        sys.exit(2)

    if args.workload == "served":
        print("served is switched off: its continuous-batching replay stopped making progress with batches "
              "in flight (cause not found)", file=sys.stderr)
        sys.exit(2)
    if args.workload:
        if args.calibrate:
            import torch

            a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)
            b.copy_(a)
            torch.cuda.synchronize()
            del a, b
        print(json.dumps(RUNNERS[args.workload](args)), flush=True)
        return

    names = [w for w in args.workloads.split(",") if w]
    head = _child("argsort_u32", args)
    others = {}
    for name in names:
        if name == "argsort_u32":
            continue
        try:
            others[name] = _child(name, args)
        except Exception as e:  # a failed secondary workload is reported, not fatal
            _log(f"{name} failed: {e}")
            others[name] = {"error": str(e)}
    head["workloads"] = others
    print(json.dumps(head), flush=True)


if __name__ == "__main__":
    main()
