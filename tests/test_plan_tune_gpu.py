"""GPU tests of HIP-graph batch plans and the autotuner: every graph replay
(with other eager work and synchronisations between replays, and after
re-pointing the plan at other tables of the same shape) equals the eager
batched group-by bit for bit; tuned launch shapes leave results unchanged."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def _bits(t):
    return t.cpu().numpy().view(np.uint8)


def _same(ra, rb, label):
    assert ra.n_groups == rb.n_groups, label
    for x, y in ((ra.keys, rb.keys), (ra.offsets, rb.offsets), (ra.counts, rb.counts)):
        assert np.array_equal(_bits(x), _bits(y)), label
    if ra.perm is not None:
        assert np.array_equal(_bits(ra.perm), _bits(rb.perm)), label
    for da, db in zip(ra.aggs, rb.aggs):
        for k in da:
            assert np.array_equal(_bits(da[k]), _bits(db[k])), (label, k)


def _batch(seed, sizes):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        k = rng.integers(0, max(n // 5, 1), n).astype(np.int32)
        out.append((k, [rng.standard_normal(n), rng.integers(-2**40, 2**40, n).astype(np.int64)]))
    return out


SIZES = [1, 4096, 0, 77_777, 2049, 300_001, 12_345]


def test_plan_replays_equal_eager(vdf):
    import torch

    host_a = _batch(1, SIZES)
    host_b = _batch(2, SIZES)
    tabs_a = [(_dev(k), [_dev(x) for x in vs]) for k, vs in host_a]
    tabs_b = [(_dev(k), [_dev(x) for x in vs]) for k, vs in host_b]
    eager_a = vdf.groupby_batch(tabs_a, key_type=N.KEY_I32)
    eager_b = vdf.groupby_batch(tabs_b, key_type=N.KEY_I32)
    pb = vdf.prepare_groupby_batch(tabs_a, key_type=N.KEY_I32).capture()
    side_k = _dev(np.arange(50_000, dtype=np.int32) % 77)
    for rep in range(3):
        _say("plan replay", rep)
        pb.replay()
        torch.cuda.synchronize()
        for i, e in enumerate(eager_a):
            _same(pb.result(i), e, f"replay {rep} table {i}")
        # the program's own eager work between replays
        vdf.groupby(side_k, [])
        torch.cuda.synchronize()
    pb.set_tables(tabs_b)
    for rep in range(2):
        pb.replay()
        torch.cuda.synchronize()
        for i, e in enumerate(eager_b):
            _same(pb.result(i), e, f"tables b replay {rep} table {i}")
        tabs_a[0][0].add_(0)  # unrelated eager kernel
    # oracle check of one replayed table
    k, vs = host_b[5]
    o = oracle.groupby(k, vs, threads=4)
    r = pb.result(5)
    assert r.n_groups == o["n_groups"]
    assert np.array_equal(r.counts.cpu().numpy().astype(np.uint64), o["counts"])
    with pytest.raises(ValueError):
        pb.set_tables(tabs_a[:-1] + [(tabs_a[0][0], tabs_a[0][1])])
    pb.release_plan()


def test_plan_without_permutation_equals_eager(vdf):
    """A plan whose outputs take no permutation replays equal to the eager call
    without one (both run the sort without row ids)."""
    import torch

    tabs = [(_dev(k), [_dev(x) for x in vs]) for k, vs in _batch(4, SIZES)]
    eager = vdf.groupby_batch(tabs, key_type=N.KEY_I32, want_perm=False)
    pb = vdf.prepare_groupby_batch(tabs, key_type=N.KEY_I32, want_perm=False).capture()
    for rep in range(2):
        pb.replay()
        torch.cuda.synchronize()
        for i, e in enumerate(eager):
            _same(pb.result(i), e, f"no-perm replay {rep} table {i}")
    pb.release_plan()


def test_plan_switches_permutation_between_replays(vdf):
    """One plan replayed with permutation outputs, then without (the variant
    captured without row ids), then with them again: every replay equals the
    eager call with the same outputs."""
    import torch

    tabs = [(_dev(k), [_dev(x) for x in vs]) for k, vs in _batch(5, SIZES)]
    eager = vdf.groupby_batch(tabs, key_type=N.KEY_I32, want_perm=True)
    pb = vdf.prepare_groupby_batch(tabs, key_type=N.KEY_I32, want_perm=True).capture()
    perms = [pb.outs[i].perm for i in range(pb.ntab)]
    for rep, with_perm in enumerate((True, False, True, False)):
        _say("plan replay", rep, "with permutation" if with_perm else "without permutation")
        for i in range(pb.ntab):
            pb.outs[i].perm = perms[i] if with_perm else None
            if not with_perm and pb.keep[i]["perm"] is not None:
                pb.keep[i]["perm"].fill_(-1)
        pb.replay()
        torch.cuda.synchronize()
        for i, e in enumerate(eager):
            r = pb.result(i)
            if not with_perm:
                # nothing writes the permutation buffers on this replay
                if r.perm is not None:
                    assert bool((r.perm == -1).all()), f"replay {rep} table {i} wrote a permutation"
                r.perm = None
            _same(r, e, f"replay {rep} table {i}")
    pb.release_plan()


def test_plan_rejects_other_shape(vdf):
    import ctypes as C

    tabs = [(_dev(k), [_dev(x) for x in vs]) for k, vs in _batch(3, [1000, 2000])]
    pb = vdf.prepare_groupby_batch(tabs, key_type=N.KEY_I32).capture()
    pb.tabs[1].n = 1999  # a row count the plan was not captured for
    with pytest.raises(N.VdfError):
        N.check(vdf.lib.vdf_plan_launch(pb.plan, pb.tabs, pb.ntab, pb.outs, C.c_void_p(0)))
    pb.tabs[1].n = 2000
    pb.release_plan()


def test_autotune_argsort_and_identity(vdf):
    import torch

    n = 3_000_000
    keys, _ = datagen.argsort_u32(n, seed=21)
    dk = _dev(keys)
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    before = vdf.argsort(dk, key_type=N.KEY_U32).cpu().numpy()
    rep = vdf.autotune_argsort(dk, perm, key_type=N.KEY_U32, reps=2)
    _say("autotune argsort:", rep)
    assert rep["trials"] >= 5 and rep["digit_bits"] == 8
    assert rep["best_ms"] <= rep["default_ms"] * 1.0001
    assert vdf.tune_get("sort_items_32") == rep["sort_items"]
    assert vdf.tune_get("sort_blocks_per_cu") == rep["sort_blocks_per_cu"]
    assert vdf.tune_get("hist_rows") == rep["hist_rows"]
    after = vdf.argsort(dk, key_type=N.KEY_U32).cpu().numpy()
    assert np.array_equal(before, after)
    assert np.array_equal(after.view(np.uint32).astype(np.int64), oracle.argsort(keys, 0, 8).astype(np.int64))
    for name, val in (("sort_items_32", 0), ("sort_blocks_per_cu", 0), ("sort_xcd", 1), ("sort_chunk", 64),
                      ("hist_rows", 65536)):
        vdf.tune(name, val)


@pytest.mark.parametrize("knob,values", [("sort_items_32", [12, 16, 24, 32, 0]), ("sort_blocks_per_cu", [1, 2, 0]),
                                         ("sort_xcd", [0, 1]), ("hist_rows", [4096, 262144, 65536])])
def test_results_independent_of_launch_shape(vdf, knob, values):
    tabs = [(_dev(k), [_dev(x) for x in vs]) for k, vs in _batch(4, [100_000, 1_000_003, 5000])]
    ref = None
    for v in values:
        vdf.tune(knob, v)
        res = vdf.groupby_batch(tabs, key_type=N.KEY_I32)
        if ref is None:
            ref = res
        else:
            for i, (a, b) in enumerate(zip(ref, res)):
                _same(a, b, f"{knob}={v} table {i}")


def test_autotune_groupby_batch(vdf):
    import torch

    sizes = datagen.batched_table_sizes(64, 1_000, 400_000, seed=9)
    host = [datagen.batched_table(int(n), i, seed=9) for i, n in enumerate(sizes)]
    tabs = [(_dev(k), [_dev(x)]) for k, x in host]
    eager = vdf.groupby_batch(tabs, key_type=N.KEY_I32)
    pb = vdf.prepare_groupby_batch(tabs, key_type=N.KEY_I32)
    rep = pb.autotune(reps=2)
    _say("autotune batch:", rep)
    assert rep["trials"] >= 5
    torch.cuda.synchronize()
    for i, e in enumerate(eager):
        _same(pb.result(i), e, f"tuned table {i}")
    for name, val in (("sort_items_32", 0), ("sort_blocks_per_cu", 0), ("sort_xcd", 1), ("sort_chunk", 64),
                      ("hist_rows", 65536)):
        vdf.tune(name, val)
