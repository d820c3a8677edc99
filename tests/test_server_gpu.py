"""GPU tests of the group-by request server: every served result is identical
to the same group-by run alone (bit for bit, all outputs) while requests of
different sizes and key types arrive during other work, queue for arena
memory, share batches and overlap on several worker streams; plus oracle
parity of a sample, budget rejection and the group-capacity rerun."""
from __future__ import annotations

import random
import sys
import threading
import time

import numpy as np
import pytest

import oracle
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu

F64_RTOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


def _bits(t):
    return None if t is None else t.cpu().numpy().view(np.uint8)


def _same_as_alone(served: dict, alone, want_perm: bool, label: str, aggs: int = N.AGG_ALL):
    assert served["n_groups"] == alone.n_groups, label
    pairs = [("keys", served["keys"], alone.keys), ("offsets", served["offsets"], alone.offsets)]
    if aggs & N.AGG_COUNT:
        pairs.append(("counts", served["counts"], alone.counts))
    for name, x, y in pairs:
        assert np.array_equal(_bits(x), _bits(y)), (label, name)
    if want_perm:
        assert np.array_equal(_bits(served["perm"]), _bits(alone.perm)), (label, "perm")
    for da, db in zip(served["aggs"], alone.aggs):
        assert set(da) == set(db), label
        for k in da:
            assert np.array_equal(_bits(da[k]), _bits(db[k])), (label, k)


def _tables():
    """Uint64 and 16-byte-string tables of uneven sizes (row counts chosen to
    hit tile edges) with one f64 value column."""
    sizes = [1, 2, 2047, 2048, 4097, 30_000, 250_001, 1_000_000, 2_500_000]
    out = []
    for i, n in enumerate(sizes):
        for kind in ("u64", "char16"):
            k, v = datagen.served_table(n, 100 + i, kind)
            out.append((kind, k, v))
    return out


@pytest.fixture(scope="module")
def served_setup():
    import torch

    from visondf_amd.device import Vdf

    tabs = _tables()
    dev = [(_dev(k), _dev(v)) for _, k, v in tabs]
    torch.cuda.synchronize()
    v = Vdf(0)
    return tabs, dev, v


def _kt(kind):
    return N.KEY_U64 if kind == "u64" else N.KEY_CHAR16


def test_served_mix_identical_to_alone(served_setup):
    from visondf_amd.server import Server

    tabs, dev, v = served_setup
    rng = random.Random(7)
    # requests: every table several times with varying aggregate masks and perm outputs
    reqs = []
    for t in range(len(tabs)):
        for aggs, perm in ((N.AGG_ALL, True), (N.AGG_ALL, False), (N.AGG_SUM | N.AGG_COUNT, False),
                           (N.AGG_MIN | N.AGG_MAX, True)):
            reqs.append((t, aggs, perm))
    reqs = reqs * 2
    rng.shuffle(reqs)
    _say("served mix:", len(reqs), "requests over", len(tabs), "tables")
    with Server(0, arena_bytes=12 << 30, workers=3, max_batch_tables=16) as srv:
        ids = [srv.add_table(k, [x], key_type=_kt(kind)) for (kind, _, _), (k, x) in zip(tabs, dev)]
        rids = []
        for t, aggs, perm in reqs:  # arrivals spread over time, so later ones join running work
            rids.append(srv.submit(ids[t], aggs, perm))
            if rng.random() < 0.3:
                time.sleep(rng.random() * 2e-3)
        served = [srv.wait(r, timeout_s=120.0) for r in rids]
        assert all(s is not None for s in served), "a request did not finish within 120 s"
        batches = [s.batch_tables for s in served]
        _say("batch sizes seen:", sorted(set(batches)), "stats:", srv.stats())
        assert max(batches) > 1, "no request shared a launch sequence"
        alone_cache = {}
        for (t, aggs, perm), s, rid in zip(reqs, served, rids):
            assert s.status == N.VDF_OK
            kind = tabs[t][0]
            key = (t, aggs, perm)
            if key not in alone_cache:
                k, x = dev[t]
                alone_cache[key] = v.groupby(k, [x], aggs=aggs, want_perm=perm, key_type=_kt(kind))
            got = srv.tensors(ids[t], s, aggs, perm)
            _same_as_alone(got, alone_cache[key], perm, f"table {t} ({kind}, n={tabs[t][1].shape[0]}) aggs {aggs}",
                           aggs)
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["queued"] == 0
        assert st["arena_used"] == 0


def test_served_vs_oracle(served_setup):
    from visondf_amd.server import Server

    tabs, dev, _ = served_setup
    with Server(0, arena_bytes=2 << 30, workers=2, max_batch_tables=8) as srv:
        picks = [i for i, (_, k, _) in enumerate(tabs) if k.shape[0] <= 1_000_000]
        ids = {i: srv.add_table(dev[i][0], [dev[i][1]], key_type=_kt(tabs[i][0])) for i in picks}
        rids = {i: srv.submit(ids[i]) for i in picks}
        for i in picks:
            kind, k, x = tabs[i]
            s = srv.wait(rids[i], timeout_s=120.0)
            assert s is not None and s.status == N.VDF_OK
            got = srv.tensors(ids[i], s)
            o = oracle.groupby(k, [x], threads=4)
            assert s.n_groups == o["n_groups"]
            gk = got["keys"].cpu().numpy()
            assert np.array_equal(gk.view(np.uint64) if kind == "u64" else gk, o["keys"])
            assert np.array_equal(got["counts"].cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"])
            a, oa = got["aggs"][0], o["aggs"][0]
            assert np.array_equal(a["min"].cpu().numpy(), oa["min"])
            assert np.array_equal(a["max"].cpu().numpy(), oa["max"])
            for name in ("sum", "mean"):
                g = a[name].cpu().numpy()
                assert np.all(np.abs(g - oa[name]) <= F64_RTOL * np.abs(oa[name])), (i, name)
            srv.release(rids[i])


def test_served_budget_queueing_and_rejection(served_setup):
    """A small arena: requests queue until earlier results are released, and
    a request that can never fit fails with VDF_ERR_BUDGET."""
    from visondf_amd.server import Server

    tabs, dev, v = served_setup
    big = max(range(len(tabs)), key=lambda i: tabs[i][1].shape[0])
    small = [i for i, (_, k, _) in enumerate(tabs) if 2048 <= k.shape[0] <= 250_001]
    with Server(0, arena_bytes=64 << 20, workers=2, max_batch_tables=4) as srv:
        ids = {i: srv.add_table(dev[i][0], [dev[i][1]], key_type=_kt(tabs[i][0])) for i in small + [big]}
        rb = srv.submit(ids[big])
        s = srv.wait(rb, timeout_s=60.0)
        assert s is not None and s.status == N.VDF_ERR_BUDGET
        srv.release(rb)
        rids = [(i, srv.submit(ids[i])) for i in small * 4]
        done = []
        # a releaser thread frees results as they finish; without it the arena fills and the queue waits
        def releaser():
            for i, r in rids:
                s = srv.wait(r, timeout_s=120.0)
                assert s is not None
                done.append((i, r, s, srv.tensors(ids[i], s)))
                srv.release(r)
        th = threading.Thread(target=releaser)
        th.start()
        th.join(timeout=100)
        assert not th.is_alive()
        assert len(done) == len(rids)
        st = srv.stats()
        _say("budget stats:", st)
        assert st["arena_peak"] <= (64 << 20)
        for i, r, s, got in done:
            assert s.status == N.VDF_OK
            kind = tabs[i][0]
            alone = v.groupby(dev[i][0], [dev[i][1]], key_type=_kt(kind), want_perm=False)
            _same_as_alone(got, alone, False, f"budget table {i}")


def test_served_group_capacity_rerun(served_setup):
    """max_groups below the true group count: the server reruns the request
    once with the exact count and the result is still identical to alone."""
    from visondf_amd.server import Server

    tabs, dev, v = served_setup
    i = next(j for j, (kind, k, _) in enumerate(tabs) if kind == "u64" and k.shape[0] == 250_001)
    with Server(0, arena_bytes=1 << 30, workers=2, max_batch_tables=4) as srv:
        tid = srv.add_table(dev[i][0], [dev[i][1]], key_type=N.KEY_U64, max_groups=100)
        rid = srv.submit(tid, N.AGG_ALL, True)
        s = srv.wait(rid, timeout_s=60.0)
        assert s is not None and s.status == N.VDF_OK
        assert srv.stats()["retries"] == 1
        alone = v.groupby(dev[i][0], [dev[i][1]], key_type=N.KEY_U64)
        _same_as_alone(srv.tensors(tid, s, N.AGG_ALL, True), alone, True, "capacity rerun")
        srv.release(rid)


def test_groupby_capacity_truncation(vdf):
    """vdf_groupby with an output capacity below the group count writes the
    first groups only and reports the true count."""
    import torch

    k, x = datagen.served_table(100_000, 999, "u64")
    dk, dx = _dev(k), _dev(x)
    full = vdf.groupby(dk, [dx], key_type=N.KEY_U64)
    g = full.n_groups
    cap = g // 3
    r = vdf.groupby(dk, [dx], key_type=N.KEY_U64, capacity=cap, trim=False)
    torch.cuda.synchronize()
    assert int(r.n_groups.item()) == g
    assert np.array_equal(r.keys[:cap].cpu().numpy(), full.keys[:cap].cpu().numpy())
    assert np.array_equal(r.offsets[:cap].cpu().numpy(), full.offsets[:cap].cpu().numpy())
    assert np.array_equal(r.counts[:cap].cpu().numpy(), full.counts[:cap].cpu().numpy())
    for name in ("sum", "mean", "min", "max"):
        assert np.array_equal(_bits(r.aggs[0][name][:cap]), _bits(full.aggs[0][name][:cap])), name
    with pytest.raises(N.VdfError):
        vdf.groupby(dk, [dx], key_type=N.KEY_U64, capacity=cap)


def test_served_large_tables_vs_oracle():
    """Served tables at the configuration's sizes: the benchmark trace's two
    largest tables (32.7M rows of 16-byte keys, 29.1M rows of uint64 keys)
    and a 50M-row uint64 table (the configuration's upper bound), submitted
    between small requests to a 3-worker server; every result against the
    oracle (keys, counts, min, max exact; sums and means within 1e-12)."""
    import torch

    from visondf_amd.server import Server

    sizes, kinds, _, _ = datagen.served_trace()
    order = np.argsort(sizes)[::-1]
    big_c16 = int(next(i for i in order if kinds[i] == "char16"))
    big_u64 = int(next(i for i in order if kinds[i] == "u64"))
    specs = [("char16", int(sizes[big_c16]), big_c16), ("u64", int(sizes[big_u64]), big_u64),
             ("u64", 50_000_000, 1000), ("char16", 20_000, 1001), ("u64", 70_001, 1002)]
    tabs = []
    for kind, n, tid in specs:
        k, x = datagen.served_table(n, tid, kind)
        tabs.append((kind, k, x, _dev(k), _dev(x)))
    torch.cuda.synchronize()
    with Server(0, arena_bytes=48 << 30, workers=3, max_batch_tables=64) as srv:
        ids = [srv.add_table(dk, [dx], key_type=_kt(kind), max_groups=k.shape[0] // 16 + 1)
               for kind, k, _, dk, dx in tabs]
        order_sub = [3, 0, 4, 1, 3, 2, 4]
        rids = [(t, srv.submit(ids[t])) for t in order_sub]
        for t, rid in rids:
            kind, k, x, _, _ = tabs[t]
            _say("served large: table", t, kind, k.shape[0])
            s = srv.wait(rid, timeout_s=120.0)
            assert s is not None and s.status == N.VDF_OK, t
            got = srv.tensors(ids[t], s)
            o = oracle.groupby(k, [x], threads=16)
            assert s.n_groups == o["n_groups"], t
            gk = got["keys"].cpu().numpy()
            assert np.array_equal(gk.view(np.uint64) if kind == "u64" else gk, o["keys"]), t
            assert np.array_equal(got["offsets"].cpu().numpy().view(np.uint32).astype(np.uint64), o["offsets"]), t
            assert np.array_equal(got["counts"].cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"]), t
            a, oa = got["aggs"][0], o["aggs"][0]
            assert np.array_equal(a["min"].cpu().numpy(), oa["min"]), t
            assert np.array_equal(a["max"].cpu().numpy(), oa["max"]), t
            for name in ("sum", "mean"):
                g = a[name].cpu().numpy()
                assert np.all(np.abs(g - oa[name]) <= F64_RTOL * np.abs(oa[name])), (t, name)
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(rids) and st["failed"] == 0 and st["arena_used"] == 0
