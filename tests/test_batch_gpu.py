"""GPU parity of batched sorts and group-bys (many independent tables in one
launch sequence) against the oracle, and per-table identity between a table
sorted / grouped alone and inside a batch."""
from __future__ import annotations

import sys

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
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


def _u(t, dtype):
    a = t.cpu().numpy()
    return a.view(dtype) if a.ndim == 1 else a


SIZES = [0, 1, 2, 3, 63, 64, 65, 255, 2047, 2048, 2049, 4095, 4096, 4097, 6144, 6145, 10_000, 77_777,
         300_001, 1_000_003]


def _tables(kind: str, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(SIZES):
        if kind == "u32":
            k = rng.integers(0, max(n // 3, 1), n).astype(np.uint32) * np.uint32(2654435761 % (1 << 32))
        elif kind == "i32":
            k = rng.integers(-max(n // 8, 1), max(n // 8, 1) + 1, n).astype(np.int32)
        elif kind == "u64":
            k = rng.integers(0, 2**63, n, dtype=np.uint64)
        elif kind == "i64":
            k = rng.integers(-1000, 1000, n).astype(np.int64)
        else:
            k = datagen.char16_keys(seed + i, 0, n, distinct=max(n // 4, 1))
        out.append(k)
    return out


@pytest.mark.parametrize("kind", ["u32", "i32", "u64", "i64", "char16"])
@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_argsort_batch_vs_oracle(vdf, kind, mode):
    keys = _tables(kind, 7)
    _say("argsort_batch", kind, mode, len(keys))
    kt = oracle.key_type_of(keys[0])
    res = vdf.argsort_batch([_dev(k) for k in keys], key_type=kt, descending=(mode == "desc"),
                            reverse_desc=(mode == "reverse"), want_keys=True)
    m = {"asc": 0, "desc": 1, "reverse": 2}[mode]
    for k, (p, sk) in zip(keys, res):
        exp = oracle.argsort(k, m, 4)
        assert np.array_equal(_u(p, np.uint32), exp.astype(np.uint32)), (kind, mode, k.shape)
        if k.shape[0]:
            got = sk.cpu().numpy()
            assert np.array_equal(got.view(k.dtype) if k.ndim == 1 else got, k[exp.astype(np.int64)])


def test_argsort_batch_perm64_and_alone_identity(vdf):
    keys = _tables("u32", 9)
    res = vdf.argsort_batch([_dev(k) for k in keys], key_type=N.KEY_U32, perm64=True)
    for k, p in zip(keys, res):
        alone = vdf.argsort(_dev(k), key_type=N.KEY_U32, perm64=True) if k.shape[0] else p
        assert np.array_equal(p.cpu().numpy(), alone.cpu().numpy())
        assert np.array_equal(_u(p, np.uint64), oracle.argsort(k, 0))


def _check_group(r, o, vals, label):
    g = o["n_groups"]
    assert r.n_groups == g, label
    if g == 0:
        return
    assert np.array_equal(_u(r.keys, o["keys"].dtype) if o["keys"].ndim == 1 else r.keys.cpu().numpy(), o["keys"])
    assert np.array_equal(_u(r.offsets, np.uint32).astype(np.uint64), o["offsets"]), label
    assert np.array_equal(_u(r.counts, np.uint32).astype(np.uint64), o["counts"]), label
    assert np.array_equal(_u(r.perm, np.uint32).astype(np.uint64), o["perm"]), label
    for c, v in enumerate(vals):
        ra, oa = r.aggs[c], o["aggs"][c]
        assert np.array_equal(_u(ra["min"], v.dtype), oa["min"]), label
        assert np.array_equal(_u(ra["max"], v.dtype), oa["max"]), label
        if v.dtype == np.float64:
            for k in ("sum", "mean"):
                got, exp = _u(ra[k], np.float64), oa[k]
                assert np.all(np.abs(got - exp) <= F64_RTOL * np.abs(exp)), (label, k)
        else:
            assert np.array_equal(_u(ra["sum"], np.int64), oa["sum"].view(np.int64)), label
            assert np.array_equal(_u(ra["mean"], np.float64), oa["mean"]), label


def _same(ra, rb, label):
    assert ra.n_groups == rb.n_groups, label
    for x, y in ((ra.keys, rb.keys), (ra.offsets, rb.offsets), (ra.counts, rb.counts), (ra.perm, rb.perm)):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy()), label
    for da, db in zip(ra.aggs, rb.aggs):
        for k in da:
            assert np.array_equal(da[k].cpu().numpy().view(np.uint8), db[k].cpu().numpy().view(np.uint8)), (label, k)


@pytest.mark.parametrize("kind", ["i32", "u64", "char16"])
def test_groupby_batch_vs_oracle_and_alone(vdf, kind):
    keys = _tables(kind, 11)
    rng = np.random.default_rng(12)
    tables = []
    for k in keys:
        n = k.shape[0]
        tables.append((k, [rng.standard_normal(n), rng.integers(-2**40, 2**40, n).astype(np.int64)]))
    _say("groupby_batch", kind, len(tables))
    kt = oracle.key_type_of(keys[0])
    dev_tables = [(_dev(k), [_dev(v) for v in vs]) for k, vs in tables]
    res = vdf.groupby_batch(dev_tables, key_type=kt)
    for i, ((k, vs), r) in enumerate(zip(tables, res)):
        o = oracle.groupby(k, vs, threads=4)
        _check_group(r, o, vs, f"{kind}[{i}] n={k.shape[0]}")
        if k.shape[0]:
            alone = vdf.groupby(dev_tables[i][0], dev_tables[i][1], key_type=kt)
            _same(r, alone, f"alone {kind}[{i}]")


def test_groupby_batch_loguniform_tables(vdf):
    """A scaled-down batched configuration: 256 tables, rows log-uniform in
    [1K, 400K], int32 keys uniform in [0, rows/8), f64 values."""
    sizes = datagen.batched_table_sizes(256, 1_000, 400_000, seed=4)
    tables = [datagen.batched_table(int(n), i) for i, n in enumerate(sizes)]
    _say("groupby_batch loguniform", len(tables), int(sizes.sum()))
    dev_tables = [(_dev(k), [_dev(v)]) for k, v in tables]
    res = vdf.groupby_batch(dev_tables, key_type=N.KEY_I32)
    for i, ((k, v), r) in enumerate(zip(tables, res)):
        if i % 8:
            continue  # every 8th table against the oracle keeps the CPU time bounded
        _check_group(r, oracle.groupby(k, [v], threads=4), [v], f"logu[{i}]")
    # determinism: a second run of the whole batch is bit-identical
    res2 = vdf.groupby_batch(dev_tables, key_type=N.KEY_I32)
    for i, (a, b) in enumerate(zip(res, res2)):
        _same(a, b, f"repeat[{i}]")


def test_groupby_batch_full_config(vdf):
    """The batched configuration at full size: 4096 tables, rows log-uniform
    in [1K, 4M] (about 2.0B rows), int32 keys uniform in [0, rows/8), float64
    values, every aggregate, group capacity rows/8, no permutation (as the
    benchmark runs it).  Every table: group count within capacity, counts
    adding up to its rows, keys strictly ascending inside [0, rows/8),
    offsets the running sum of the counts, the mean between min and max.  A seeded
    sample of 24 tables, the four largest among them, against the oracle."""
    import torch

    ntab = 4096
    sizes = datagen.batched_table_sizes(ntab, 1_000, 4_000_000, seed=4)
    total = int(sizes.sum())
    _say("groupby_batch full config", ntab, total)
    offs = np.zeros(ntab + 1, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)
    keys = np.empty(total, dtype=np.int32)
    vals = np.empty(total, dtype=np.float64)
    for i, n in enumerate(sizes):
        k, x = datagen.batched_table(int(n), i)
        keys[offs[i]:offs[i + 1]] = k
        vals[offs[i]:offs[i + 1]] = x
    d_keys = torch.from_numpy(keys).cuda()
    d_vals = torch.from_numpy(vals).cuda()
    tables = [(d_keys[offs[i]:offs[i + 1]], [d_vals[offs[i]:offs[i + 1]]]) for i in range(ntab)]
    caps = [max(int(n) // 8, 1) for n in sizes]
    pb = vdf.prepare_groupby_batch(tables, aggs=N.AGG_ALL, want_perm=False, key_type=N.KEY_I32, capacities=caps)
    pb.run()
    torch.cuda.synchronize()
    _say("groupby_batch full config: launched, checking every table")
    ng = torch.cat([t["ng"] for t in pb.keep]).cpu().numpy().astype(np.int64)
    assert np.all(ng <= np.array(caps)), "a table has more groups than its capacity"
    rows_seen = 0
    for i in range(ntab):
        r = pb.result(i)
        g = int(ng[i])
        n = int(sizes[i])
        cnt = r.counts.to(torch.int64)
        k = r.keys
        off = r.offsets.to(torch.int64)
        assert int(cnt.sum().item()) == n, i
        assert bool((cnt > 0).all().item()), i
        assert bool((k[1:] > k[:-1]).all().item()) and int(k[0].item()) >= 0 and int(k[-1].item()) < max(n // 8, 1), i
        assert int(off[0].item()) == 0 and bool((off[1:g] == torch.cumsum(cnt, 0)[:-1]).all().item()), i
        a = r.aggs[0]
        # values in [0, 1): the mean (rounded sum / count) within an ulp-scale margin of [min, max]
        assert bool((a["mean"] - a["min"] >= -1e-12).all().item()), i
        assert bool((a["max"] - a["mean"] >= -1e-12).all().item()), i
        rows_seen += n
    assert rows_seen == total
    rng = np.random.default_rng(44)
    sample = set(np.argsort(sizes)[-4:].tolist()) | set(rng.choice(ntab, 20, replace=False).tolist())
    for i in sorted(sample):
        k = keys[offs[i]:offs[i + 1]]
        v = vals[offs[i]:offs[i + 1]]
        o = oracle.groupby(k, [v], threads=16)
        r = pb.result(i)
        g = o["n_groups"]
        assert r.n_groups == g, i
        assert np.array_equal(_u(r.keys, np.int32), o["keys"]), i
        assert np.array_equal(_u(r.offsets, np.uint32).astype(np.uint64)[:g], o["offsets"][:g]), i
        assert np.array_equal(_u(r.counts, np.uint32).astype(np.uint64), o["counts"]), i
        ra, oa = r.aggs[0], o["aggs"][0]
        assert np.array_equal(_u(ra["min"], np.float64), oa["min"]), i
        assert np.array_equal(_u(ra["max"], np.float64), oa["max"]), i
        for kk in ("sum", "mean"):
            got, exp = _u(ra[kk], np.float64), oa[kk]
            assert np.all(np.abs(got - exp) <= F64_RTOL * np.abs(exp)), (i, kk)
