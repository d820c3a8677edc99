"""GPU parity of the sort-based group-by against the oracle.

Bar: group keys, offsets, counts, the group-contiguous permutation, integer
sums (mod 2^64), integer means (exact sum / count), and min / max are
bit-exact; f64 sums and means agree with the correctly rounded exact sum
within a relative 1e-12 (F64_RTOL)."""
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


def _host(t, dtype):
    a = t.cpu().numpy()
    return a.view(dtype) if a.ndim == 1 else a


def _check(vdf, keys, vals, aggs=N.AGG_ALL, label="", threads=8):
    _say("groupby", label, keys.dtype, keys.shape, [v.dtype for v in vals])
    kt = oracle.key_type_of(keys)
    vts = [oracle.val_type_of(v) for v in vals]
    r = vdf.groupby(_dev(keys), [_dev(v) for v in vals], aggs=aggs, key_type=kt, val_types=vts)
    o = oracle.groupby(keys, vals, threads=threads)
    g = o["n_groups"]
    assert r.n_groups == g, label
    if g == 0:
        return r, o
    assert np.array_equal(_host(r.keys, keys.dtype).view(np.uint8), o["keys"].view(np.uint8)), label  # NaN keys: bits
    assert np.array_equal(_host(r.offsets, np.uint32).astype(np.uint64), o["offsets"]), label
    assert np.array_equal(_host(r.perm, np.uint32).astype(np.uint64), o["perm"]), label
    if aggs & N.AGG_COUNT:
        assert np.array_equal(_host(r.counts, np.uint32).astype(np.uint64), o["counts"]), label
    for c, v in enumerate(vals):
        ra, oa = r.aggs[c], o["aggs"][c]
        if aggs & N.AGG_MIN:
            assert np.array_equal(_host(ra["min"], v.dtype).view(np.uint8), oa["min"].view(np.uint8)), (label, c)
        if aggs & N.AGG_MAX:
            assert np.array_equal(_host(ra["max"], v.dtype).view(np.uint8), oa["max"].view(np.uint8)), (label, c)
        if v.dtype == np.float64:
            for k in ("sum", "mean"):
                if (aggs & (N.AGG_SUM if k == "sum" else N.AGG_MEAN)):
                    got, exp = _host(ra[k], np.float64), oa[k]
                    err = np.abs(got - exp)
                    assert np.all(err <= F64_RTOL * np.abs(exp)), (label, c, k, float(err.max()))
        else:
            if aggs & N.AGG_SUM:
                assert np.array_equal(_host(ra["sum"], np.int64), oa["sum"].view(np.int64)), (label, c)
            if aggs & N.AGG_MEAN:
                assert np.array_equal(_host(ra["mean"], np.float64), oa["mean"]), (label, c)
    return r, o


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_groupby_float_keys(vdf, dt):
    """Float keys group by their bits in the order of sort_flt.inl:44-50:
    -0.0 and +0.0 are separate groups, equal NaNs one group."""
    rng = np.random.default_rng(31)
    n = 250_001
    k = np.round(rng.standard_normal(n) * 100).astype(dt) / dt(4)
    sp = np.array([np.nan, np.copysign(np.nan, -1), np.inf, -np.inf, 0.0, -0.0], dtype=dt)
    k[rng.integers(0, n, 5000)] = sp[rng.integers(0, sp.shape[0], 5000)]
    _check(vdf, k, [rng.standard_normal(n), rng.integers(-1000, 1000, n).astype(np.int32)], label=f"float {dt}")


def test_groupby_small_config(vdf):
    """1M rows, int64 key in [0, 1000), f64 and int32 values: all aggregates."""
    key, vf, vi = datagen.groupby_small(1_000_000, seed=1)
    _check(vdf, key, [vf, vi], label="c1")


@pytest.mark.parametrize("kt", ["u32", "i32", "u64", "i64", "char16"])
def test_groupby_key_types(vdf, kt):
    rng = np.random.default_rng(31)
    n = 300_007
    if kt == "u32":
        keys = (rng.integers(0, 5000, n) * 858_993).astype(np.uint32)
    elif kt == "i32":
        keys = rng.integers(-3000, 3000, n).astype(np.int32)
    elif kt == "u64":
        keys = rng.integers(0, 7000, n).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    elif kt == "i64":
        keys = rng.integers(-2**62, 2**62, 4000)[rng.integers(0, 4000, n)].astype(np.int64)
    else:
        keys = datagen.char16_keys(5, 0, n, distinct=2000)
    vals = [rng.standard_normal(n), rng.integers(-2**40, 2**40, n), rng.integers(-2**31, 2**31, n).astype(np.int32),
            rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)]
    _check(vdf, keys, vals, label=kt)


def test_groupby_u64_values_and_wraparound(vdf):
    n = 50_000
    keys = (np.arange(n) % 3).astype(np.int32)
    big = np.full(n, 2**62 + 12345, dtype=np.int64)  # group sums wrap mod 2^64
    u = np.full(n, 2**64 - 7, dtype=np.uint64)
    _check(vdf, keys, [big, u], label="wrap")


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 8191, 8193, 100_000])
def test_groupby_tile_edges(vdf, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, max(n // 5, 1), n).astype(np.int32)
    _check(vdf, keys, [rng.standard_normal(n), rng.integers(-100, 100, n)], label=f"edge{n}")


def test_groupby_single_group_spans_tiles(vdf):
    n = 3_000_001  # ~730 tiles, one group: the fix-up walks every tile
    vals = datagen.normal01(3, 1, n)
    _check(vdf, np.zeros(n, dtype=np.int64), [vals, np.arange(n, dtype=np.int64)], label="one-group")


def test_groupby_group_spans_fixup_windows(vdf):
    """One group of 8.5M rows (about 2,080 tiles, more than one fix-up window of
    2,048 tiles) between groups of a few rows, and a second long group that
    ends exactly at the table's end."""
    rng = np.random.default_rng(21)
    keys = np.concatenate([rng.integers(0, 1000, 5000), np.full(8_500_000, 5000), rng.integers(6000, 7000, 7001),
                           np.full(3 * 4096 + 17, 9000)]).astype(np.int32)
    perm = rng.permutation(keys.size)
    keys = keys[perm]
    n = keys.size
    _check(vdf, keys, [datagen.normal01(21, 1, n), rng.integers(-(2**40), 2**40, n)], label="windows")


@pytest.mark.parametrize("items", [8, 12, 16])
def test_groupby_carried_pass_tiles(vdf, items):
    """The group-by's passes carry a 16-byte payload (two 8-byte columns) under
    every pass tile size, including the split-staging ones: results equal the
    oracle's."""
    key, vf, vi = datagen.zipf_groupby(3_000_017, seed=31)
    old = vdf.tune_get("sort_items_32")
    try:
        vdf.tune("sort_items_32", items)
        _check(vdf, key, [vf, vi], label=f"carried items {items}")
        _check(vdf, key, [vf], label=f"carried f64 items {items}")
    finally:
        vdf.tune("sort_items_32", old)


def test_groupby_wide_keys_split_tiles(vdf):
    """u64 and char[16] keys with an f64 column carried through split-staged
    16-row-per-thread passes: equal to the oracle."""
    old = (vdf.tune_get("sort_items_64"), vdf.tune_get("sort_items_128"))
    try:
        vdf.tune("sort_items_64", 16)
        vdf.tune("sort_items_128", 16)
        for kind in ("u64", "char16"):
            key, val = datagen.served_table(300_007, 3, kind)
            _check(vdf, key, [val], label=f"split {kind}")
    finally:
        vdf.tune("sort_items_64", old[0])
        vdf.tune("sort_items_128", old[1])


def test_groupby_every_row_distinct(vdf):
    n = 262_147
    keys = np.random.default_rng(1).permutation(n).astype(np.uint32)
    _check(vdf, keys, [np.random.default_rng(2).standard_normal(n)], label="distinct")


def test_groupby_group_sizes_near_tile(vdf):
    """Runs of 4095..4097 equal keys so heads land on every tile offset."""
    sizes = np.array([4095, 4096, 4097, 1, 8192, 12289, 3, 4096] * 6)
    keys = np.repeat(np.arange(sizes.size, dtype=np.int64), sizes)
    perm = np.random.default_rng(5).permutation(keys.size)
    keys = keys[perm]
    _check(vdf, keys, [np.random.default_rng(6).standard_normal(keys.size)], label="near-tile")


def test_groupby_counts_only_and_masks(vdf):
    key, vf, vi = datagen.groupby_small(200_000, seed=2)
    _check(vdf, key, [], aggs=N.AGG_COUNT, label="counts-only")
    _check(vdf, key, [vf, vi], aggs=N.AGG_SUM | N.AGG_MIN, label="sum+min")


def test_groupby_empty(vdf):
    import torch

    r = vdf.groupby(torch.empty(0, dtype=torch.int32, device="cuda"), [])
    assert r.n_groups == 0


def test_groupby_zipf_20m(vdf):
    """Zipf(1.1) keys over 2^24 values (the skewed configuration at 20M rows)."""
    key, vf, vi = datagen.zipf_groupby(20_000_000, seed=3)
    _check(vdf, key, [vf, vi], label="zipf20m")


def test_groupby_zipf_200m(vdf):
    """The skewed configuration at full size: 200M rows, int32 keys Zipf(1.1)
    over 2^24 values (about 8.6M groups, the top one about 23M rows), float64
    N(0,1) and int64 uniform [0, 2^40) values.  Every group's key, offsets,
    count, row ids, min, max, sums and means against the oracle; then the
    1000 largest groups' int64 sums (the top group's wraps past 2^63) from
    their rows by numpy's wrapping sum, and the 20 largest groups' float64
    sums against math.fsum of their rows."""
    import math

    key, vf, vi = datagen.zipf_groupby(200_000_000, seed=3)
    r, o = _check(vdf, key, [vf, vi], label="zipf200m", threads=16)
    cnt = o["counts"].astype(np.int64)
    off = o["offsets"].astype(np.int64)
    perm = _host(r.perm, np.uint32).astype(np.int64)
    gsum_f = _host(r.aggs[0]["sum"], np.float64)
    gsum_i = _host(r.aggs[1]["sum"], np.int64)
    top = np.argsort(cnt, kind="stable")[::-1][:1000]
    assert cnt[top[0]] > 22_000_000
    wrapped = 0
    for j, g in enumerate(top):
        rows = perm[off[g]:off[g + 1]]
        s = int(vi[rows].sum())  # numpy int64 sum wraps modulo 2^64
        exact = int(vi[rows].astype(object).sum()) if j == 0 else None
        if exact is not None:
            assert exact > 2**63 - 1  # the top group's true sum overflows int64
            wrapped += 1
            assert (exact - s) % 2**64 == 0
        assert gsum_i[g] == s, (g, int(gsum_i[g]), s)
        if j < 20:
            e = math.fsum(vf[rows].tolist())
            assert abs(gsum_f[g] - e) <= F64_RTOL * abs(e), (g, float(gsum_f[g]), e)
    assert wrapped == 1


def test_groupby_100m_i32(vdf):
    """The 100M-row int32-key group-by (key uniform in [0, 1000), float64 and
    int32 values): every output against the oracle at full size."""
    key, vf, vi = datagen.groupby_i32(100_000_000)
    _check(vdf, key, [vf, vi], label="i32 100m")


@pytest.mark.parametrize("items", [0, 8, 12, 16])
@pytest.mark.parametrize("n", [1, 4095, 4097, 12_289, 1_000_003])
def test_groupby_carried_8_4_pair(vdf, items, n):
    """4-byte keys carrying an 8-byte and a 4-byte column (12-byte payload)
    under every pass tile size, with the permutation (row-id kernels) and
    without it (row-id-free kernel): the two calls agree bit for bit outside
    the permutation, and the first matches the oracle."""
    rng = np.random.default_rng(n + items)
    keys = rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32) // max(n // 50, 1)
    vals = [rng.standard_normal(n), rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32)]
    d_keys, d_vals = _dev(keys), [_dev(v) for v in vals]
    old = vdf.tune_get("sort_items_32")
    try:
        vdf.tune("sort_items_32", items)
        with_p = vdf.groupby(d_keys, d_vals, want_perm=True)
        without = vdf.groupby(d_keys, d_vals, want_perm=False)
    finally:
        vdf.tune("sort_items_32", old)
    ref = vdf.groupby(d_keys, d_vals, want_perm=True)  # default tiles
    assert with_p.n_groups == without.n_groups == ref.n_groups
    assert _outputs(with_p, vals) == _outputs(ref, vals), (n, items)
    a, b = _outputs(with_p, vals), _outputs(without, vals)
    del a[3], b[3]
    assert all(x == y for x, y in zip(a, b)), (n, items)
    _check(vdf, keys, vals, label=f"8+4 pair n={n} items={items}")


def test_groupby_deterministic_repeat(vdf):
    """Same table twice (and after other work) gives bit-identical results."""
    key, vf, _ = datagen.zipf_groupby(2_000_000, seed=8)
    r1 = vdf.groupby(_dev(key), [_dev(vf)])
    other = datagen.groupby_small(300_000, seed=9)
    vdf.groupby(_dev(other[0]), [_dev(other[1])])
    r2 = vdf.groupby(_dev(key), [_dev(vf)])
    for k in ("sum", "mean", "min", "max"):
        assert np.array_equal(r1.aggs[0][k].cpu().numpy().view(np.uint64), r2.aggs[0][k].cpu().numpy().view(np.uint64))


def _outputs(r, vals):
    """Every output of a group-by result as bytes (for bit-identity checks)."""
    out = [r.keys.cpu().numpy().tobytes(), r.offsets.cpu().numpy().tobytes(), r.counts.cpu().numpy().tobytes(),
           r.perm.cpu().numpy().tobytes() if r.perm is not None else b""]
    for c in range(len(vals)):
        for k in ("sum", "mean", "min", "max"):
            out.append(r.aggs[c][k].cpu().numpy().tobytes())
    return out


@pytest.mark.parametrize("kind", ["u64", "char16"])
@pytest.mark.parametrize("items", [0, 8, 12, 16])
def test_groupby_wide_keys_without_permutation(vdf, kind, items):
    """8- and 16-byte keys with carried columns and no permutation (the passes
    move no row ids): every other output is bit-identical to the call with the
    permutation, which matches the oracle."""
    knob = "sort_items_64" if kind == "u64" else "sort_items_128"
    key, val = datagen.served_table(400_009, 7, kind)
    ival = np.random.default_rng(7).integers(-(2**62), 2**62, key.shape[0], dtype=np.int64)
    old = vdf.tune_get(knob)
    for vals in ([val], [val, ival]):  # 8- and 16-byte payloads
        d_keys, d_vals = _dev(key), [_dev(v) for v in vals]
        try:
            if items and (kind == "u64" or items in (8, 16)):  # 16-byte keys: 8 or 16 rows per thread
                vdf.tune(knob, items)
            with_p = vdf.groupby(d_keys, d_vals, want_perm=True)
            without = vdf.groupby(d_keys, d_vals, want_perm=False)
        finally:
            vdf.tune(knob, old)
        assert with_p.n_groups == without.n_groups
        a, b = _outputs(with_p, vals), _outputs(without, vals)
        del a[3], b[3]
        assert all(x == y for x, y in zip(a, b)), (kind, items, len(vals))
        _check(vdf, key, vals, label=f"{kind} reference")


@pytest.mark.parametrize("combo", ["f64", "i32", "u32", "i64", "f64+i64", "f64+u64+i32", "i32+f64", "u64+f64+i64+i32",
                                   "f64+i32", "f64+u32", "i64+i32+f64"])
def test_groupby_carried_columns_bit_identical(vdf, combo):
    """Value columns carried through the sort (the default) give results
    bit-identical to gathering every column through the permutation, for every
    mix of carried and gathered columns; and both match the oracle."""
    rng = np.random.default_rng(41)
    n = 300_007
    keys = datagen.scramble24(datagen.zipf_ranks(5, 0, n, 1.1, 1 << 16)).astype(np.int32)
    make = {"f64": lambda: rng.standard_normal(n) * 1e3,
            "i32": lambda: rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32),
            "u32": lambda: rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
            "i64": lambda: rng.integers(-(2**62), 2**62, n, dtype=np.int64),
            "u64": lambda: rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2)}
    vals = [make[t]() for t in combo.split("+")]
    _check(vdf, keys, vals, label=f"carried {combo}")
    kt = oracle.key_type_of(keys)
    vts = [oracle.val_type_of(v) for v in vals]
    d_keys, d_vals = _dev(keys), [_dev(v) for v in vals]
    assert vdf.tune_get("groupby_carry") == 1
    a = _outputs(vdf.groupby(d_keys, d_vals, key_type=kt, val_types=vts), vals)
    vdf.tune("groupby_carry", 0)
    try:
        b = _outputs(vdf.groupby(d_keys, d_vals, key_type=kt, val_types=vts), vals)
    finally:
        vdf.tune("groupby_carry", 1)
    assert all(x == y for x, y in zip(a, b)), combo


@pytest.mark.parametrize("combo", ["", "f64", "i32", "f64+i64", "f64+u64+i32"])
def test_groupby_without_permutation_identical(vdf, combo):
    """A group-by whose caller takes no permutation (and whose value columns are
    all carried, or which counts only) sorts without row ids; every other output
    is bit-identical to the call that takes the permutation."""
    rng = np.random.default_rng(43)
    n = 200_003
    keys = datagen.scramble24(datagen.zipf_ranks(6, 0, n, 1.1, 1 << 14)).astype(np.int32)
    make = {"f64": lambda: rng.standard_normal(n),
            "i32": lambda: rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32),
            "i64": lambda: rng.integers(-(2**62), 2**62, n, dtype=np.int64),
            "u64": lambda: rng.integers(0, 2**63, n, dtype=np.uint64)}
    vals = [make[t]() for t in combo.split("+") if t]
    kt = oracle.key_type_of(keys)
    vts = [oracle.val_type_of(v) for v in vals]
    d_keys, d_vals = _dev(keys), [_dev(v) for v in vals]
    with_p = vdf.groupby(d_keys, d_vals, key_type=kt, val_types=vts, want_perm=True)
    without = vdf.groupby(d_keys, d_vals, key_type=kt, val_types=vts, want_perm=False)
    assert with_p.n_groups == without.n_groups
    a, b = _outputs(with_p, vals), _outputs(without, vals)
    del a[3], b[3]  # the permutation itself
    assert all(x == y for x, y in zip(a, b)), combo
    o = oracle.groupby(keys, vals, threads=8)
    assert np.array_equal(_host(without.counts, np.uint32).astype(np.uint64), o["counts"])


def test_groupby_100m_i32_full_range(vdf):
    """100M rows of int32 keys uniform over the full 32-bit range (about 98M
    groups, four radix passes, output-heavy), float64 and int32 values:
    every output against the oracle at full size."""
    key, vf, vi = datagen.groupby_i32_full(100_000_000)
    r, o = _check(vdf, key, [vf, vi], label="i32 full 100m", threads=16)
    assert o["n_groups"] > 95_000_000
