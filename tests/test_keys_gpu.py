"""GPU parity of multi-column keys (group-by and argsort through the packed
composite key) and of the per-row broadcast of group results, against the
oracle."""
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


def _bytes(t):
    return t.cpu().numpy().view(np.uint8).reshape(-1)


def _key_cols(combo, n, rng):
    out = []
    for kind in combo:
        if kind == "i32":
            out.append(rng.integers(-30, 30, n).astype(np.int32))
        elif kind == "u32":
            out.append(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) % np.uint32(40))
        elif kind == "i64":
            out.append(rng.integers(-2**62, 2**62, n).astype(np.int64) % 17 - 8)
        elif kind == "u64":
            out.append(datagen.splitmix64(9, 0, n) % np.uint64(25) * np.uint64(0x9E3779B97F4A7C15))
        elif kind == "f32":
            f = rng.integers(-4, 4, n).astype(np.float32) / np.float32(2)
            f[rng.integers(0, n, n // 50)] = np.float32(-0.0)
            out.append(f)
        elif kind == "f64":
            f = rng.integers(-6, 6, n).astype(np.float64) * 0.25
            f[rng.integers(0, n, n // 50)] = np.nan
            out.append(f)
        else:
            out.append(datagen.char16_keys(11, 0, n, distinct=12))
    return out


COMBOS = [("i32", "i32"), ("u32",), ("i64",), ("u32", "i64"), ("i64", "f64"), ("char16",), ("f32", "i32", "u32"),
          ("u64", "i32"), ("i32", "u32", "i32", "u32")]


@pytest.mark.parametrize("combo", COMBOS, ids=lambda c: "+".join(c))
def test_groupby_multi_vs_oracle(vdf, combo):
    rng = np.random.default_rng(len(combo) * 7 + 1)
    n = 300_007
    keys = _key_cols(combo, n, rng)
    vals = [rng.standard_normal(n), rng.integers(-2**40, 2**40, n).astype(np.int64)]
    _say("groupby_multi", combo, n)
    kts = [oracle.key_type_of(k) for k in keys]  # device tensors of unsigned data are signed views
    r = vdf.groupby_multi([_dev(k) for k in keys], [_dev(v) for v in vals], key_types=kts)
    o = oracle.groupby_multi(keys, vals, threads=8)
    assert r.n_groups == o["n_groups"]
    for c, k in enumerate(keys):
        assert np.array_equal(_bytes(r.keys[c]), o["keys"][c].view(np.uint8).reshape(-1)), (combo, c)
    assert np.array_equal(r.counts.cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"])
    assert np.array_equal(r.offsets.cpu().numpy().view(np.uint32).astype(np.uint64), o["offsets"])
    assert np.array_equal(r.perm.cpu().numpy().view(np.uint32).astype(np.uint64), o["perm"])
    a, oa = r.aggs[0], o["aggs"][0]
    got = a["sum"].cpu().numpy()
    assert np.all(np.abs(got - oa["sum"]) <= F64_RTOL * np.abs(oa["sum"]))
    assert np.array_equal(_bytes(a["min"]), oa["min"].view(np.uint8))
    assert np.array_equal(r.aggs[1]["sum"].cpu().numpy(), o["aggs"][1]["sum"].view(np.int64))


@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
@pytest.mark.parametrize("combo", [("i32", "u32"), ("u64", "i32"), ("char16",), ("f64", "i32")],
                         ids=lambda c: "+".join(c) if isinstance(c, tuple) else c)
def test_argsort_multi_vs_oracle(vdf, combo, mode):
    rng = np.random.default_rng(3)
    keys = _key_cols(combo, 200_003, rng)
    m = {"asc": 0, "desc": 1, "reverse": 2}[mode]
    perm = vdf.argsort_multi([_dev(k) for k in keys], key_types=[oracle.key_type_of(k) for k in keys],
                             descending=(m == 1), reverse_desc=(m == 2))
    assert np.array_equal(perm.cpu().numpy().view(np.uint32).astype(np.uint64), oracle.argsort_multi(keys, m))


def _bytes_col(n: int, width: int, seed: int, maxlen: int = 3) -> np.ndarray:
    """Padded byte strings of up to maxlen bytes from a small alphabet with
    signed-char extremes (a byte-string key column of `width` bytes)."""
    rng = np.random.default_rng(seed)
    alphabet = np.array([0x00, 0x41, 0x7F, 0x80, 0xFF], dtype=np.uint8)
    c = alphabet[rng.integers(0, alphabet.size, (n, width))]
    c[:, maxlen:] = 0
    return c


WIDE = {
    "char16+char16": lambda n, rng: [datagen.char16_keys(1, 0, n, distinct=6), datagen.char16_keys(2, 0, n, distinct=5)],
    "bytes23+i32": lambda n, rng: [_bytes_col(n, 23, 3), rng.integers(-4, 4, n).astype(np.int32)],
    "i64+char16+bytes5": lambda n, rng: [rng.integers(-2, 2, n).astype(np.int64),
                                         datagen.char16_keys(4, 0, n, distinct=3), _bytes_col(n, 5, 5, 2)],
    "bytes5+u32": lambda n, rng: [_bytes_col(n, 5, 6), rng.integers(0, 3, n).astype(np.uint32)],  # 9 bytes
    "bytes300": lambda n, rng: [_bytes_col(n, 300, 7)],
    "f64+bytes40+u64+char16": lambda n, rng: [np.where(rng.random(n) < 0.5, -0.0, 1.5),
                                              _bytes_col(n, 40, 8, 2),
                                              rng.integers(0, 2, n).astype(np.uint64) << np.uint64(40),
                                              datagen.char16_keys(9, 0, n, distinct=2)],
}


@pytest.mark.parametrize("combo", list(WIDE))
def test_groupby_multi_wide_keys_vs_oracle(vdf, combo):
    """Multi-column keys of any width (byte-string columns and composites
    over 16 bytes, sorted and grouped through the byte-string path): keys of
    every column, offsets, counts, permutation and aggregates against the
    oracle; the argsort in every order mode."""
    rng = np.random.default_rng(13)
    n = 150_007
    keys = WIDE[combo](n, rng)
    vals = [rng.standard_normal(n), rng.integers(-2**40, 2**40, n).astype(np.int64)]
    _say("groupby_multi wide", combo, [k.shape for k in keys])
    r = vdf.groupby_multi([_dev(k) for k in keys], [_dev(v) for v in vals])
    o = oracle.groupby_multi(keys, vals, threads=8)
    assert r.n_groups == o["n_groups"], combo
    for c, k in enumerate(keys):
        got = r.keys[c].cpu().numpy()
        assert np.array_equal(got.view(np.uint8).reshape(got.shape[0], -1),
                              o["keys"][c].view(np.uint8).reshape(got.shape[0], -1)), (combo, c)
    assert np.array_equal(r.offsets.cpu().numpy().view(np.uint32).astype(np.uint64), o["offsets"]), combo
    assert np.array_equal(r.counts.cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"]), combo
    assert np.array_equal(r.perm.cpu().numpy().view(np.uint32).astype(np.uint64), o["perm"]), combo
    ra, oa = r.aggs[1], o["aggs"][1]
    assert np.array_equal(ra["sum"].cpu().numpy().view(np.int64), oa["sum"].view(np.int64)), combo
    assert np.array_equal(ra["min"].cpu().numpy(), oa["min"]), combo
    got = r.aggs[0]["sum"].cpu().numpy()
    assert np.all(np.abs(got - o["aggs"][0]["sum"]) <= F64_RTOL * np.abs(o["aggs"][0]["sum"])), combo
    for m, kw in ((0, {}), (1, {"descending": True}), (2, {"reverse_desc": True})):
        perm = vdf.argsort_multi([_dev(k) for k in keys], **kw)
        assert np.array_equal(perm.cpu().numpy().view(np.uint32).astype(np.uint64),
                              oracle.argsort_multi(keys, m)), (combo, m)


def test_multi_key_too_wide_rejected(vdf):
    import torch

    k = torch.zeros(10, dtype=torch.int64, device="cuda")
    c = torch.zeros((10, 4090), dtype=torch.uint8, device="cuda")
    with pytest.raises(N.VdfError) as e:
        vdf.groupby_multi([k, c])
    assert e.value.status == N.VDF_ERR_UNSUPPORTED


def _agg_vs_oracle(r, o, vals, label):
    """Per-group outputs of a GPU result against the oracle's."""
    g = o["n_groups"]
    assert r.n_groups == g, label
    k = r.keys.cpu().numpy()
    assert np.array_equal(k.view(np.uint8).reshape(-1), o["keys"].view(np.uint8).reshape(-1)), label
    assert np.array_equal(r.offsets.cpu().numpy().view(np.uint32).astype(np.uint64), o["offsets"]), label
    assert np.array_equal(r.counts.cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"]), label
    assert np.array_equal(r.perm.cpu().numpy().view(np.uint32).astype(np.uint64), o["perm"]), label
    for c, v in enumerate(vals):
        ra, oa = r.aggs[c], o["aggs"][c]
        assert np.array_equal(_bytes(ra["min"]), oa["min"].view(np.uint8)), (label, c)
        assert np.array_equal(_bytes(ra["max"]), oa["max"].view(np.uint8)), (label, c)
        if v.dtype == np.float64:
            for kk in ("sum", "mean"):
                got = ra[kk].cpu().numpy()
                assert np.all(np.abs(got - oa[kk]) <= F64_RTOL * np.abs(oa[kk])), (label, c, kk)
        else:
            assert np.array_equal(ra["sum"].cpu().numpy().view(np.int64), oa["sum"].view(np.int64)), (label, c)
            assert np.array_equal(ra["mean"].cpu().numpy(), oa["mean"]), (label, c)


@pytest.mark.parametrize("src", ["f64_sum", "i64_sum", "count", "i32_min"])
def test_broadcast_per_row(vdf, src):
    """vdf_broadcast writes each row its group's value: against the oracle's
    group-by broadcast to rows by key (exact for counts, integer sums and
    min; float64 sums within 1e-12 of the oracle's exact sums), and
    bit-identical to the GPU group value it was given."""
    rng = np.random.default_rng(12)
    n = 500_001
    k = rng.integers(0, 7000, n).astype(np.int64)
    vf = rng.standard_normal(n)
    vi = rng.integers(-1000, 1000, n).astype(np.int32)
    r = vdf.groupby(_dev(k), [_dev(vf), _dev(vi)], trim=False)
    o = oracle.groupby(k, [vf, vi], threads=8)
    if src == "f64_sum":
        gv, to64, ov = r.aggs[0]["sum"], False, o["aggs"][0]["sum"]
    elif src == "i64_sum":
        gv, to64, ov = r.aggs[1]["sum"], True, o["aggs"][1]["sum"].view(np.int64)
    elif src == "count":
        gv, to64, ov = r.counts, True, o["counts"]  # counts (int32 cells) as a float64 column
    else:
        gv, to64, ov = r.aggs[1]["min"], False, o["aggs"][1]["min"]
    out = vdf.broadcast(gv, r, n, to_f64=to64).cpu().numpy()
    g = int(r.n_groups.item())
    assert g == o["n_groups"]
    row_group = np.searchsorted(o["keys"], k)  # the oracle's group of every row
    exp = ov[row_group]
    if to64:
        exp = exp.astype(np.float64)
    if src == "f64_sum":
        assert np.all(np.abs(out - exp) <= F64_RTOL * np.abs(exp)), src
    else:
        assert np.array_equal(out.view(np.uint8), exp.view(np.uint8)), src
    gval = gv[:g].cpu().numpy()
    same = gval[row_group].astype(np.float64) if to64 else gval[row_group]
    assert np.array_equal(out.view(np.uint8), same.view(np.uint8)), src


@pytest.mark.parametrize("kind", ["i32", "u64", "char16", "f64"])
def test_cached_grouping_vs_oracle(vdf, kind):
    """Aggregating against a kept grouping, for several value columns in
    turn (no re-sort between them): every output against the oracle's
    group-by of the same columns, and bit-identical to a fresh group-by."""
    rng = np.random.default_rng(17)
    n = 400_003
    keys = _key_cols((kind,), n, rng)[0]
    kt = oracle.key_type_of(keys)
    dk = _dev(keys)
    g = vdf.grouping(dk, key_type=kt)
    for c, vals in enumerate([[rng.standard_normal(n)], [rng.integers(-2**40, 2**40, n).astype(np.int64)],
                              [rng.integers(0, 1000, n).astype(np.int32), rng.standard_normal(n)], []]):
        dv = [_dev(v) for v in vals]
        got = g.aggregate(dv)
        _agg_vs_oracle(got, oracle.groupby(keys, vals, threads=8), vals, f"{kind}[{c}]")
        exp = vdf.groupby(dk, dv, key_type=kt)
        assert got.n_groups == exp.n_groups, (kind, c)
        for x, y in ((got.keys, exp.keys), (got.offsets, exp.offsets), (got.counts, exp.counts),
                     (got.perm, exp.perm)):
            assert np.array_equal(_bytes(x), _bytes(y)), (kind, c)
        for da, db in zip(got.aggs, exp.aggs):
            for k in da:
                assert np.array_equal(_bytes(da[k]), _bytes(db[k])), (kind, c, k)
    g.close()


# ------------------------------------------------------- fixed-width strings
def _strings(n: int, width: int, seed: int) -> np.ndarray:
    """n padded byte strings of up to `width` bytes: lengths vary (padding
    0x00, as the reference's PaddingChar), bytes from a small alphabet with
    signed-char extremes, so ties and shared prefixes are common."""
    rng = np.random.default_rng(seed)
    alphabet = np.array([0x00, 0x01, 0x41, 0x42, 0x7F, 0x80, 0xC3, 0xFF], dtype=np.uint8)
    keys = alphabet[rng.integers(0, alphabet.size, (n, width))]
    lens = rng.integers(0, width + 1, n)
    keys[np.arange(width)[None, :] >= lens[:, None]] = 0
    return keys


@pytest.mark.parametrize("width", [1, 3, 16, 17, 32, 33, 70, 255])
@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_argsort_bytes_vs_oracle(vdf, width, mode):
    import torch

    m = {"asc": 0, "desc": 1, "reverse": 2}[mode]
    for n in (1, 2, 4097, 100_003):
        keys = _strings(n, width, width * 7 + n)
        _say("argsort_bytes", width, mode, n)
        exp = oracle.argsort_bytes(keys, m).astype(np.int64)
        perm, rows = vdf.argsort_bytes(torch.from_numpy(keys).cuda(), descending=(mode == "desc"),
                                       reverse_desc=(mode == "reverse"), want_sorted=True)
        assert np.array_equal(perm.cpu().numpy().view(np.uint32).astype(np.int64), exp), (width, mode, n)
        assert np.array_equal(rows.cpu().numpy(), keys[exp]), (width, mode, n)


def test_argsort_bytes_width16_equals_char16_and_perm64(vdf):
    import torch

    keys = datagen.char16_keys(5, 0, 200_001, distinct=5000)
    d = torch.from_numpy(keys).cuda()
    a = vdf.argsort_bytes(d, perm64=True)
    b = vdf.argsort(d, key_type=N.KEY_CHAR16, perm64=True)
    assert torch.equal(a, b)
    assert np.array_equal(a.cpu().numpy().astype(np.uint64), oracle.argsort(keys, 0))


def test_argsort_bytes_rejects_bad_width_and_empty_is_fine(vdf):
    import torch

    with pytest.raises(N.VdfError) as e:
        vdf.argsort_bytes(torch.zeros((4, 4097), dtype=torch.uint8, device="cuda"))
    assert e.value.status == N.VDF_ERR_INVALID
    assert vdf.argsort_bytes(torch.zeros((0, 8), dtype=torch.uint8, device="cuda")).shape[0] == 0


def _groupby_bytes_oracle(keys: np.ndarray, vals):
    """The oracle's group-by over each string's rank among the distinct
    strings (ascending signed-char byte order, i.e. bytes ^ 0x80 compared as
    unsigned): the same groups, order, row view and aggregates as grouping by
    the strings themselves."""
    n, w = keys.shape
    t = np.ascontiguousarray(keys ^ np.uint8(0x80)).view(np.dtype((np.void, w))).ravel()
    _, ids = np.unique(t, return_inverse=True)
    o = oracle.groupby(ids.reshape(-1).astype(np.uint32), vals)
    g = o["n_groups"]
    o["skeys"] = keys[o["perm"][o["offsets"][:g].astype(np.int64)].astype(np.int64)]
    return o


@pytest.mark.parametrize("width", [1, 7, 16, 40, 100])
def test_groupby_bytes_vs_oracle(vdf, width):
    import torch

    rng = np.random.default_rng(width)
    for n, distinct in ((1, 1), (4097, 50), (300_007, 20_000)):
        pool = _strings(distinct, width, width + distinct)
        keys = pool[rng.integers(0, distinct, n)]
        vf = rng.standard_normal(n)
        vi = rng.integers(-(2**40), 2**40, n)
        _say("groupby_bytes", width, n, distinct)
        r = vdf.groupby_bytes(torch.from_numpy(keys).cuda(), [torch.from_numpy(vf).cuda(), torch.from_numpy(vi).cuda()])
        o = _groupby_bytes_oracle(keys, [vf, vi])
        g = o["n_groups"]
        assert r.n_groups == g
        assert np.array_equal(r.keys.cpu().numpy(), o["skeys"]), (width, n)
        assert np.array_equal(r.offsets.cpu().numpy().view(np.uint32).astype(np.uint64), o["offsets"])
        assert np.array_equal(r.counts.cpu().numpy().view(np.uint32).astype(np.uint64), o["counts"])
        assert np.array_equal(r.perm.cpu().numpy().view(np.uint32).astype(np.uint64), o["perm"])
        fa, ia = r.aggs[0], r.aggs[1]
        of, oi = o["aggs"][0], o["aggs"][1]
        assert np.array_equal(ia["sum"].cpu().numpy(), oi["sum"].view(np.int64))
        assert np.array_equal(ia["min"].cpu().numpy(), oi["min"]) and np.array_equal(ia["max"].cpu().numpy(), oi["max"])
        assert np.array_equal(fa["min"].cpu().numpy().view(np.uint8), of["min"].view(np.uint8))
        assert np.array_equal(fa["max"].cpu().numpy().view(np.uint8), of["max"].view(np.uint8))
        for k in ("sum", "mean"):
            got, exp = fa[k].cpu().numpy(), of[k]
            assert np.all(np.abs(got - exp) <= F64_RTOL * np.abs(exp)), (width, n, k)


def test_groupby_bytes_empty_and_one_group(vdf):
    import torch

    r = vdf.groupby_bytes(torch.zeros((0, 5), dtype=torch.uint8, device="cuda"), [])
    assert r.n_groups == 0
    keys = np.full((10_000, 33), 0x41, dtype=np.uint8)
    r = vdf.groupby_bytes(torch.from_numpy(keys).cuda(), [], aggs=N.AGG_COUNT)
    assert r.n_groups == 1 and int(r.counts[0]) == 10_000
    assert np.array_equal(r.keys.cpu().numpy(), keys[:1])
