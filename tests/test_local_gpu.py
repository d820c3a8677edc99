"""GPU parity of the local finish of 4-byte-key sorts: two global radix passes
over the top two non-constant digits, then each bucket of equal top digits
sorted in LDS.  Every case is checked bit-exact against the oracle's stable
argsort (the reference's radix_sort_uint32 order, radix_sort_uint32.inl) and
against the same call with the local finish switched off; the profile shows
which path ran (the local finish, or its fallback to the plain passes)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from tests.test_radix_gpu import _dev, _key_type, _perm_host
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu

MODES = {"asc": 0, "desc": 1, "reverse": 2}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _u32(seed: int, n: int, bits: int = 32) -> np.ndarray:
    return (datagen.splitmix64(seed, 0, n) >> np.uint64(64 - bits)).astype(np.uint32)


def _vals(n: int, width: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if width == 8:
        return rng.standard_normal(n)
    return rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32)


def _run(vdf, local: int, fn):
    old = vdf.tune_get("sort_local")
    vdf.profile(True)
    vdf.profile_reset()
    try:
        vdf.tune("sort_local", local)
        out = fn()
        _, launches, _ = vdf.profile_read("radix_local")
        _, passes, _ = vdf.profile_read("radix_pass")
    finally:
        vdf.tune("sort_local", old)
        vdf.profile(False)
        vdf.profile_reset()
    return out, launches, passes


def _sort_by(vdf, keys, vals, mode, perm64=False, want_perm=True):
    return vdf.sort_by(_dev(keys), _dev(vals), key_type=_key_type(keys), descending=(mode == "desc"),
                       reverse_desc=(mode == "reverse"), perm64=perm64, want_perm=want_perm, want_keys=True)


def _check_sort_by(vdf, keys, width, mode="asc", perm64=False, want_perm=True, expect_local=True):
    vals = _vals(keys.shape[0], width, keys.shape[0] + width)
    exp = oracle.argsort(keys, MODES[mode], 4).astype(np.int64)
    (perm, out, skeys), launches, passes = _run(vdf, 1, lambda: _sort_by(vdf, keys, vals, mode, perm64, want_perm))
    assert launches == (1 if expect_local else 0), f"local finish ran {launches} times"
    if expect_local:
        assert passes == 2
    if want_perm:
        assert np.array_equal(_perm_host(perm).astype(np.int64), exp)
    assert np.array_equal(out.cpu().numpy().view(np.uint8), vals[exp].view(np.uint8))
    assert np.array_equal(skeys.cpu().numpy().view(np.uint8), keys[exp].view(np.uint8))
    # the same call through the plain passes
    (perm0, out0, skeys0), launches0, _ = _run(vdf, 0, lambda: _sort_by(vdf, keys, vals, mode, perm64, want_perm))
    assert launches0 == 0
    import torch

    if want_perm:
        assert torch.equal(perm0, perm)
    assert torch.equal(out0.view(torch.uint8), out.view(torch.uint8))
    assert torch.equal(skeys0.view(torch.uint8), skeys.view(torch.uint8))


@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
@pytest.mark.parametrize("width", [4, 8])
def test_local_sort_by_uniform(vdf, mode, width):
    for n in (70_001, 1_000_003, 3_000_017):
        _say("local sort_by uniform", mode, width, n)
        _check_sort_by(vdf, _u32(n, n), width, mode)


def test_local_sort_by_key_types(vdf):
    n = 1_500_007
    rng = np.random.default_rng(5)
    i32 = rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32)
    _check_sort_by(vdf, i32, 8)
    f32 = (rng.standard_normal(n) * 1e4).astype(np.float32)
    sp = np.array([np.nan, np.copysign(np.nan, -1), np.inf, -np.inf, 0.0, -0.0, 1e-40, -1e-40], dtype=np.float32)
    f32[rng.integers(0, n, 5000)] = sp[rng.integers(0, sp.shape[0], 5000)]
    _check_sort_by(vdf, f32, 8)
    _check_sort_by(vdf, f32, 4, "reverse")


def test_local_sort_by_perm64_and_without_perm(vdf):
    keys = _u32(31, 2_000_003)
    _check_sort_by(vdf, keys, 8, perm64=True)
    _check_sort_by(vdf, keys, 4, perm64=True)
    _check_sort_by(vdf, keys, 8, want_perm=False)


@pytest.mark.parametrize("bits", [25, 28, 31])
def test_local_narrower_keys(vdf, bits):
    """Keys whose top digit holds only a few bits (buckets of equal top
    digits cover several values of it)."""
    n = 1_200_011
    _say("local bits", bits)
    _check_sort_by(vdf, _u32(bits, n, bits), 8)


@pytest.mark.parametrize("bits", [17, 24])
def test_three_digit_keys_stay_plain(vdf, bits):
    """With three non-constant digits the local finish would replace one pass
    at the cost of more than one: the plain passes run."""
    n = 1_200_011
    keys = _u32(bits, n, bits)
    (perm, out, sk), launches, passes = _run(vdf, 1, lambda: _sort_by(vdf, keys, _vals(n, 8, 2), "asc"))
    assert launches == 0 and passes == 3
    exp = oracle.argsort(keys, 0, 4).astype(np.int64)
    assert np.array_equal(_perm_host(perm).astype(np.int64), exp)


def test_local_argsort(vdf):
    import torch

    for n in (65_536, 999_999, 2_500_001):
        keys = _u32(n + 1, n)
        exp = oracle.argsort(keys, 0, 4).astype(np.uint32)
        (perm, skeys), launches, _ = _run(vdf, 1, lambda: vdf.argsort(_dev(keys), key_type=N.KEY_U32, want_keys=True))
        assert launches == 1
        assert np.array_equal(_perm_host(perm), exp)
        assert np.array_equal(skeys.cpu().numpy().view(np.uint32), keys[exp.astype(np.int64)])
        (perm0, _), _, _ = _run(vdf, 0, lambda: vdf.argsort(_dev(keys), key_type=N.KEY_U32, want_keys=True))
        assert torch.equal(perm0, perm)


def _equal_buckets(n_per: int, seed: int) -> np.ndarray:
    """256 buckets of exactly n_per rows: digits 2 and 3 equal (so each digit
    alone looks uniform), the low 16 bits random."""
    n = 256 * n_per
    top = (np.arange(n, dtype=np.uint64) % 256).astype(np.uint32)
    rng = np.random.default_rng(seed)
    rng.shuffle(top)
    low = _u32(seed, n, 16)
    return (top << np.uint32(24)) | (top << np.uint32(16)) | low


def test_local_bucket_size_limit(vdf):
    """Stride rule (tables under 2^24 rows): buckets of exactly 3840 rows
    (4096 - 256, the largest it takes) finish locally; 3841-row buckets fall
    back to the plain passes, which then read the caller's columns again and
    must give the same result."""
    _say("local limit 3840")
    _check_sort_by(vdf, _equal_buckets(3840, 1), 8)
    _say("local fallback 3841")
    _check_sort_by(vdf, _equal_buckets(3841, 2), 8, expect_local=False)
    _check_sort_by(vdf, _equal_buckets(3841, 3), 4, "reverse", expect_local=False)


def _pair_buckets(n_per: int, seed: int) -> np.ndarray:
    """4096 buckets of exactly n_per rows (at least 2^24 rows in all, so the
    chunks are packed greedily); every value of digits 2 and 3 occurs in 16
    of them, so each digit alone looks uniform."""
    i = np.arange(4096, dtype=np.uint32)
    y = i % 256
    x = i // 256 + 16 * (i % 16)
    top = np.repeat((y << np.uint32(8)) | x, n_per)
    np.random.default_rng(seed).shuffle(top)
    return (top << np.uint32(16)) | _u32(seed, top.shape[0], 16)


def test_local_greedy_bucket_size_limit(vdf):
    """Greedy packing (tables of 2^24 rows or more): buckets of exactly 4096
    rows, a whole chunk each, finish locally; 4097-row buckets fall back."""
    _say("local greedy limit 4096")
    _check_sort_by(vdf, _pair_buckets(4096, 5), 8)
    _say("local greedy fallback 4097")
    _check_sort_by(vdf, _pair_buckets(4097, 6), 4, "desc", expect_local=False)


@pytest.mark.parametrize("mode", ["asc", "reverse"])
def test_local_greedy_uniform(vdf, mode):
    """20M uniform keys: chunks packed greedily hold about 13 buckets each."""
    _say("local greedy uniform", mode)
    _check_sort_by(vdf, _u32(20, 20_000_003), 8, mode)


def _two_bucket_keys(n: int, seed: int) -> np.ndarray:
    """Digit 3 in [0, 48): 12,288 buckets of about 1,630 rows, so greedy
    chunks hold two adjacent buckets and finish with the 9-bit pass."""
    top = (datagen.splitmix64(seed, 0, n) % np.uint64(48 * 256)).astype(np.uint32)
    return (top << np.uint32(16)) | _u32(seed + 1, n, 16)


@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_local_two_bucket_chunks(vdf, mode):
    _say("local two-bucket chunks", mode)
    keys = _two_bucket_keys(20_000_011, 7)
    if mode != "asc":
        keys = keys ^ np.uint32(0xFFFF0000)  # the same buckets at the top of the range
    _check_sort_by(vdf, keys, 8, mode)
    _check_sort_by(vdf, keys.view(np.int32), 4, mode)


def test_local_two_bucket_chunks_perm64_and_without_perm(vdf):
    """The greedy rule and the 9-bit pass with 64-bit row ids and with the
    row-id-free kernel (no permutation asked)."""
    keys = _two_bucket_keys(20_000_011, 8)
    _check_sort_by(vdf, keys, 8, perm64=True)
    _check_sort_by(vdf, keys, 4, "reverse", want_perm=False)


def test_local_skewed_keys_stay_plain(vdf):
    """A heavy key makes the digit histograms rule the local finish out
    before any pass runs (no wasted passes)."""
    n = 1_000_003
    keys = _u32(77, n)
    keys[::3] = 123456789
    (perm, out, sk), launches, passes = _run(vdf, 1, lambda: _sort_by(vdf, keys, _vals(n, 8, 1), "asc"))
    assert launches == 0 and passes == 4
    exp = oracle.argsort(keys, 0, 4).astype(np.int64)
    assert np.array_equal(_perm_host(perm).astype(np.int64), exp)


@pytest.mark.parametrize("combo", ["f64+i32", "f64+f64", "i32", "f64"])
def test_local_groupby_full_range(vdf, combo):
    """Group-bys over full-range int32 keys sort with the local finish (the
    carried columns ride through it); results equal the plain path's and the
    oracle's."""
    import torch

    n = 1_000_003
    keys = _u32(9, n).view(np.int32)
    rng = np.random.default_rng(4)
    cols = []
    for c in combo.split("+"):
        cols.append(rng.standard_normal(n) if c == "f64" else rng.integers(-10**6, 10**6, n).astype(np.int32))
    d_cols = [_dev(c) for c in cols]
    for want_perm in (False, True):
        r1, launches, _ = _run(vdf, 1, lambda: vdf.groupby(_dev(keys), d_cols, want_perm=want_perm))
        assert launches == 1
        r0, _, _ = _run(vdf, 0, lambda: vdf.groupby(_dev(keys), d_cols, want_perm=want_perm))
        assert r1.n_groups == r0.n_groups
        assert torch.equal(r1.keys, r0.keys) and torch.equal(r1.counts, r0.counts)
        if want_perm:
            assert torch.equal(r1.perm, r0.perm)
        for a1, a0 in zip(r1.aggs, r0.aggs):
            for k in a1:
                if a1[k] is not None:
                    assert torch.equal(a1[k].view(torch.uint8), a0[k].view(torch.uint8)), k
    o = oracle.groupby(keys, cols)
    assert r1.n_groups == o["n_groups"]
    assert np.array_equal(r1.keys.cpu().numpy(), o["keys"])
    assert np.array_equal(r1.counts.cpu().numpy().astype(np.uint64), o["counts"])


@pytest.mark.parametrize("case", ["sort_by_local", "sort_by_plain", "sort_by_two_bucket", "argsort_u64",
                                  "groupby_full", "zipf"])
def test_ballot_and_atomic_ranking_identical(vdf, case):
    """The per-row LDS atomic ranking (used where the device serves same-address
    lanes in lane order) and the wave-ballot ranking give bit-identical
    results on every path that ranks: plain passes, the local finish, 8-byte
    keys, group-bys."""
    import torch

    if vdf.tune_get("sort_rank_atomic") != 1:
        pytest.skip("the device check did not verify the atomic ranking: both runs would use ballots")
    n = 1_000_003
    rng = np.random.default_rng(17)

    def run():
        if case == "sort_by_two_bucket":
            keys = _two_bucket_keys(20_000_011, 3)
            out, launches, _ = _run(vdf, 1, lambda: _sort_by(vdf, keys, _vals(keys.shape[0], 8, 3), "asc"))
            assert launches == 1
            return out
        if case.startswith("sort_by"):
            keys = _u32(5, n)
            vals = _vals(n, 8, 5)
            return _run(vdf, 1 if case == "sort_by_local" else 0,
                        lambda: _sort_by(vdf, keys, vals, "reverse"))[0]
        if case == "argsort_u64":
            keys = datagen.splitmix64(6, 0, n)
            return vdf.argsort(_dev(keys), key_type=N.KEY_U64, want_keys=True)
        if case == "groupby_full":
            keys = _u32(9, n).view(np.int32)
            r = vdf.groupby(_dev(keys), [_dev(rng.standard_normal(n))])
            return (r.keys, r.counts, r.perm, r.aggs[0]["sum"], r.aggs[0]["min"])
        keys = datagen.scramble24(datagen.zipf_ranks(2, 0, n, 1.1, 1 << 24)).astype(np.int32)
        r = vdf.groupby(_dev(keys), [_dev(rng.integers(-10**6, 10**6, n).astype(np.int32))])
        return (r.keys, r.counts, r.perm, r.aggs[0]["sum"], r.aggs[0]["max"])

    old = vdf.tune_get("sort_rank")
    try:
        vdf.tune("sort_rank", 0)
        assert vdf.tune_get("sort_rank_atomic") == 0
        ballot = run()
        rng = np.random.default_rng(17)
        vdf.tune("sort_rank", 1)
        atomic = run()
    finally:
        vdf.tune("sort_rank", old)
    for b, a in zip(ballot, atomic):
        if b is None:
            assert a is None
            continue
        assert torch.equal(b.view(torch.uint8), a.view(torch.uint8))
