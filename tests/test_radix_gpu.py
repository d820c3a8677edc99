"""GPU parity of the radix argsort and gather against the oracle and the
reference's golden vectors (bit-exact permutations)."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle
from tests.golden import cases
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "radix_golden.json")
KT = {"u32": N.KEY_U32, "i32": N.KEY_I32, "u64": N.KEY_U64, "i64": N.KEY_I64, "char16": N.KEY_CHAR16}


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


def _key_type(a: np.ndarray) -> int:
    return oracle.key_type_of(a)


def _perm_host(t) -> np.ndarray:
    p = t.cpu().numpy()
    return p.view(np.uint32) if p.dtype == np.int32 else p.view(np.uint64)


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden(), ids=lambda c: c["name"])
def test_argsort_reference_golden(vdf, case):
    keys = cases.make_keys(case)
    n = keys.shape[0]
    _say("golden", case["name"], n)
    if n == 0:
        return
    perm = _perm_host(vdf.argsort(_dev(keys), key_type=KT[case["kind"]]))
    assert hashlib.sha256(perm.astype(np.uint32).tobytes()).hexdigest() == case["perm_sha256"]
    if "perm" in case:
        assert perm.tolist() == case["perm"]


def _key_sets():
    rng = np.random.default_rng(21)
    yield "u32_uniform", (datagen.splitmix64(1, 0, 300_001) >> np.uint64(32)).astype(np.uint32)
    yield "u32_two_values", rng.integers(0, 2, 100_000).astype(np.uint32) * np.uint32(0x01000001)
    yield "i32_small", rng.integers(-50, 50, 123_457).astype(np.int32)
    yield "i32_zipf24", datagen.scramble24(datagen.zipf_ranks(2, 0, 200_000, 1.1, 1 << 24)).astype(np.int32)
    yield "u64_uniform", datagen.splitmix64(3, 0, 150_001)
    yield "i64_c1", datagen.groupby_small(200_000, seed=1)[0]
    yield "i64_equal", np.full(77_777, -5, dtype=np.int64)
    yield "char16", datagen.char16_keys(4, 0, 60_001)
    yield "char16_few", datagen.char16_keys(4, 0, 60_001, distinct=9)
    for dt in (np.float32, np.float64):
        f = np.round(rng.standard_normal(100_003) * 20).astype(dt)  # ties, +0.0
        sp = np.array([np.nan, np.copysign(np.nan, -1), np.inf, -np.inf, 0.0, -0.0, 1e-40, -1e-40], dtype=dt)
        f[rng.integers(0, f.shape[0], 4000)] = sp[rng.integers(0, sp.shape[0], 4000)]
        yield f"{np.dtype(dt).name}_specials", f


@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_argsort_modes_vs_oracle(vdf, mode):
    for name, keys in _key_sets():
        _say("modes", mode, name, keys.shape)
        m = {"asc": 0, "desc": 1, "reverse": 2}[mode]
        exp = oracle.argsort(keys, m, 4)
        perm, skeys = vdf.argsort(_dev(keys), key_type=_key_type(keys), descending=(mode == "desc"),
                                  reverse_desc=(mode == "reverse"), want_keys=True)
        got = _perm_host(perm)
        assert np.array_equal(got, exp.astype(np.uint32)), name
        sk = skeys.cpu().numpy()  # bytes compared: NaN keys come back bit-exact
        assert np.array_equal(sk.view(np.uint8), keys[exp.astype(np.int64)].view(np.uint8)), name


@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_argsort_wide_keys_split_tiles(vdf, mode):
    """8- and 16-byte keys with 16 rows per thread: 16-byte keys use split
    staging (a key word per phase, then the row ids); bit-exact permutation and
    sorted keys against the oracle."""
    m = {"asc": 0, "desc": 1, "reverse": 2}[mode]
    old = (vdf.tune_get("sort_items_64"), vdf.tune_get("sort_items_128"))
    try:
        vdf.tune("sort_items_64", 16)
        vdf.tune("sort_items_128", 16)
        for name, keys in _key_sets():
            if keys.dtype.itemsize == 4 and keys.ndim == 1:
                continue
            _say("wide split", mode, name, keys.shape)
            exp = oracle.argsort(keys, m, 4)
            perm, skeys = vdf.argsort(_dev(keys), key_type=_key_type(keys), descending=(mode == "desc"),
                                      reverse_desc=(mode == "reverse"), want_keys=True)
            assert np.array_equal(_perm_host(perm), exp.astype(np.uint32)), name
            assert np.array_equal(skeys.cpu().numpy().view(np.uint8), keys[exp.astype(np.int64)].view(np.uint8)), name
    finally:
        vdf.tune("sort_items_64", old[0])
        vdf.tune("sort_items_128", old[1])


def test_argsort_perm64(vdf):
    keys = datagen.groupby_small(50_001, seed=5)[0]
    perm = vdf.argsort(_dev(keys), perm64=True)
    assert np.array_equal(_perm_host(perm), oracle.argsort(keys, 0))


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097, 65535, 1_000_013])
def test_argsort_tile_edges(vdf, n):
    _say("edges", n)
    keys = (datagen.splitmix64(n, 0, n) >> np.uint64(40)).astype(np.uint32)  # 24-bit: 3 passes
    perm = _perm_host(vdf.argsort(_dev(keys), key_type=N.KEY_U32))
    assert np.array_equal(perm, oracle.argsort(keys, 0).astype(np.uint32))


def test_internals_doc_example_on_gpu(vdf):
    col3 = np.array([2, 7, 2, 7, 2, 7, 2, 7, 7, 2, 7, 7, 7, 2, 7], dtype=np.uint32)
    asc = _perm_host(vdf.argsort(_dev(col3), key_type=N.KEY_U32)) + 1
    desc = _perm_host(vdf.argsort(_dev(col3), key_type=N.KEY_U32, descending=True)) + 1
    assert asc.tolist() == [1, 3, 5, 7, 10, 14, 2, 4, 6, 8, 9, 11, 12, 13, 15]
    assert desc.tolist() == [2, 4, 6, 8, 9, 11, 12, 13, 15, 1, 3, 5, 7, 10, 14]


def test_gather_vs_oracle(vdf):
    import torch

    rng = np.random.default_rng(4)
    n = 1_000_003
    perm = rng.permutation(n).astype(np.uint32)
    for src in (rng.standard_normal(n), rng.integers(0, 2**31, n).astype(np.int32), datagen.char16_keys(6, 0, n)):
        _say("gather", src.dtype, src.shape)
        got = vdf.gather(_dev(src), torch.from_numpy(perm.view(np.int32)).cuda()).cpu().numpy()
        assert np.array_equal(got.view(src.dtype) if src.ndim == 1 else got, oracle.gather(src, perm.astype(np.uint64)))
    p64 = torch.from_numpy(perm.astype(np.int64)).cuda()
    src = rng.standard_normal(n)
    assert np.array_equal(vdf.gather(_dev(src), p64).cpu().numpy(), src[perm.astype(np.int64)])


@pytest.mark.parametrize("width", [1, 3, 12, 20, 33, 256])
def test_gather_any_width_vs_oracle(vdf, width):
    """Cells of widths without a vectorised kernel (padded strings): row copies
    in words or bytes, 32- and 64-bit row ids, and an unaligned source view."""
    import torch

    rng = np.random.default_rng(width)
    n = 20_011
    src = rng.integers(0, 256, (n, width), dtype=np.uint8)
    perm = rng.permutation(n)
    exp = oracle.gather(src, perm.astype(np.uint64))
    d = torch.from_numpy(src).cuda()
    for p in (torch.from_numpy(perm.astype(np.int32)).cuda(), torch.from_numpy(perm.astype(np.int64)).cuda()):
        assert np.array_equal(vdf.gather(d, p).cpu().numpy(), exp), width
    # the same rows starting one byte into a larger buffer
    big = torch.zeros(n * width + 1, dtype=torch.uint8, device="cuda")
    big[1:] = d.reshape(-1)
    view = big[1:].view(n, width)
    assert np.array_equal(vdf.gather(view, torch.from_numpy(perm.astype(np.int32)).cuda()).cpu().numpy(), exp)


def test_reuse_handle_and_side_stream(vdf):
    import torch

    big = (datagen.splitmix64(8, 0, 2_000_000) >> np.uint64(32)).astype(np.uint32)
    small = big[:1000].copy()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        pb = vdf.argsort(_dev(big), key_type=N.KEY_U32, stream=s)
    s.synchronize()
    ps = vdf.argsort(_dev(small), key_type=N.KEY_U32)
    assert np.array_equal(_perm_host(pb), oracle.argsort(big, 0, 4).astype(np.uint32))
    assert np.array_equal(_perm_host(ps), oracle.argsort(small, 0).astype(np.uint32))


def test_profile_counts_active_passes(vdf):
    keys = (datagen.splitmix64(9, 0, 1_000_000) >> np.uint64(32)).astype(np.uint32)
    k24 = keys >> np.uint32(8)
    d_keys, d_k24 = _dev(keys), _dev(k24)
    old_local = vdf.tune_get("sort_local")
    vdf.tune("sort_local", 0)  # the plain passes (the local finish: tests/test_local_gpu.py)
    vdf.profile(True)
    vdf.profile_reset()
    vdf.argsort(d_keys, key_type=N.KEY_U32)
    ms, launches, nbytes = vdf.profile_read("radix_pass")
    assert launches == 4 and ms > 0
    n = 1_000_000
    assert nbytes == n * (4 + 4 + 4) + 2 * n * 16 + n * (4 + 4 + 4)
    vdf.profile_reset()
    vdf.argsort(d_k24, key_type=N.KEY_U32)  # top digit constant: 3 passes
    _, launches, _ = vdf.profile_read("radix_pass")
    assert launches == 3
    # a batch: each pass launch is charged the tables whose plan has that pass
    small = (keys[:300_000] & np.uint32(0xFFFF))  # 2 passes
    vdf.profile_reset()
    vdf.argsort_batch([_dev(small), d_keys], key_type=N.KEY_U32)
    _, launches, nbytes = vdf.profile_read("radix_pass")
    assert launches == 4
    assert nbytes == 300_000 * (12 + 12) + n * (12 + 16 + 16 + 12)
    # local finish: two passes, then the LDS finish reads (key, row id) records
    # and writes the permutation
    vdf.tune("sort_local", 1)
    vdf.profile_reset()
    vdf.argsort(d_keys, key_type=N.KEY_U32)
    _, launches, nbytes = vdf.profile_read("radix_pass")
    assert launches == 2 and nbytes == n * (12 + 16)
    _, launches, nbytes = vdf.profile_read("radix_local")
    assert launches == 1 and nbytes == n * (8 + 4)
    vdf.tune("sort_local", old_local)
    vdf.profile(False)
    vdf.profile_reset()


def test_error_paths(vdf):
    import ctypes as C

    lib = N.load()
    rc = lib.vdf_argsort(vdf.h, N.KEY_U32, C.c_void_p(256), N.MAX_ROWS + 1, 0, C.c_void_p(256), None, None)
    assert rc == N.VDF_ERR_INVALID and b"VDF_MAX_ROWS" in lib.vdf_last_error()
    rc = lib.vdf_argsort(vdf.h, N.KEY_U32, C.c_void_p(256), 10, N.SORT_DESC | N.SORT_DESC_REVERSE,
                         C.c_void_p(256), None, None)
    assert rc == N.VDF_ERR_INVALID
    from visondf_amd.device import Vdf

    tiny = Vdf(0, hbm_budget_bytes=1 << 16)
    keys = _dev(np.arange(100_000, dtype=np.uint32))
    with pytest.raises(N.VdfError) as e:
        tiny.argsort(keys, key_type=N.KEY_U32)
    assert e.value.status == N.VDF_ERR_BUDGET
    tiny.close()


def test_argsort_100m_u32_bitexact(vdf):
    """Full-size uniform uint32 argsort (100M rows) + f64 gather, bit-exact vs the oracle."""
    import torch

    n = 100_000_000
    _say("100M: generating")
    keys, val = datagen.argsort_u32(n)
    _say("100M: gpu sort")
    d_keys = _dev(keys)
    perm, skeys = vdf.argsort(d_keys, key_type=N.KEY_U32, want_keys=True)
    out = vdf.gather(_dev(val), perm)
    torch.cuda.synchronize()
    # size-independent properties on the device
    sk = skeys.to(torch.int64) & 0xFFFFFFFF
    assert bool((sk[1:] >= sk[:-1]).all())
    p = perm.to(torch.int64)
    eq = sk[1:] == sk[:-1]
    assert bool((p[1:][eq] > p[:-1][eq]).all())  # stable
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_fill_(0, p, 1)
    assert int(seen.sum()) == n
    _say("100M: oracle")
    exp = oracle.argsort(keys, 0, min(os.cpu_count() or 8, 32))
    got = _perm_host(perm)
    assert np.array_equal(got, exp.astype(np.uint32))
    assert np.array_equal(out.cpu().numpy(), val[exp.astype(np.int64)])
