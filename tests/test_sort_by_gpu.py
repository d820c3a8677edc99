"""GPU parity of vdf_sort_by, the sort that carries one value column through
its radix passes, against the oracle's argsort + gather (bit-exact: the
permutation, the sorted keys and the value column in sorted order)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from tests.test_radix_gpu import _dev, _key_sets, _key_type, _perm_host
from visondf_amd import _native as N
from visondf_amd import datagen

pytestmark = pytest.mark.gpu

MODES = {"asc": 0, "desc": 1, "reverse": 2}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _vals(n: int, width: int, seed: int) -> np.ndarray:
    """8 bytes: float64 with NaN payloads, signed zeros and infinities (moved as
    bits); 4 bytes: int32 over the full range."""
    rng = np.random.default_rng(seed)
    if width == 8:
        v = rng.standard_normal(n)
        sp = np.array([np.nan, -np.nan, np.inf, -np.inf, 0.0, -0.0, 5e-324], dtype=np.float64)
        v[rng.integers(0, n, max(n // 50, 1))] = sp[rng.integers(0, sp.shape[0], max(n // 50, 1))]
        return v
    return rng.integers(-(2**31), 2**31, n, dtype=np.int64).astype(np.int32)


def _check(vdf, keys, vals, mode="asc", perm64=False, want_perm=True):
    m = MODES[mode]
    exp = oracle.argsort(keys, m, 4).astype(np.int64)
    res = vdf.sort_by(_dev(keys), _dev(vals), key_type=_key_type(keys), descending=(mode == "desc"),
                      reverse_desc=(mode == "reverse"), perm64=perm64, want_perm=want_perm, want_keys=True)
    perm, out, skeys = res
    if want_perm:
        got = _perm_host(perm)
        assert np.array_equal(got.astype(np.int64), exp)
    else:
        assert perm is None
    assert np.array_equal(out.cpu().numpy().view(np.uint8), vals[exp].view(np.uint8))
    assert np.array_equal(skeys.cpu().numpy().view(np.uint8), keys[exp].view(np.uint8))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("mode", ["asc", "desc", "reverse"])
def test_sort_by_vs_oracle(vdf, mode, width):
    for i, (name, keys) in enumerate(_key_sets()):
        _say("sort_by", mode, width, name, keys.shape)
        _check(vdf, keys, _vals(keys.shape[0], width, 100 + i), mode)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 3071, 3072, 3073, 4095, 4096, 4097, 65537, 1_000_013])
def test_sort_by_tile_edges(vdf, n):
    keys = (datagen.splitmix64(n, 0, n) >> np.uint64(40)).astype(np.uint32)  # 24-bit: 3 passes
    for width in (4, 8):
        _say("sort_by edges", n, width)
        _check(vdf, keys, _vals(n, width, n))


def test_sort_by_single_pass_and_constant_keys(vdf):
    n = 50_001
    low = (datagen.splitmix64(7, 0, n) & np.uint64(0xFF)).astype(np.uint32)  # one active digit
    _check(vdf, low, _vals(n, 8, 1))
    _check(vdf, np.full(n, 9, dtype=np.uint32), _vals(n, 8, 2))  # every digit constant


def test_sort_by_without_perm_and_perm64(vdf):
    keys = datagen.groupby_small(200_003, seed=3)[0]
    _check(vdf, keys, _vals(keys.shape[0], 8, 3), want_perm=False)
    _check(vdf, keys, _vals(keys.shape[0], 4, 4), perm64=True)
    _check(vdf, keys, _vals(keys.shape[0], 8, 5), perm64=True)  # split staging, 64-bit row ids
    u32 = (datagen.splitmix64(6, 0, 300_007) >> np.uint64(32)).astype(np.uint32)
    _check(vdf, u32, _vals(u32.shape[0], 8, 6), perm64=True)
    _check(vdf, u32, _vals(u32.shape[0], 8, 7), want_perm=False)


@pytest.mark.parametrize("items,xcd", [(12, 1), (16, 1), (16, 0), (24, 1)])
def test_sort_by_independent_of_launch_shape(vdf, items, xcd):
    """Equal to argsort + gather under every pass tile size and tile order
    (a tile size without a payload instantiation falls back to the default)."""
    import torch

    keys = (datagen.splitmix64(11, 0, 2_000_003) >> np.uint64(32)).astype(np.uint32)
    vals = _vals(keys.shape[0], 8, 11)
    d_keys, d_vals = _dev(keys), _dev(vals)
    ref_perm = vdf.argsort(d_keys, key_type=N.KEY_U32)
    ref_out = vdf.gather(d_vals, ref_perm)
    old = (vdf.tune_get("sort_items_32"), vdf.tune_get("sort_xcd"))
    try:
        vdf.tune("sort_items_32", items)
        vdf.tune("sort_xcd", xcd)
        perm, out = vdf.sort_by(d_keys, d_vals, key_type=N.KEY_U32)
        torch.cuda.synchronize()
    finally:
        vdf.tune("sort_items_32", old[0])
        vdf.tune("sort_xcd", old[1])
    assert torch.equal(perm, ref_perm)
    assert torch.equal(out.view(torch.int64), ref_out.view(torch.int64))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("items", [24])
@pytest.mark.parametrize("mode", ["asc", "reverse"])
def test_sort_by_split_staging(vdf, width, items, mode):
    """Large pass tiles stage a tile one 8-byte column at a time (4-byte keys
    with a payload); bit-exact against the oracle on every key set, the last
    pass writing the permutation, sorted keys and column (reversed too)."""
    old = vdf.tune_get("sort_items_32")
    try:
        vdf.tune("sort_items_32", items)
        for i, (name, keys) in enumerate(_key_sets()):
            if keys.dtype.itemsize != 4 or keys.ndim != 1:
                continue
            _say("sort_by split", items, width, mode, name, keys.shape)
            _check(vdf, keys, _vals(keys.shape[0], width, 300 + i), mode)
        for n in (1, 6143, 6144, 6145, 8191, 8193, 1_000_003):
            keys = (datagen.splitmix64(n, 5, n) >> np.uint64(40)).astype(np.uint32)
            _check(vdf, keys, _vals(n, width, n), mode)
    finally:
        vdf.tune("sort_items_32", old)


@pytest.mark.parametrize("mode", ["asc", "reverse"])
def test_sort_by_wide_keys_split_tiles(vdf, mode):
    """8- and 16-byte keys carrying an 8-byte column with 16 rows per thread
    (split staging: key, row ids and the column in separate phases)."""
    old = (vdf.tune_get("sort_items_64"), vdf.tune_get("sort_items_128"))
    try:
        vdf.tune("sort_items_64", 16)
        vdf.tune("sort_items_128", 16)
        for i, (name, keys) in enumerate(_key_sets()):
            if keys.dtype.itemsize == 4 and keys.ndim == 1:
                continue
            _say("sort_by wide split", mode, name, keys.shape)
            _check(vdf, keys, _vals(keys.shape[0], 8, 500 + i), mode)
    finally:
        vdf.tune("sort_items_64", old[0])
        vdf.tune("sort_items_128", old[1])


def test_sort_by_rejects_bad_width(vdf):
    import torch

    keys = torch.zeros(16, dtype=torch.int32, device="cuda")
    with pytest.raises(N.VdfError) as e:
        vdf.sort_by(keys, torch.zeros(16, dtype=torch.int16, device="cuda"), key_type=N.KEY_U32)
    assert e.value.status == N.VDF_ERR_INVALID


def test_sort_by_100m_u32_f64(vdf):
    """The full-size uniform uint32 sort (100M rows) carrying a float64 column:
    identical to argsort + gather, and the permutation bit-exact vs the oracle."""
    import os

    import torch

    n = 100_000_000
    _say("sort_by 100M: generating")
    keys, val = datagen.argsort_u32(n)
    d_keys, d_val = _dev(keys), _dev(val)
    perm, out = vdf.sort_by(d_keys, d_val, key_type=N.KEY_U32)
    ref_perm = vdf.argsort(d_keys, key_type=N.KEY_U32)
    ref_out = vdf.gather(d_val, ref_perm)
    torch.cuda.synchronize()
    assert torch.equal(perm, ref_perm)
    assert torch.equal(out.view(torch.int64), ref_out.view(torch.int64))
    del ref_out
    _say("sort_by 100M: oracle")
    exp = oracle.argsort(keys, 0, min(os.cpu_count() or 8, 32)).astype(np.int64)
    assert np.array_equal(_perm_host(perm).astype(np.int64), exp)
    assert np.array_equal(out.cpu().numpy(), val[exp])
