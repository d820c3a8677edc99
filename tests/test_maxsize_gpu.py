"""GPU parity at the row limit and near the largest local-finish size.

A table holds up to VDF_MAX_ROWS = 2^30 - 1 rows (the look-back words carry
30-bit counts).  At that size the oracle would take minutes, so the checks
are the properties that fix a stable sort and do not depend on size, all on
the device: the permutation is a permutation (every row id once), the sorted
keys are the keys at the permutation and never decrease, equal keys keep
their row order (the stability of the reference's LSD radix,
radix_sort_uint32_mt.inl:6-164), and the carried column (the row index as
f64) follows its rows."""
from __future__ import annotations

import sys

import pytest

from visondf_amd import _native as N

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _keys(n: int, seed: int):
    """n pseudo-random u32 keys (as int64 in [0, 2^32)), with repeats."""
    import torch

    x = torch.arange(n, dtype=torch.int64, device="cuda") + seed
    x = x * -7046029254386353131  # 0x9E3779B97F4A7C15, wrapped
    x = x ^ (x >> 31)
    x = x * -4658895280553007687  # 0xBF58476D1CE4E5B9, wrapped
    return (x >> 32) & 0xFFFFFFFF


def _check(keys64, perm, vals_out, keys_out):
    import torch

    n = keys64.numel()
    p = perm.long()
    del perm
    # the carried column is the row index: it must follow its rows
    assert torch.equal(vals_out, p.double())
    cnt = torch.bincount(p, minlength=n)
    assert cnt.numel() == n and int(cnt.min()) == 1 and int(cnt.max()) == 1
    del cnt
    sk = keys64[p]
    assert torch.equal(sk, keys_out.long() & 0xFFFFFFFF)
    d = sk[1:] - sk[:-1]
    del sk
    assert bool((d >= 0).all())
    tie = d == 0
    del d
    assert int(tie.sum()) > 0  # the keys repeat, so stability is exercised
    assert bool((p[1:][tie] > p[:-1][tie]).all())


def _sort_by(vdf, keys64, local: int):
    import torch

    n = keys64.numel()
    keys = keys64.to(torch.int32)  # the u32 bit patterns
    vals = torch.arange(n, dtype=torch.float64, device="cuda")
    old = vdf.tune_get("sort_local")
    vdf.profile(True)
    vdf.profile_reset()
    try:
        vdf.tune("sort_local", local)
        perm, vals_out, keys_out = vdf.sort_by(keys, vals, key_type=N.KEY_U32, want_keys=True)
        torch.cuda.synchronize()
        _, launches, _ = vdf.profile_read("radix_local")
        _, passes, _ = vdf.profile_read("radix_pass")
    finally:
        vdf.tune("sort_local", old)
        vdf.profile(False)
        vdf.profile_reset()
    del keys, vals
    return perm, vals_out, keys_out, launches, passes


def test_sort_by_max_rows(vdf):
    """2^30 - 1 rows: buckets of equal top 16 bits hold about 16K rows, so the
    plan keeps the four plain passes."""
    import torch

    n = N.MAX_ROWS
    _say("max rows sort_by", n)
    keys64 = _keys(n, 1)
    perm, vals_out, keys_out, launches, passes = _sort_by(vdf, keys64, 1)
    assert launches == 0 and passes == 4
    _check(keys64, perm, vals_out, keys_out)
    del keys64, perm, vals_out, keys_out
    torch.cuda.empty_cache()


def test_sort_by_local_finish_190m(vdf):
    """190M rows, near the largest uniform table the plan finishes locally
    (its estimate of the largest bucket, 3,072 rows, is reached near 200M):
    buckets of about 2,900 rows, one per greedy chunk."""
    import torch

    n = 190_000_007
    _say("local finish sort_by", n)
    keys64 = _keys(n, 2)
    perm, vals_out, keys_out, launches, passes = _sort_by(vdf, keys64, 1)
    assert launches == 1 and passes == 2
    _check(keys64, perm, vals_out, keys_out)
    del keys64, perm, vals_out, keys_out
    torch.cuda.empty_cache()
