"""vdf_sort_by_noalloc (the allocation-free vdf_sort_by) through the C ABI:
bit-identical to vdf_sort_by on the same inputs, on two calls in a row (the
second smaller, so the fixed scratch holds the first call's leftovers), for
4-, 8- and 16-byte keys, 4- and 8-byte columns and every order mode; sizes
beyond the fixed scratch are rejected before any work; and the calls leave
device memory use unchanged."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from visondf_amd import _native as N

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _keys(kind: str, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if kind == "u32":
        return rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if kind == "i32":
        return rng.integers(-50, 50, n).astype(np.int32)  # many ties
    if kind == "u64":
        return rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64) * np.uint64(3)
    return rng.integers(32, 127, (n, 16)).astype(np.uint8)  # char[16]


def _dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


KEY_TYPES = {"u32": N.KEY_U32, "i32": N.KEY_I32, "u64": N.KEY_U64, "char16": N.KEY_CHAR16}


def _run(lib, h, fn, kt, d_keys, n, flags, d_vals, vb, want_keys, perm64):
    import torch

    perm = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device="cuda")
    out = torch.empty(n * vb, dtype=torch.uint8, device="cuda")
    sk = torch.empty_like(d_keys) if want_keys else None
    rc = getattr(lib, fn)(h, kt, C.c_void_p(d_keys.data_ptr()), n, flags, C.c_void_p(perm.data_ptr()),
                          C.c_void_p(sk.data_ptr()) if sk is not None else None, vb,
                          C.c_void_p(d_vals.data_ptr()), C.c_void_p(out.data_ptr()),
                          C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, lib.vdf_last_error()
    return perm, out, sk


@pytest.mark.parametrize("kind", ["u32", "i32", "u64", "char16"])
def test_noalloc_matches_sort_by(kind):
    import torch

    lib = N.load()
    h = C.c_void_p()
    N.check(lib.vdf_create(C.byref(h), 0, 0))
    try:
        kt = KEY_TYPES[kind]
        for (n1, n2), (flags, vb, want_keys) in zip(
                [(1_000_003, 4_099), (1_100_000, 300_000), (262_144, 1)],
                [(0, 8, True), (N.SORT_DESC, 4, False), (N.SORT_DESC_REVERSE | N.SORT_PERM64, 8, True)]):
            for call, n in enumerate((n1, n2)):
                _say("noalloc", kind, "flags", flags, "val_bytes", vb, "call", call, "rows", n)
                keys = _keys(kind, n, 7 * n + call)
                vals = np.random.default_rng(n).integers(0, 256, n * vb).astype(np.uint8)
                d_keys = _dev(keys)
                d_vals = torch.from_numpy(vals).cuda()
                perm64 = bool(flags & N.SORT_PERM64)
                a = _run(lib, h, "vdf_sort_by", kt, d_keys, n, flags, d_vals, vb, want_keys, perm64)
                b = _run(lib, h, "vdf_sort_by_noalloc", kt, d_keys, n, flags, d_vals, vb, want_keys, perm64)
                torch.cuda.synchronize()
                assert torch.equal(a[0], b[0]), "permutation"
                assert torch.equal(a[1], b[1]), "value column"
                if want_keys:
                    assert torch.equal(a[2], b[2]), "sorted keys"
                # and the result is a stable sort of the keys (rows in order within equal keys)
                if kind in ("u32", "i32", "u64") and flags == 0:
                    p = b[0].cpu().numpy().astype(np.int64)
                    assert np.array_equal(p, np.argsort(keys, kind="stable"))
    finally:
        N.check(lib.vdf_destroy(h))


def test_noalloc_rejects_above_fixed_scratch():
    import torch

    lib = N.load()
    h = C.c_void_p()
    N.check(lib.vdf_create(C.byref(h), 0, 0))
    try:
        n = 8_000_000  # u32 keys + 8-byte column: ~33 B/row of scratch > 64 MiB
        d_keys = torch.zeros(n, dtype=torch.int32, device="cuda")
        d_vals = torch.zeros(n, dtype=torch.float64, device="cuda")
        out = torch.empty_like(d_vals)
        rc = lib.vdf_sort_by_noalloc(h, N.KEY_U32, C.c_void_p(d_keys.data_ptr()), n, 0, None, None, 8,
                                     C.c_void_p(d_vals.data_ptr()), C.c_void_p(out.data_ptr()), None)
        assert rc == N.VDF_ERR_INVALID
        assert b"VDF_NOALLOC_SCRATCH_BYTES" in lib.vdf_last_error()
    finally:
        N.check(lib.vdf_destroy(h))


def test_noalloc_allocates_nothing():
    """A fresh handle's no-alloc calls leave the device's free memory as it
    was (vdf_sort_by on the same handle grows a ~30 MB scratch arena)."""
    import torch

    lib = N.load()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    hip.hipMemGetInfo.restype = C.c_int

    def free_bytes():
        torch.cuda.synchronize()
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    n = 1_000_000
    keys = _keys("u32", n, 3)
    d_keys = _dev(keys)
    d_vals = torch.from_numpy(np.arange(n, dtype=np.float64)).cuda()
    out = torch.empty_like(d_vals)
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    h = C.c_void_p()
    N.check(lib.vdf_create(C.byref(h), 0, 0))
    try:
        args = (N.KEY_U32, C.c_void_p(d_keys.data_ptr()), n, 0, C.c_void_p(perm.data_ptr()), None, 8,
                C.c_void_p(d_vals.data_ptr()), C.c_void_p(out.data_ptr()), None)
        N.check(lib.vdf_sort_by_noalloc(h, *args))  # first call on the device: code object loaded
        before = free_bytes()
        for _ in range(3):
            N.check(lib.vdf_sort_by_noalloc(h, *args))
        after = free_bytes()
        assert before - after < (1 << 20), (before, after)
        assert np.array_equal(out.cpu().numpy(), np.arange(n, dtype=np.float64)[np.argsort(keys, kind="stable")])
        N.check(lib.vdf_sort_by(h, *args))
        assert before - free_bytes() >= 16 * n  # the ordinary call does allocate its arena
    finally:
        N.check(lib.vdf_destroy(h))
