"""Input recipes of the golden radix vectors (shared by make_golden.py and tests)."""
from __future__ import annotations

import numpy as np

from visondf_amd import datagen


def _u32(seed, n):
    return (datagen.splitmix64(seed, 0, n) >> np.uint64(32)).astype(np.uint32)


def _u64(seed, n):
    return datagen.splitmix64(seed, 0, n)


def make_keys(case: dict) -> np.ndarray:
    kind, n, seed, dist = case["kind"], case["n"], case["seed"], case["dist"]
    if kind in ("u32", "i32"):
        if dist == "uniform":
            k = _u32(seed, n)
        elif dist == "small":  # few distinct values
            k = (datagen.splitmix64(seed, 0, n) % np.uint64(4)).astype(np.uint32)
        elif dist == "equal":
            k = np.full(n, 0xDEADBEEF, dtype=np.uint32)
        elif dist == "sorted":
            k = np.sort(_u32(seed, n))
        elif dist == "reverse":
            k = np.sort(_u32(seed, n))[::-1].copy()
        elif dist == "byte2":  # only the third byte varies
            k = ((datagen.splitmix64(seed, 0, n) % np.uint64(256)).astype(np.uint32) << np.uint32(16)) | np.uint32(7)
        else:
            raise ValueError(dist)
        return k.view(np.int32) if kind == "i32" else k
    if kind in ("u64", "i64"):
        if dist == "uniform":
            k = _u64(seed, n)
        elif dist == "c1":  # int64 key uniform in [0, 1000)
            k = (datagen.splitmix64(seed, 0, n) % np.uint64(1000)).astype(np.uint64)
        elif dist == "signed_small":
            k = (datagen.uniform_int(seed, 0, n, -500, 500)).astype(np.int64).view(np.uint64)
        else:
            raise ValueError(dist)
        return k.view(np.int64) if kind == "i64" else k
    if kind == "char16":
        if dist == "uniform":
            return datagen.char16_keys(seed, 0, n)
        if dist == "few":
            return datagen.char16_keys(seed, 0, n, distinct=37)
        raise ValueError(dist)
    raise ValueError(kind)


CASES = []
for n in (0, 1, 2, 17, 255, 256, 257, 4095, 4096, 65537, 1_000_003):
    CASES.append({"name": f"u32_uniform_{n}", "kind": "u32", "n": n, "seed": 11, "dist": "uniform"})
for dist in ("small", "equal", "sorted", "reverse", "byte2"):
    CASES.append({"name": f"u32_{dist}", "kind": "u32", "n": 70_001, "seed": 12, "dist": dist})
for n in (17, 4096, 100_003):
    CASES.append({"name": f"i32_uniform_{n}", "kind": "i32", "n": n, "seed": 13, "dist": "uniform"})
for n in (1, 17, 4096, 65537, 300_007):
    CASES.append({"name": f"u64_uniform_{n}", "kind": "u64", "n": n, "seed": 14, "dist": "uniform"})
CASES.append({"name": "i64_c1_1000000", "kind": "i64", "n": 1_000_000, "seed": 1, "dist": "c1"})
CASES.append({"name": "i64_signed_small", "kind": "i64", "n": 50_000, "seed": 15, "dist": "signed_small"})
for n in (17, 4096, 20_000):
    CASES.append({"name": f"char16_uniform_{n}", "kind": "char16", "n": n, "seed": 16, "dist": "uniform"})
CASES.append({"name": "char16_few", "kind": "char16", "n": 30_000, "seed": 17, "dist": "few"})
