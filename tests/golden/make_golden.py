"""Writes tests/golden/radix_golden.json from the reference's own radix sorts.

Run in a container where the reference tree exists:
    make -C oracle ref && python tests/golden/make_golden.py

oracle/_ref/libvref.so is the reference's radix_sort_uint32 / uint32_mt /
uint64 / uint64_mt / charbuf_flat (.inl files compiled where they lie, with
Simd=false).  Inputs come from visondf_amd.datagen (seeded), transformed the
way the reference's sort_by dispatch transforms them (sort_integers.inl:46,
sort_char.inl:73), so each record holds: how to regenerate the input, a
SHA-256 of the input bytes, and the reference's permutation (in full for small
n, as a SHA-256 of the uint32 little-endian permutation otherwise).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests.golden import cases  # noqa: E402

FULL_LIMIT = 4096


def ref_perm(kind: str, keys: np.ndarray, mt: bool) -> np.ndarray:
    L = oracle.ref_lib()
    n = keys.shape[0]
    idx = np.arange(n, dtype=np.uint64)
    if kind in ("u32", "i32"):
        t = keys.view(np.uint32).copy()
        if kind == "i32":
            t ^= np.uint32(0x80000000)
        (L.vref_radix_sort_u32_mt8 if mt else L.vref_radix_sort_u32)(t.ctypes.data, idx.ctypes.data, n)
    elif kind in ("u64", "i64"):
        t = keys.view(np.uint64).copy()
        if kind == "i64":
            t ^= np.uint64(0x8000000000000000)
        (L.vref_radix_sort_u64_mt8 if mt else L.vref_radix_sort_u64)(t.ctypes.data, idx.ctypes.data, n)
    elif kind == "char16":
        flat = np.ascontiguousarray(keys) ^ np.uint8(0x80)
        L.vref_radix_sort_charbuf(flat.ctypes.data, n, 16, idx.ctypes.data)
    else:
        raise ValueError(kind)
    return idx


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libvref.so is not built: run `make -C oracle ref` first")
    out = []
    for case in cases.CASES:
        keys = cases.make_keys(case)
        rec = dict(case)
        rec["key_sha256"] = hashlib.sha256(np.ascontiguousarray(keys).tobytes()).hexdigest()
        perm = ref_perm(case["kind"], keys, mt=False)
        if case["kind"] != "char16" and keys.shape[0] > 0:
            perm_mt = ref_perm(case["kind"], keys, mt=True)
            assert np.array_equal(perm, perm_mt), f"reference mt/serial disagree on {case}"
        p32 = perm.astype(np.uint32)
        rec["perm_sha256"] = hashlib.sha256(p32.tobytes()).hexdigest()
        if keys.shape[0] <= FULL_LIMIT:
            rec["perm"] = [int(x) for x in p32]
        out.append(rec)
        print(case["name"], keys.shape[0], rec["perm_sha256"][:16])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radix_golden.json")
    with open(path, "w") as f:
        json.dump({"source": "reference radix .inl files (Simd=false), compiled by oracle/Makefile ref",
                   "cases": out}, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
