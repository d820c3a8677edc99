"""The C ABI: the library loads, exports every symbol include/vdf.h declares,
and rejects bad arguments before touching the device (CPU-only checks)."""
from __future__ import annotations

import ctypes as C
import os
import re

from visondf_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vdf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(vdf_[a-z0-9_]+)\s*\(", src))


def test_header_and_binding_agree():
    assert _declared() == set(N.EXPORTS)


def test_library_exports_every_symbol():
    lib = N.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_string():
    assert b"gfx950" in N.load().vdf_version()


def test_argument_validation_without_device():
    lib = N.load()
    # null handle is rejected before any HIP call
    rc = lib.vdf_argsort(None, N.KEY_U32, C.c_void_p(16), 10, 0, C.c_void_p(16), None, None)
    assert rc == N.VDF_ERR_INVALID
    assert b"null handle" in lib.vdf_last_error()
    rc = lib.vdf_gather(None, 8, None, None, 0, 1, None, None)
    assert rc == N.VDF_ERR_INVALID
    rc = lib.vdf_groupby(None, N.KEY_I32, None, 0, None, 0, 0, None, None)
    assert rc == N.VDF_ERR_INVALID
    rc = lib.vdf_create(None, 0, 0)
    assert rc == N.VDF_ERR_INVALID


def test_scratch_sizes():
    lib = N.load()
    assert lib.vdf_argsort_scratch_bytes(99, 10, 0, 0) == 0
    small = lib.vdf_argsort_scratch_bytes(N.KEY_U32, 1000, 0, 0)
    big = lib.vdf_argsort_scratch_bytes(N.KEY_U32, 100_000_000, 0, 0)
    assert 0 < small < big
    # two key buffers + two row-id buffers dominate: 16 B/row for u32 keys
    assert 16 * 100_000_000 <= big < 18 * 100_000_000
    assert lib.vdf_argsort_scratch_bytes(N.KEY_CHAR16, 1000, 0, 0) > small


def test_header_constants_match_binding():
    """Every VDF_KEY_/VAL_/AGG_/SORT_/ERR_ constant of include/vdf.h has the
    same value in the Python binding."""
    src = open(os.path.join(ROOT, "include", "vdf.h")).read()
    defs = dict(re.findall(r"#define\s+VDF_([A-Z0-9_]+)\s+([0-9xA-Fa-fu]+)", src))
    checked = 0
    for name, text in defs.items():
        val = int(text.rstrip("uU"), 0)
        prefix = name.split("_")[0]
        if prefix in ("KEY", "VAL", "AGG", "SORT"):
            assert getattr(N, name) == val, name
            checked += 1
        elif prefix == "ERR" or name == "OK":
            assert getattr(N, "VDF_" + name) == val, name
            checked += 1
    assert checked >= 25
