"""Seeded synthetic inputs for the radix-sort / group-by workloads.

All columns come from SplitMix64 streams keyed by (seed, column id), generated
on the host (libvdf_datagen.so, OpenMP; numpy restatements below when it is
not built), so the same arrays feed the GPU path and the CPU baseline
bit-for-bit.  The library and the numpy code agree bit-for-bit except for
normal01, where libm and numpy's log / cos may differ by a couple of ulps.

Workloads (row counts as used by the benchmark and the parity tests):
  groupby_small   n rows, int64 key uniform in [0, 1000), f64 value uniform
                  [0, 1), int32 value uniform in [-1e6, 1e6]
  argsort_u32     n rows, uint32 key uniform over 2^32, f64 value column
  zipf_groupby    n rows, int32 key = bijective scramble of a Zipf(s) rank over
                  2^24 values, f64 value N(0, 1), int64 value uniform [0, 2^40)
  batched_tables  table row counts log-uniform in [lo, hi], int32 keys uniform
                  in [0, rows / 8), f64 values uniform [0, 1)
  served_tables   row counts log-uniform in [10K, 50M]; uint64 or 16-byte ASCII
                  keys from rows / 16 distinct values, f64 values uniform [0, 1)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvdf_datagen.so")
_lib = None


def _gen():
    """The OpenMP generator library (libvdf_datagen.so) or None when unbuilt."""
    global _lib
    if _lib is None and os.path.exists(_LIB_PATH):
        L = C.CDLL(_LIB_PATH)
        u64, vp, i32, dbl = C.c_uint64, C.c_void_p, C.c_int, C.c_double
        L.vdg_splitmix64.argtypes = [u64, u64, u64, u64, vp, i32]
        L.vdg_uniform01.argtypes = [u64, u64, u64, vp, i32]
        L.vdg_normal01.argtypes = [u64, u64, u64, vp, i32]
        L.vdg_zipf_ranks.argtypes = [u64, u64, u64, dbl, u64, vp, i32]
        L.vdg_csv_airline.argtypes = [u64, u64, vp, u64, i32]
        L.vdg_csv_airline.restype = u64
        _lib = L
    return _lib


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, col: int, n: int, start: int = 0, use_lib: bool = True) -> np.ndarray:
    """n outputs of the SplitMix64 stream for (seed, col), from index `start`."""
    L = _gen() if use_lib else None
    if L is not None:
        out = np.empty(n, dtype=np.uint64)
        L.vdg_splitmix64(seed & 0xFFFFFFFFFFFFFFFF, col, start, n, out.ctypes.data, _threads())
        return out
    with np.errstate(over="ignore"):
        s0 = _mix(np.array([np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLDEN + np.uint64(col & 0xFFFF)],
                           dtype=np.uint64))[0]
        out = np.empty(n, dtype=np.uint64)
        chunk = 1 << 24
        for b in range(0, n, chunk):
            m = min(chunk, n - b)
            i = np.arange(start + b + 1, start + b + m + 1, dtype=np.uint64)
            out[b:b + m] = _mix(s0 + i * _GOLDEN)
    return out


def uniform01(seed: int, col: int, n: int, use_lib: bool = True) -> np.ndarray:
    """Doubles uniform in [0, 1): top 53 bits of each draw times 2^-53."""
    L = _gen() if use_lib else None
    if L is not None:
        out = np.empty(n, dtype=np.float64)
        L.vdg_uniform01(seed & 0xFFFFFFFFFFFFFFFF, col, n, out.ctypes.data, _threads())
        return out
    return (splitmix64(seed, col, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_int(seed: int, col: int, n: int, lo: int, hi: int) -> np.ndarray:
    """Integers uniform in [lo, hi] (inclusive) by reduction modulo the range."""
    span = np.uint64(hi - lo + 1)
    return (splitmix64(seed, col, n) % span).astype(np.int64) + lo


def normal01(seed: int, col: int, n: int, use_lib: bool = True) -> np.ndarray:
    """N(0, 1) by the Box-Muller transform of two uniform streams."""
    L = _gen() if use_lib else None
    if L is not None:
        out = np.empty(n, dtype=np.float64)
        L.vdg_normal01(seed & 0xFFFFFFFFFFFFFFFF, col, n, out.ctypes.data, _threads())
        return out
    u1 = uniform01(seed, col, n, use_lib=False)
    u2 = uniform01(seed, col + 1000, n, use_lib=False)
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def zipf_ranks(seed: int, col: int, n: int, s: float = 1.1, domain: int = 1 << 24,
               use_lib: bool = True) -> np.ndarray:
    """Zipf(s) ranks in [0, domain) by exact inverse-CDF table lookup."""
    L = _gen() if use_lib else None
    if L is not None:
        out = np.empty(n, dtype=np.int64)
        L.vdg_zipf_ranks(seed & 0xFFFFFFFFFFFFFFFF, col, n, s, domain, out.ctypes.data, _threads())
        return out
    w = np.arange(1, domain + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = uniform01(seed, col, n, use_lib=False)
    r = np.searchsorted(cdf, u, side="right")
    return np.minimum(r, domain - 1).astype(np.int64)


def scramble24(r: np.ndarray) -> np.ndarray:
    """Bijection on [0, 2^24): odd multiplier plus offset modulo 2^24."""
    return ((r.astype(np.uint64) * np.uint64(0x9E3779) + np.uint64(0x5BD1E9)) & np.uint64(0xFFFFFF)).astype(np.int64)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def groupby_small(n: int = 1_000_000, seed: int = 1):
    key = uniform_int(seed, 0, n, 0, 999).astype(np.int64)
    vf = uniform01(seed, 1, n)
    vi = uniform_int(seed, 2, n, -1_000_000, 1_000_000).astype(np.int32)
    return key, vf, vi


def groupby_i32(n: int = 100_000_000, seed: int = 6):
    """The CPU reference configuration's group-by (key uniform in [0, 1000),
    float64 uniform [0, 1), int32 uniform [-1e6, 1e6]) at 100M rows with an
    int32 key column."""
    key = uniform_int(seed, 0, n, 0, 999).astype(np.int32)
    vf = uniform01(seed, 1, n)
    vi = uniform_int(seed, 2, n, -1_000_000, 1_000_000).astype(np.int32)
    return key, vf, vi


def groupby_i32_full(n: int = 100_000_000, seed: int = 8):
    """100M rows, int32 key uniform over the full 32-bit range (about 98M
    groups at 100M rows: four radix passes, output-heavy), float64 uniform
    [0, 1) and int32 uniform [-1e6, 1e6] values."""
    key = (splitmix64(seed, 0, n) >> np.uint64(32)).astype(np.uint32).view(np.int32)
    vf = uniform01(seed, 1, n)
    vi = uniform_int(seed, 2, n, -1_000_000, 1_000_000).astype(np.int32)
    return key, vf, vi


def argsort_u32(n: int = 100_000_000, seed: int = 2):
    key = (splitmix64(seed, 0, n) >> np.uint64(32)).astype(np.uint32)
    val = uniform01(seed, 1, n)
    return key, val


def zipf_groupby(n: int = 200_000_000, seed: int = 3, s: float = 1.1, domain: int = 1 << 24):
    key = scramble24(zipf_ranks(seed, 0, n, s, domain)).astype(np.int32)
    vf = normal01(seed, 1, n)
    vi = (splitmix64(seed, 2, n) >> np.uint64(24)).astype(np.int64)
    return key, vf, vi


def batched_table_sizes(ntables: int = 4096, lo: int = 1_000, hi: int = 4_000_000, seed: int = 4):
    u = uniform01(seed, 0, ntables)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batched_table(rows: int, table_id: int, seed: int = 4):
    key = uniform_int(seed * 1_000_003 + table_id, 1, rows, 0, max(rows // 8, 1) - 1).astype(np.int32)
    val = uniform01(seed * 1_000_003 + table_id, 2, rows)
    return key, val


def served_table_sizes(ntables: int = 64, lo: int = 10_000, hi: int = 50_000_000, seed: int = 5):
    """Row counts of the served tables, log-uniform in [lo, hi]."""
    u = uniform01(seed, 0, ntables)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def served_trace(ntables: int = 32, requests: int = 1000, rate: float = 2000.0, seed: int = 6):
    """The served workload's request trace: table row counts (log-uniform in
    [10K, 50M]), key kinds (uint64 for even tables, 16-byte ASCII for odd
    ones), Poisson arrival times (seconds, `rate` per second) and the table
    each request groups."""
    sizes = served_table_sizes(ntables, 10_000, 50_000_000, seed=5)
    kinds = ["u64" if i % 2 == 0 else "char16" for i in range(ntables)]
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    picks = rng.integers(0, ntables, requests)
    return sizes, kinds, arrivals, picks


def served_table(rows: int, table_id: int, key_kind: str, seed: int = 5):
    """One served table: `rows` keys drawn from rows / 16 distinct values
    (uint64 spread over all 64 bits, or 16-byte printable ASCII) and a f64
    value column uniform in [0, 1)."""
    s = seed * 1_000_003 + table_id
    distinct = max(rows // 16, 1)
    if key_kind == "u64":
        ids = splitmix64(s, 1, rows) % np.uint64(distinct)
        with np.errstate(over="ignore"):
            key = _mix(ids * _GOLDEN + np.uint64(s))
    elif key_kind == "char16":
        key = char16_keys(s, 1, rows, distinct=distinct)
    else:
        raise ValueError(key_kind)
    return key, uniform01(s, 2, rows)


def char16_keys(seed: int, col: int, n: int, distinct: int | None = None) -> np.ndarray:
    """n fixed-width 16-byte printable-ASCII keys (uint8 array of shape (n, 16)).

    With `distinct`, keys are drawn from that many distinct strings."""
    if distinct is not None:
        ids = (splitmix64(seed, col, n) % np.uint64(distinct)).astype(np.uint64)
        with np.errstate(over="ignore"):
            a = _mix(ids * _GOLDEN + np.uint64(seed))
            b = _mix(a ^ np.uint64(0x5DEECE66D))
    else:
        a = splitmix64(seed, col, n)
        b = splitmix64(seed, col + 500, n)
    raw = np.empty((n, 16), dtype=np.uint8)
    for j in range(8):
        raw[:, j] = ((a >> np.uint64(8 * j)) & np.uint64(0xFF)).astype(np.uint8)
        raw[:, 8 + j] = ((b >> np.uint64(8 * j)) & np.uint64(0xFF)).astype(np.uint8)
    return (raw % np.uint8(95) + np.uint8(32)).astype(np.uint8)


def csv_airline(nrows: int, seed: int = 9) -> np.ndarray:
    """CSV text (uint8 array) shaped like the US airline delay-cause table:
    a header and `nrows` rows of 21 columns -- year, month, a 2-character
    carrier code, carrier name, 3-letter airport code, a quoted airport name
    holding a comma, and 15 two-decimal numbers.  Needs libvdf_datagen.so."""
    L = _gen()
    if L is None:
        raise RuntimeError("csv_airline needs libvdf_datagen.so (make -C visondf_amd/csrc)")
    cap = 1024 + nrows * 400
    buf = np.empty(cap, dtype=np.uint8)
    n = L.vdg_csv_airline(seed, nrows, buf.ctypes.data, cap, _threads())
    if n == 0:
        raise RuntimeError("csv_airline: buffer too small")
    return buf[:n]
