"""ctypes access to the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
baseline.  Only tests/, the smoke check in __graft_entry__.py and the
cpu_baseline leg of bench.py import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libvref.so")

KEY_U32, KEY_I32, KEY_U64, KEY_I64, KEY_CHAR16, KEY_F32, KEY_F64 = 0, 1, 2, 3, 4, 5, 6
VAL_F64, VAL_I64, VAL_I32, VAL_U32, VAL_U64 = 0, 1, 2, 3, 4

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
        L.orc_argsort.restype = i32
        L.orc_argsort.argtypes = [i32, vp, sz, i32, i32, vp]
        L.orc_gather.restype = None
        L.orc_gather.argtypes = [sz, vp, vp, sz, vp, i32]
        L.orc_exact_sum.restype = C.c_double
        L.orc_exact_sum.argtypes = [vp, sz]
        L.orc_groupby.restype = C.c_uint64
        L.orc_groupby.argtypes = [i32, vp, sz, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]
        L.orc_hash_groupby_f64.restype = C.c_uint64
        L.orc_hash_groupby_f64.argtypes = [i32, vp, sz, vp, i32, i32, vp, vp, vp, vp, vp, vp]
        L.orc_hash_bytes16.restype = C.c_uint64
        L.orc_hash_bytes16.argtypes = [vp, sz]
        for name in ("orc_radix_sort_u32", "orc_radix_sort_u64"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [vp, vp, sz]
        for name in ("orc_radix_sort_u32_mt", "orc_radix_sort_u64_mt"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [vp, vp, sz, i32]
        L.orc_radix_sort_charbuf.restype = None
        L.orc_radix_sort_charbuf.argtypes = [vp, sz, sz, vp]
        ch = C.c_char
        L.orc_csv_infer.restype = i32
        L.orc_csv_infer.argtypes = [vp, sz, ch, ch, i32, i32, vp, i32]
        L.orc_csv_lines.restype = C.c_uint64
        L.orc_csv_lines.argtypes = [vp, sz, vp, i32]
        L.orc_csv_measure.restype = C.c_int64
        L.orc_csv_measure.argtypes = [vp, sz, vp, C.c_uint64, ch, ch, i32, vp, vp, i32]
        L.orc_csv_convert.restype = None
        L.orc_csv_convert.argtypes = [vp, sz, vp, C.c_uint64, ch, ch, ch, i32, vp, vp, vp, i32]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own radix .inl files compiled in place (None if unbuilt)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = C.CDLL(REF_LIB_PATH)
        vp, sz = C.c_void_p, C.c_size_t
        for name in ("vref_radix_sort_u32", "vref_radix_sort_u32_mt8", "vref_radix_sort_u32_mt16",
                     "vref_radix_sort_u32_mt32", "vref_radix_sort_u64", "vref_radix_sort_u64_mt8"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [vp, vp, sz]
        L.vref_radix_sort_charbuf.restype = None
        L.vref_radix_sort_charbuf.argtypes = [vp, sz, sz, vp]
        L.vref_hash_groupby.restype = C.c_uint64
        L.vref_hash_groupby.argtypes = [vp, sz, sz, vp, C.c_int, vp, C.c_int, C.c_int] + [vp] * 8
        L.vref_simd_hash.restype = C.c_uint64
        L.vref_simd_hash.argtypes = [vp, sz]
        _ref = L
    return _ref


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def key_type_of(a: np.ndarray) -> int:
    if a.dtype == np.uint8 and a.ndim == 2 and a.shape[1] == 16:
        return KEY_CHAR16
    return {np.dtype(np.uint32): KEY_U32, np.dtype(np.int32): KEY_I32,
            np.dtype(np.uint64): KEY_U64, np.dtype(np.int64): KEY_I64,
            np.dtype(np.float32): KEY_F32, np.dtype(np.float64): KEY_F64}[a.dtype]


def val_type_of(a: np.ndarray) -> int:
    return {np.dtype(np.float64): VAL_F64, np.dtype(np.int64): VAL_I64, np.dtype(np.int32): VAL_I32,
            np.dtype(np.uint32): VAL_U32, np.dtype(np.uint64): VAL_U64}[a.dtype]


def argsort(keys: np.ndarray, mode: int = 0, threads: int = 1) -> np.ndarray:
    """Stable argsort as the reference computes it; mode 0 asc, 1 stable desc, 2 reversed asc."""
    keys = np.ascontiguousarray(keys)
    n = keys.shape[0]
    perm = np.empty(n, dtype=np.uint64)
    rc = lib().orc_argsort(key_type_of(keys), _p(keys), n, mode, threads, _p(perm))
    assert rc == 0
    return perm


def argsort_bytes(keys: np.ndarray, mode: int = 0) -> np.ndarray:
    """Stable argsort of fixed-width byte strings ((n, width) uint8, padded):
    each byte ^ 0x80 (signed char order), then the byte-wise LSD from the last
    byte to the first, as sort_string.inl:30-85 + radix_sort_charbuf_flat.inl:
    4-107 (Simd=false); mode as argsort."""
    keys = np.ascontiguousarray(keys, dtype=np.uint8)
    n, width = keys.shape
    flat = keys ^ np.uint8(0x80 ^ (0xFF if mode == 1 else 0))
    perm = np.arange(n, dtype=np.uint64)
    if n and width:
        lib().orc_radix_sort_charbuf(_p(flat), n, width, _p(perm))
    if mode == 2:
        perm = perm[::-1].copy()
    return perm


def gather(src: np.ndarray, perm: np.ndarray, threads: int = 1) -> np.ndarray:
    src = np.ascontiguousarray(src)
    perm = np.ascontiguousarray(perm, dtype=np.uint64)
    out = np.empty((perm.shape[0],) + src.shape[1:], dtype=src.dtype)
    eb = src.dtype.itemsize * (src.shape[1] if src.ndim == 2 else 1)
    lib().orc_gather(eb, _p(src), _p(perm), perm.shape[0], _p(out), threads)
    return out


def exact_sum(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_exact_sum(_p(x), x.shape[0])


def groupby(keys: np.ndarray, vals=(), threads: int = 1) -> dict:
    """Sort-based group-by (ascending keys): keys, offsets, counts, perm and per
    value column sum / mean / min / max."""
    keys = np.ascontiguousarray(keys)
    n = keys.shape[0]
    nv = len(vals)
    vals = [np.ascontiguousarray(v) for v in vals]
    vt = np.array([val_type_of(v) for v in vals] or [0], dtype=np.int32)
    vptr = (C.c_void_p * max(nv, 1))(*[v.ctypes.data for v in vals])
    gk = np.empty_like(keys) if n else np.empty((0,) + keys.shape[1:], dtype=keys.dtype)
    off = np.empty(n + 1, dtype=np.uint64)
    cnt = np.empty(max(n, 1), dtype=np.uint64)
    perm = np.empty(max(n, 1), dtype=np.uint64)
    sums, means, mins, maxs = [], [], [], []
    for v in vals:
        sums.append(np.empty(max(n, 1), dtype=np.float64 if v.dtype == np.float64 else np.int64))
        means.append(np.empty(max(n, 1), dtype=np.float64))
        mins.append(np.empty(max(n, 1), dtype=v.dtype))
        maxs.append(np.empty(max(n, 1), dtype=v.dtype))

    def arr(lst):
        return (C.c_void_p * max(nv, 1))(*[a.ctypes.data for a in lst])

    g = lib().orc_groupby(key_type_of(keys), _p(keys), n, nv, _p(vt), vptr, _p(gk), _p(off), _p(cnt), _p(perm),
                          arr(sums), arr(means), arr(mins), arr(maxs), threads)
    g = int(g)
    return {
        "n_groups": g, "keys": gk[:g], "offsets": off[:g + 1], "counts": cnt[:g], "perm": perm[:n],
        "aggs": [{"sum": sums[c][:g], "mean": means[c][:g], "min": mins[c][:g], "max": maxs[c][:g]}
                 for c in range(nv)],
    }


def ref_hash_groupby(keys: np.ndarray, v0: np.ndarray, v1: np.ndarray | None = None, threads: int = 1,
                     refapi: bool = False, outputs: bool = True) -> dict:
    """The reference's hash group-by (its ankerl map and simd_hash, compiled
    in oracle/_ref, under loops restated from onecol_functions.inl and
    merge_functions.inl) over a key column, a float64 column and an optional
    int32/int64 column.  refapi: one hash pass per aggregate, as the
    reference's API runs them; else one fused pass.  Groups come in the
    reference's iteration order.  outputs=False only times the work."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libvref.so is not built")
    keys = np.ascontiguousarray(keys)
    v0 = np.ascontiguousarray(v0, dtype=np.float64)
    n = keys.shape[0]
    kb = keys.itemsize * (keys.shape[1] if keys.ndim == 2 else 1)
    k1 = -1 if v1 is None else {np.dtype(np.int64): 1, np.dtype(np.int32): 2}[v1.dtype]
    cap = max(n, 1) if outputs else 0
    gk = np.empty((cap,) + keys.shape[1:], dtype=keys.dtype) if outputs else None
    cnt = np.empty(cap, dtype=np.uint64) if outputs else None
    f = [np.empty(cap, dtype=np.float64) if outputs else None for _ in range(3)]
    i = [np.empty(cap, dtype=np.int64) if outputs and v1 is not None else None for _ in range(3)]
    g = int(L.vref_hash_groupby(_p(keys), kb, n, _p(v0), k1, _p(v1), threads, 1 if refapi else 0, _p(gk), _p(cnt),
                                _p(f[0]), _p(f[1]), _p(f[2]), _p(i[0]), _p(i[1]), _p(i[2])))
    if not outputs:
        return {"n_groups": g}
    r = {"n_groups": g, "keys": gk[:g], "counts": cnt[:g], "sum0": f[0][:g], "min0": f[1][:g], "max0": f[2][:g]}
    if v1 is not None:
        r.update({"sum1": i[0][:g], "min1": i[1][:g], "max1": i[2][:g]})
    return r


def ref_simd_hash(key: bytes) -> int:
    """The reference's simd_hash of a key (oracle/_ref)."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libvref.so is not built")
    b = np.frombuffer(key, dtype=np.uint8).copy()
    return int(L.vref_simd_hash(_p(b), len(key)))


def hash_bytes16(key: bytes) -> int:
    """The oracle's restatement of the reference's key hash (orc_hash_bytes16)."""
    b = np.frombuffer(key, dtype=np.uint8).copy()
    return int(lib().orc_hash_bytes16(_p(b), len(key)))


def hash_groupby_f64(keys: np.ndarray, val: np.ndarray, threads: int = 1, refapi: bool = False) -> dict:
    """The reference's hash group-by restated (first-occurrence group order)."""
    keys = np.ascontiguousarray(keys)
    val = np.ascontiguousarray(val, dtype=np.float64)
    n = keys.shape[0]
    gk = np.empty_like(keys)
    cnt = np.empty(max(n, 1), dtype=np.uint64)
    s, m, mn, mx = (np.empty(max(n, 1), dtype=np.float64) for _ in range(4))
    g = int(lib().orc_hash_groupby_f64(key_type_of(keys), _p(keys), n, _p(val), threads, 1 if refapi else 0,
                                       _p(gk), _p(cnt), _p(s), _p(m), _p(mn), _p(mx)))
    return {"n_groups": g, "keys": gk[:g], "counts": cnt[:g], "sum": s[:g], "mean": m[:g], "min": mn[:g],
            "max": mx[:g]}


# ---------------------------------------------------------------------------
# multi-column keys
# ---------------------------------------------------------------------------
def order_bits(a: np.ndarray) -> np.ndarray:
    """Order-preserving unsigned form of one key column, as the sort maps keys:
    unsigned as is, signed with the sign bit flipped (sort_integers.inl:46,53),
    floats by the transform of sort_flt.inl:44-50, char[16] bytes ^ 0x80
    (sort_char.inl:73; returned as a (n, 16) uint8 array)."""
    if a.ndim == 2:
        return a ^ np.uint8(0x80)
    if a.dtype == np.float32:
        b = a.view(np.uint32)
        return b ^ np.where(b >> np.uint32(31), np.uint32(0xFFFFFFFF), np.uint32(0x80000000))
    if a.dtype == np.float64:
        b = a.view(np.uint64)
        return b ^ np.where(b >> np.uint64(63), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(1 << 63))
    if a.dtype == np.int32:
        return a.view(np.uint32) ^ np.uint32(0x80000000)
    if a.dtype == np.int64:
        return a.view(np.uint64) ^ np.uint64(1 << 63)
    return a


def argsort_multi(keys_list, mode: int = 0) -> np.ndarray:
    """Stable argsort by several key columns, lexicographic with the first
    column most significant; mode as argsort (1: stable descending over the
    whole key, 2: ascending reversed)."""
    cols = []  # numpy lexsort: the last entry is the primary key
    for k in reversed(list(keys_list)):
        ob = order_bits(np.ascontiguousarray(k))
        if mode == 1:
            ob = ~ob
        if ob.ndim == 2:  # char[16] or a byte-string column: bytes ^ 0x80, first byte most significant
            cols.extend(ob[:, j] for j in range(ob.shape[1] - 1, -1, -1))
        else:
            cols.append(ob)
    perm = np.lexsort(cols) if cols and len(cols[0]) else np.zeros(0, dtype=np.int64)
    if mode == 2:
        perm = perm[::-1]
    return perm.astype(np.uint64)


def groupby_multi(keys_list, vals=(), threads: int = 1) -> dict:
    """Group-by on several key columns: rows are equal when every column's
    bytes are equal (the reference's concatenated key bytes, key_build.inl:3-110);
    groups in lexicographic column order.  Computed as the single-key group-by
    of each row's dense rank among the distinct keys, so the aggregates, the
    permutation and the offsets come from the pinned single-key oracle."""
    keys_list = [np.ascontiguousarray(k) for k in keys_list]
    n = keys_list[0].shape[0]
    perm = argsort_multi(keys_list, 0).astype(np.int64)
    diff = np.zeros(max(n - 1, 0), dtype=bool)
    for k in keys_list:
        b = k.reshape(n, -1).view(np.uint8) if k.ndim == 2 else k.view(np.uint8).reshape(n, -1)
        s = b[perm]
        diff |= np.any(s[1:] != s[:-1], axis=1)
    rank_sorted = np.concatenate([[0], np.cumsum(diff)]).astype(np.uint32) if n else np.zeros(0, np.uint32)
    rank = np.empty(n, dtype=np.uint32)
    rank[perm] = rank_sorted
    r = groupby(rank, list(vals), threads)
    first = r["perm"][r["offsets"][:-1].astype(np.int64)].astype(np.int64)
    r["keys"] = [k[first] for k in keys_list]
    return r


CSV_STRING, CSV_CHAR, CSV_BOOL, CSV_INT, CSV_UINT, CSV_DOUBLE = range(6)


def read_csv(text, delim: str = ",", quote: str = "'", header: bool = True, int_bytes: int = 8,
             char_width: int = 0, trailing: str = "0", threads: int = 1) -> dict:
    """The oracle's CSV ingest (orc_csv_*): returns {"names", "kinds",
    "widths", "cols"}; cols[c] is a numpy array (float64 / int / uint32 for
    bool) or, for string and char columns, an (nrow, width) uint8 array."""
    t = np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text
    t = np.ascontiguousarray(t, dtype=np.uint8)
    L = lib()
    d, q, tr = delim.encode(), quote.encode(), trailing.encode()
    kinds = np.zeros(65536, dtype=np.int32)
    ncol = L.orc_csv_infer(_p(t), t.size, d, q, int(header), char_width, _p(kinds), kinds.size)
    if ncol < 0:
        raise ValueError("more than 65535 columns")
    kinds = kinds[:ncol].copy()
    line = np.zeros(int((t == 10).sum()) + 2, dtype=np.uint64)
    nl = int(L.orc_csv_lines(_p(t), t.size, _p(line), threads))
    first = 1 if header else 0
    nrow = max(nl - first, 0)
    lv = line[first:]
    names = []
    if ncol:
        e = int(line[1]) - 1 if nl else t.size
        hl = bytes(t[: min(e, t.size)]).rstrip(b"\r").decode("latin-1")
        raw, cur, inq = [], "", False
        for ch_ in hl:
            if ch_ == quote:
                inq = not inq
                cur += ch_
            elif ch_ == delim and not inq:
                raw.append(cur)
                cur = ""
            else:
                cur += ch_
        raw.append(cur)
        names = [(x if x else "NA") if header else "" for x in raw]
    maxlen = np.zeros(max(ncol, 1), dtype=np.uint32)
    bad = int(L.orc_csv_measure(_p(t), t.size, _p(lv), nrow, d, q, ncol, _p(kinds), _p(maxlen), threads)) \
        if nrow else -1
    if bad >= 0:
        raise ValueError(f"data row {bad} does not have {ncol} fields")
    widths = np.zeros(max(ncol, 1), dtype=np.uint32)
    cols = []
    for c in range(ncol):
        k = int(kinds[c])
        if k == CSV_DOUBLE:
            w, a = 8, np.zeros(nrow, dtype=np.float64)
        elif k == CSV_INT:
            w, a = int_bytes, np.zeros(nrow, dtype=np.int64 if int_bytes == 8 else np.int32)
        elif k == CSV_UINT:
            w, a = int_bytes, np.zeros(nrow, dtype=np.uint64 if int_bytes == 8 else np.uint32)
        elif k == CSV_BOOL:
            w, a = 4, np.zeros(nrow, dtype=np.uint32)
        elif k == CSV_CHAR:
            w = char_width
            a = np.zeros((nrow, w), dtype=np.uint8)
        else:
            w = max(int(maxlen[c]), 1)
            a = np.zeros((nrow, w), dtype=np.uint8)
        widths[c] = w
        cols.append(a)
    ptrs = (C.c_void_p * max(ncol, 1))(*[a.ctypes.data for a in cols])
    if nrow:
        L.orc_csv_convert(_p(t), t.size, _p(lv), nrow, d, q, tr, ncol, _p(kinds), _p(widths), ptrs, threads)
    return {"names": names, "kinds": [int(k) for k in kinds], "widths": [int(w) for w in widths[:ncol]],
            "cols": cols, "nrow": nrow}
