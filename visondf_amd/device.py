"""Python access to libvdf_hip.so over torch device tensors.

torch supplies device memory and the current HIP stream; every operation runs
in the HIP library through the C ABI.  There is no CPU fallback: without a
GPU or without the built library these calls raise.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N

_KEY_DTYPES = {
    N.KEY_U32: (torch.uint32, 4),
    N.KEY_I32: (torch.int32, 4),
    N.KEY_U64: (torch.uint64, 8),
    N.KEY_I64: (torch.int64, 8),
    N.KEY_CHAR16: (torch.uint8, 16),
}
_VAL_DTYPES = {
    N.VAL_F64: torch.float64,
    N.VAL_I64: torch.int64,
    N.VAL_I32: torch.int32,
    N.VAL_U32: torch.uint32,
    N.VAL_U64: torch.uint64,
}


def key_type_of(t: torch.Tensor) -> int:
    if t.dtype == torch.uint8 and t.dim() == 2 and t.shape[1] == 16:
        return N.KEY_CHAR16
    for k, (dt, _) in _KEY_DTYPES.items():
        if t.dtype == dt and k != N.KEY_CHAR16:
            return k
    raise TypeError(f"unsupported key dtype {t.dtype}")


def val_type_of(t: torch.Tensor) -> int:
    for k, dt in _VAL_DTYPES.items():
        if t.dtype == dt:
            return k
    raise TypeError(f"unsupported value dtype {t.dtype}")


def _stream_ptr(stream) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


class GroupbyResult:
    """Per-group outputs, trimmed to the number of groups."""

    def __init__(self, n_groups, keys, offsets, counts, perm, aggs):
        self.n_groups = n_groups
        self.keys = keys
        self.offsets = offsets
        self.counts = counts
        self.perm = perm
        self.aggs = aggs  # list of dicts: sum / mean / min / max


class Vdf:
    """One library handle (one scratch arena) on a device."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = N.load()
        self.device = device
        h = C.c_void_p()
        N.check(self.lib.vdf_create(C.byref(h), device, hbm_budget_bytes))
        self.h = h

    def close(self):
        if self.h:
            N.check(self.lib.vdf_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- argsort
    def argsort(self, keys: torch.Tensor, *, key_type: int | None = None, descending: bool = False,
                reverse_desc: bool = False, perm64: bool = False, want_keys: bool = False, perm_out=None,
                keys_out=None, stream=None):
        """key_type overrides the dtype-derived type (e.g. an int32 tensor holding u32 keys)."""
        kt = key_type_of(keys) if key_type is None else key_type
        n = keys.shape[0]
        flags = (N.SORT_DESC if descending else 0) | (N.SORT_DESC_REVERSE if reverse_desc else 0) \
            | (N.SORT_PERM64 if perm64 else 0)
        if perm_out is None:
            perm_out = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device=keys.device)
        if want_keys and keys_out is None:
            keys_out = torch.empty_like(keys)
        N.check(self.lib.vdf_argsort(self.h, kt, _ptr(keys), n, flags, _ptr(perm_out),
                                     _ptr(keys_out) if want_keys else None, C.c_void_p(_stream_ptr(stream))))
        return (perm_out, keys_out) if want_keys else perm_out

    # ----------------------------------------------------------------- gather
    def gather(self, src: torch.Tensor, perm: torch.Tensor, out=None, stream=None):
        n = perm.shape[0]
        eb = src.element_size() * (src.shape[1] if src.dim() == 2 else 1)
        if out is None:
            out = torch.empty((n,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
        perm64 = 1 if perm.element_size() == 8 else 0
        N.check(self.lib.vdf_gather(self.h, eb, _ptr(src), _ptr(perm), perm64, n, _ptr(out),
                                    C.c_void_p(_stream_ptr(stream))))
        return out

    # ---------------------------------------------------------------- groupby
    def groupby(self, keys: torch.Tensor, vals=(), aggs: int = N.AGG_ALL, want_perm: bool = True,
                stream=None, trim: bool = True, key_type: int | None = None, val_types=None):
        kt = key_type_of(keys) if key_type is None else key_type
        n = keys.shape[0]
        dev = keys.device
        out = N.VdfGroupbyOut()
        ng = torch.zeros(1, dtype=torch.int64, device=dev)
        gkeys = torch.empty_like(keys)
        offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
        counts = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        perm = torch.empty(max(n, 1), dtype=torch.int32, device=dev) if want_perm else None
        out.n_groups = ng.data_ptr()
        out.group_keys = gkeys.data_ptr()
        out.offsets = offsets.data_ptr()
        out.counts = counts.data_ptr()
        out.perm = perm.data_ptr() if perm is not None else None
        cols = (N.VdfValueCol * max(len(vals), 1))()
        agg_t = []
        for c, v in enumerate(vals):
            vt = val_type_of(v) if val_types is None else val_types[c]
            cols[c].type = vt
            cols[c].data = v.data_ptr()
            d = {}
            sum_dt = torch.float64 if vt == N.VAL_F64 else torch.int64
            if aggs & N.AGG_SUM:
                d["sum"] = torch.empty(max(n, 1), dtype=sum_dt, device=dev)
                out.aggs[c].sum = d["sum"].data_ptr()
            if aggs & N.AGG_MEAN:
                d["mean"] = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
                out.aggs[c].mean = d["mean"].data_ptr()
            if aggs & N.AGG_MIN:
                d["min"] = torch.empty(max(n, 1), dtype=v.dtype, device=dev)
                out.aggs[c].min = d["min"].data_ptr()
            if aggs & N.AGG_MAX:
                d["max"] = torch.empty(max(n, 1), dtype=v.dtype, device=dev)
                out.aggs[c].max = d["max"].data_ptr()
            agg_t.append(d)
        N.check(self.lib.vdf_groupby(self.h, kt, _ptr(keys), n, cols, len(vals), aggs, C.byref(out),
                                     C.c_void_p(_stream_ptr(stream))))
        if not trim:
            return GroupbyResult(ng, gkeys, offsets, counts, perm, agg_t)
        g = int(ng.item())
        return GroupbyResult(g, gkeys[:g], offsets[:g + 1], counts[:g],
                             perm[:n] if perm is not None else None,
                             [{k: t[:g] for k, t in d.items()} for d in agg_t])

    # ----------------------------------------------------------------- tuning
    def tune(self, name: str, value: int):
        N.check(self.lib.vdf_tune_set(self.h, name.encode(), int(value)))

    def tune_get(self, name: str) -> int:
        v = C.c_int64()
        N.check(self.lib.vdf_tune_get(self.h, name.encode(), C.byref(v)))
        return v.value

    # -------------------------------------------------------------- profiling
    def profile(self, on: bool = True):
        N.check(self.lib.vdf_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        N.check(self.lib.vdf_profile_reset(self.h))

    def profile_read(self, kernel: str):
        ms = C.c_double()
        launches = C.c_uint64()
        nbytes = C.c_uint64()
        N.check(self.lib.vdf_profile_read(self.h, kernel.encode(), C.byref(ms), C.byref(launches),
                                          C.byref(nbytes)))
        return ms.value, launches.value, nbytes.value
