"""Python access to libvdf_hip.so over torch device tensors.

torch supplies device memory and the current HIP stream; every operation runs
in the HIP library through the C ABI.  There is no CPU fallback: without a
GPU or without the built library these calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import torch

from . import _native as N

# Library objects still open when the interpreter exits are closed by an
# atexit hook (registered after torch's, so it runs before torch and the HIP
# runtime tear down), never by __del__ during interpreter teardown.
# Objects that use a handle (groupings, captured plans, servers) close
# before the handles.
_live_users = weakref.WeakSet()
_live_handles = weakref.WeakSet()


def _close_live():
    for group in (_live_users, _live_handles):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(_close_live)


def track(obj, handle: bool = False):
    """Registers a library object whose close() must run before exit."""
    (_live_handles if handle else _live_users).add(obj)
    return obj

_KEY_DTYPES = {
    N.KEY_U32: (torch.uint32, 4),
    N.KEY_I32: (torch.int32, 4),
    N.KEY_U64: (torch.uint64, 8),
    N.KEY_I64: (torch.int64, 8),
    N.KEY_CHAR16: (torch.uint8, 16),
    N.KEY_F32: (torch.float32, 4),
    N.KEY_F64: (torch.float64, 8),
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


def key_col_of(t: torch.Tensor) -> tuple[int, int]:
    """(key type, cell width) of one column of a multi-column key: a 2-D
    uint8 tensor of width other than 16 is a fixed-width byte-string column."""
    if t.dtype == torch.uint8 and t.dim() == 2 and t.shape[1] != 16:
        return N.KEY_BYTES, int(t.shape[1])
    kt = key_type_of(t)
    return kt, _KEY_DTYPES[kt][1]


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
        track(self, handle=True)

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

    def sort_by(self, keys: torch.Tensor, vals: torch.Tensor, *, key_type: int | None = None,
                descending: bool = False, reverse_desc: bool = False, perm64: bool = False,
                want_perm: bool = True, want_keys: bool = False, perm_out=None, vals_out=None, stream=None):
        """Sorts `keys` and carries one 4- or 8-byte value column through the
        passes: returns (perm or None, vals in sorted order[, sorted keys])."""
        kt = key_type_of(keys) if key_type is None else key_type
        n = keys.shape[0]
        flags = (N.SORT_DESC if descending else 0) | (N.SORT_DESC_REVERSE if reverse_desc else 0) \
            | (N.SORT_PERM64 if perm64 else 0)
        if want_perm and perm_out is None:
            perm_out = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device=keys.device)
        if vals_out is None:
            vals_out = torch.empty_like(vals)
        keys_out = torch.empty_like(keys) if want_keys else None
        N.check(self.lib.vdf_sort_by(self.h, kt, _ptr(keys), n, flags, _ptr(perm_out) if want_perm else None,
                                     _ptr(keys_out), vals.element_size(), _ptr(vals), _ptr(vals_out),
                                     C.c_void_p(_stream_ptr(stream))))
        out = (perm_out if want_perm else None, vals_out)
        return out + (keys_out,) if want_keys else out

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
    @staticmethod
    def _alloc_out(keys: torch.Tensor, vts, dtypes, aggs: int, want_perm: bool, capacity: int = 0):
        """Output buffers of one table (room for `capacity` groups, 0 = n) and the
        C descriptor."""
        n = keys.shape[0]
        cap = n if capacity <= 0 else min(int(capacity), n)
        dev = keys.device
        out = N.VdfGroupbyOut()
        out.capacity = cap
        t = {"cap": cap, "ng": torch.zeros(1, dtype=torch.int64, device=dev),
             "keys": torch.empty((max(cap, 1),) + tuple(keys.shape[1:]), dtype=keys.dtype, device=dev),
             "offsets": torch.empty(cap + 1, dtype=torch.int32, device=dev),
             "counts": torch.empty(max(cap, 1), dtype=torch.int32, device=dev),
             "perm": torch.empty(max(n, 1), dtype=torch.int32, device=dev) if want_perm else None,
             "aggs": []}
        out.n_groups = t["ng"].data_ptr()
        out.group_keys = t["keys"].data_ptr() if cap else None
        out.offsets = t["offsets"].data_ptr()
        out.counts = t["counts"].data_ptr()
        out.perm = t["perm"].data_ptr() if want_perm else None
        for c, (vt, dt) in enumerate(zip(vts, dtypes)):
            d = {}
            sum_dt = torch.float64 if vt == N.VAL_F64 else torch.int64
            if aggs & N.AGG_SUM:
                d["sum"] = torch.empty(max(cap, 1), dtype=sum_dt, device=dev)
                out.aggs[c].sum = d["sum"].data_ptr()
            if aggs & N.AGG_MEAN:
                d["mean"] = torch.empty(max(cap, 1), dtype=torch.float64, device=dev)
                out.aggs[c].mean = d["mean"].data_ptr()
            if aggs & N.AGG_MIN:
                d["min"] = torch.empty(max(cap, 1), dtype=dt, device=dev)
                out.aggs[c].min = d["min"].data_ptr()
            if aggs & N.AGG_MAX:
                d["max"] = torch.empty(max(cap, 1), dtype=dt, device=dev)
                out.aggs[c].max = d["max"].data_ptr()
            t["aggs"].append(d)
        return out, t

    @staticmethod
    def _result(t, n: int, trim: bool) -> GroupbyResult:
        if not trim:
            return GroupbyResult(t["ng"], t["keys"], t["offsets"], t["counts"], t["perm"], t["aggs"])
        g = int(t["ng"].item())
        if g > t["cap"]:
            raise N.VdfError(N.VDF_ERR_INVALID, f"{g} groups exceed the output capacity {t['cap']}")
        return GroupbyResult(g, t["keys"][:g], t["offsets"][:g + 1], t["counts"][:g],
                             t["perm"][:n] if t["perm"] is not None else None,
                             [{k: x[:g] for k, x in d.items()} for d in t["aggs"]])

    def groupby(self, keys: torch.Tensor, vals=(), aggs: int = N.AGG_ALL, want_perm: bool = True,
                stream=None, trim: bool = True, key_type: int | None = None, val_types=None,
                capacity: int = 0):
        kt = key_type_of(keys) if key_type is None else key_type
        n = keys.shape[0]
        vts = [val_type_of(v) if val_types is None else val_types[c] for c, v in enumerate(vals)]
        out, t = self._alloc_out(keys, vts, [v.dtype for v in vals], aggs, want_perm, capacity)
        cols = (N.VdfValueCol * max(len(vals), 1))()
        for c, v in enumerate(vals):
            cols[c].type = vts[c]
            cols[c].data = v.data_ptr()
        N.check(self.lib.vdf_groupby(self.h, kt, _ptr(keys), n, cols, len(vals), aggs, C.byref(out),
                                     C.c_void_p(_stream_ptr(stream))))
        return self._result(t, n, trim)

    def groupby_batch(self, tables, aggs: int = N.AGG_ALL, want_perm: bool = True, stream=None,
                      trim: bool = True, key_type: int | None = None, val_types=None, capacities=None):
        """tables: list of (keys, [value columns]); all tables share key and value types.
        capacities: per-table group capacity of the outputs (0 / None = rows)."""
        if not tables:
            return []
        k0, v0 = tables[0]
        kt = key_type_of(k0) if key_type is None else key_type
        vts = [val_type_of(v) if val_types is None else val_types[c] for c, v in enumerate(v0)]
        nv = len(v0)
        tabs = (N.VdfTable * len(tables))()
        outs = (N.VdfGroupbyOut * len(tables))()
        keep = []
        for i, (k, vs) in enumerate(tables):
            tabs[i].keys = k.data_ptr() if k.shape[0] else None
            tabs[i].n = k.shape[0]
            for c, v in enumerate(vs):
                tabs[i].vals[c] = v.data_ptr() if k.shape[0] else None
            o, t = self._alloc_out(k, vts, [v.dtype for v in vs], aggs, want_perm,
                                   capacities[i] if capacities is not None else 0)
            outs[i] = o
            keep.append(t)
        vt_arr = (C.c_int * max(nv, 1))(*vts)
        N.check(self.lib.vdf_groupby_batch(self.h, kt, tabs, len(tables), vt_arr, nv, aggs, outs,
                                           C.c_void_p(_stream_ptr(stream))))
        return [self._result(t, k.shape[0], trim) for t, (k, _) in zip(keep, tables)]

    def argsort_batch(self, keys_list, *, key_type: int | None = None, descending: bool = False,
                      reverse_desc: bool = False, perm64: bool = False, want_keys: bool = False, stream=None):
        if not keys_list:
            return []
        kt = key_type_of(keys_list[0]) if key_type is None else key_type
        flags = (N.SORT_DESC if descending else 0) | (N.SORT_DESC_REVERSE if reverse_desc else 0) \
            | (N.SORT_PERM64 if perm64 else 0)
        segs = (N.VdfSortSeg * len(keys_list))()
        res = []
        for i, k in enumerate(keys_list):
            n = k.shape[0]
            p = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device=k.device)
            ko = torch.empty_like(k) if want_keys else None
            segs[i].keys = k.data_ptr() if n else None
            segs[i].n = n
            segs[i].perm_out = p.data_ptr() if n else None
            segs[i].sorted_keys_out = ko.data_ptr() if (want_keys and n) else None
            res.append((p, ko) if want_keys else p)
        N.check(self.lib.vdf_argsort_batch(self.h, kt, segs, len(keys_list), flags,
                                           C.c_void_p(_stream_ptr(stream))))
        return res

    def groupby_multi(self, keys, vals=(), aggs: int = N.AGG_ALL, want_perm: bool = True, stream=None,
                      trim: bool = True, key_types=None, val_types=None, capacity: int = 0):
        """Group by several key columns (any key type; a (n, w) uint8 column
        with w != 16 is a byte-string column); the result's keys is a list
        with one tensor of group keys per column."""
        keys = list(keys)
        n = keys[0].shape[0]
        kcs = [key_col_of(k) if key_types is None else (key_types[c], int(k.shape[1]) if k.dim() == 2 else 0)
               for c, k in enumerate(keys)]
        kts = [t for t, _ in kcs]
        vts = [val_type_of(v) if val_types is None else val_types[c] for c, v in enumerate(vals)]
        out, t = self._alloc_out(keys[0], vts, [v.dtype for v in vals], aggs, want_perm, capacity)
        cap = t["cap"]
        gk = [torch.empty((max(cap, 1),) + tuple(k.shape[1:]), dtype=k.dtype, device=k.device) for k in keys]
        kc = (N.VdfKeyCol * len(keys))()
        for c, k in enumerate(keys):
            kc[c].type = kts[c]
            kc[c].data = _ptr(k)
            kc[c].width = kcs[c][1] if kts[c] == N.KEY_BYTES else 0
        cols = (N.VdfValueCol * max(len(vals), 1))()
        for c, v in enumerate(vals):
            cols[c].type = vts[c]
            cols[c].data = v.data_ptr()
        gptrs = (C.c_void_p * len(keys))(*[g.data_ptr() for g in gk])
        N.check(self.lib.vdf_groupby_multi(self.h, kc, len(keys), n, cols, len(vals), aggs, C.byref(out), gptrs,
                                           C.c_void_p(_stream_ptr(stream))))
        t["keys"] = gk
        if not trim:
            return GroupbyResult(t["ng"], gk, t["offsets"], t["counts"], t["perm"], t["aggs"])
        g = int(t["ng"].item())
        if g > cap:
            raise N.VdfError(N.VDF_ERR_INVALID, f"{g} groups exceed the output capacity {cap}")
        return GroupbyResult(g, [x[:g] for x in gk], t["offsets"][:g + 1], t["counts"][:g],
                             t["perm"][:n] if t["perm"] is not None else None,
                             [{k: x[:g] for k, x in d.items()} for d in t["aggs"]])

    def argsort_multi(self, keys, *, key_types=None, descending: bool = False, reverse_desc: bool = False,
                      perm64: bool = False, stream=None) -> torch.Tensor:
        keys = list(keys)
        n = keys[0].shape[0]
        kcs = [key_col_of(k) if key_types is None else (key_types[c], int(k.shape[1]) if k.dim() == 2 else 0)
               for c, k in enumerate(keys)]
        flags = (N.SORT_DESC if descending else 0) | (N.SORT_DESC_REVERSE if reverse_desc else 0) \
            | (N.SORT_PERM64 if perm64 else 0)
        perm = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device=keys[0].device)
        kc = (N.VdfKeyCol * len(keys))()
        for c, k in enumerate(keys):
            kc[c].type = kcs[c][0]
            kc[c].data = _ptr(k)
            kc[c].width = kcs[c][1] if kcs[c][0] == N.KEY_BYTES else 0
        N.check(self.lib.vdf_argsort_multi(self.h, kc, len(keys), n, flags, _ptr(perm),
                                           C.c_void_p(_stream_ptr(stream))))
        return perm

    def argsort_bytes(self, keys: torch.Tensor, *, descending: bool = False, reverse_desc: bool = False,
                      perm64: bool = False, want_sorted: bool = False, stream=None):
        """Stable argsort of fixed-width byte strings: a (n, width) uint8
        tensor, one padded string per row, ordered bytewise as signed chars.
        Returns the permutation, or (permutation, rows in sorted order)."""
        if keys.dtype != torch.uint8 or keys.dim() != 2:
            raise N.VdfError(N.VDF_ERR_INVALID, "argsort_bytes: keys must be a (n, width) uint8 tensor")
        keys = keys.contiguous()
        n, width = keys.shape
        flags = (N.SORT_DESC if descending else 0) | (N.SORT_DESC_REVERSE if reverse_desc else 0) \
            | (N.SORT_PERM64 if perm64 else 0)
        perm = torch.empty(n, dtype=torch.int64 if perm64 else torch.int32, device=keys.device)
        out = torch.empty_like(keys) if want_sorted else None
        N.check(self.lib.vdf_argsort_bytes(self.h, _ptr(keys), width, n, flags, _ptr(perm),
                                           _ptr(out) if out is not None else None,
                                           C.c_void_p(_stream_ptr(stream))))
        return (perm, out) if want_sorted else perm

    def groupby_bytes(self, keys: torch.Tensor, vals=(), aggs: int = N.AGG_ALL, want_perm: bool = True,
                      stream=None, trim: bool = True, val_types=None, capacity: int = 0):
        """Group-by over fixed-width byte strings ((n, width) uint8, padded);
        group keys come back as (groups, width) uint8 rows, ascending."""
        if keys.dtype != torch.uint8 or keys.dim() != 2:
            raise N.VdfError(N.VDF_ERR_INVALID, "groupby_bytes: keys must be a (n, width) uint8 tensor")
        keys = keys.contiguous()
        n, width = keys.shape
        vts = [val_type_of(v) if val_types is None else val_types[c] for c, v in enumerate(vals)]
        out, t = self._alloc_out(keys, vts, [v.dtype for v in vals], aggs, want_perm, capacity)
        cols = (N.VdfValueCol * max(len(vals), 1))()
        for c, v in enumerate(vals):
            cols[c].type = vts[c]
            cols[c].data = v.data_ptr()
        N.check(self.lib.vdf_groupby_bytes(self.h, _ptr(keys) if n else None, width, n, cols, len(vals), aggs,
                                           C.byref(out), C.c_void_p(_stream_ptr(stream))))
        return self._result(t, n, trim)

    def broadcast(self, group_vals: torch.Tensor, result: "GroupbyResult", n: int, to_f64: bool = False,
                  stream=None) -> torch.Tensor:
        """Per-row column: row r gets the value of its group (the group-by's
        offsets / n_groups / perm from an untrimmed result)."""
        st = val_type_of(group_vals)
        dt = N.VAL_F64 if to_f64 else st
        out = torch.empty(n, dtype=torch.float64 if to_f64 else group_vals.dtype, device=group_vals.device)
        ng = result.n_groups
        if not isinstance(ng, torch.Tensor):
            raise ValueError("broadcast needs an untrimmed group-by result (device group count)")
        N.check(self.lib.vdf_broadcast(self.h, st, dt, group_vals.data_ptr(), result.offsets.data_ptr(),
                                       ng.data_ptr(), result.offsets.shape[0] - 1, result.perm.data_ptr(), n,
                                       _ptr(out), C.c_void_p(_stream_ptr(stream))))
        return out

    def grouping(self, keys: torch.Tensor, *, key_type: int | None = None, stream=None) -> "Grouping":
        """Sort a key column once; aggregate value columns against it later."""
        return Grouping(self, keys, key_type_of(keys) if key_type is None else key_type, stream)

    def prepare_groupby_batch(self, tables, aggs: int = N.AGG_ALL, want_perm: bool = True, key_type=None,
                              val_types=None, capacities=None) -> "PreparedBatch":
        """Descriptors and outputs of one batch, built once for repeated runs."""
        return PreparedBatch(self, tables, aggs, want_perm, key_type, val_types, capacities)

    def autotune_argsort(self, keys: torch.Tensor, perm_out: torch.Tensor, *, key_type: int | None = None,
                         flags: int = 0, reps: int = 3, stream=None) -> dict:
        kt = key_type_of(keys) if key_type is None else key_type
        rep = N.VdfTuneReport()
        N.check(self.lib.vdf_autotune_argsort(self.h, kt, _ptr(keys), keys.shape[0], flags, _ptr(perm_out), reps,
                                              C.c_void_p(_stream_ptr(stream)), C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in N.VdfTuneReport._fields_ if k != "reserved"}

    def autotune_sort_by(self, keys: torch.Tensor, vals: torch.Tensor, perm_out: torch.Tensor | None,
                         vals_out: torch.Tensor, *, key_type: int | None = None, flags: int = 0, reps: int = 3,
                         stream=None) -> dict:
        kt = key_type_of(keys) if key_type is None else key_type
        rep = N.VdfTuneReport()
        N.check(self.lib.vdf_autotune_sort_by(self.h, kt, _ptr(keys), keys.shape[0], flags, _ptr(perm_out),
                                              vals.element_size(), _ptr(vals), _ptr(vals_out), reps,
                                              C.c_void_p(_stream_ptr(stream)), C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in N.VdfTuneReport._fields_ if k != "reserved"}

    # ------------------------------------------------------------- CSV ingest
    def read_csv(self, text, delim: str = ",", quote: str = "'", header: bool = True, int_bytes: int = 8,
                 char_width: int = 16, trailing: str = "0") -> list:
        """Parses CSV text (bytes, or a uint8 device tensor) on the device
        (vdf_read_csv).  Returns [(name, kind, tensor)] per column: DOUBLE
        float64, INT int32/int64, UINT uint32/uint64 (as int tensors of the
        same bits), BOOL int32 0/1, CHAR and STRING (n, width) uint8."""
        if isinstance(text, (bytes, bytearray)):
            buf = torch.frombuffer(bytearray(text), dtype=torch.uint8) if len(text) else torch.empty(0, dtype=torch.uint8)
            dt = buf.cuda()
        else:
            dt = text
        torch.cuda.synchronize()
        opt = N.VdfCsvOptions(delim=delim.encode(), quote=quote.encode(), trailing=trailing.encode(), reserved0=b"\0",
                              header=1 if header else 0, int_bytes=int_bytes, char_width=char_width)
        t = C.c_void_p()
        N.check(self.lib.vdf_read_csv(self.h, C.c_void_p(dt.data_ptr()) if dt.numel() else None, dt.numel(),
                                      C.byref(opt), C.byref(t)))
        try:
            nrow, ncol = C.c_uint64(), C.c_int()
            N.check(self.lib.vdf_csv_info(t, C.byref(nrow), C.byref(ncol)))
            out = []
            for c in range(ncol.value):
                col = N.VdfCsvColumn()
                N.check(self.lib.vdf_csv_column_get(t, c, C.byref(col)))
                n = nrow.value
                if col.kind == N.CSV_DOUBLE:
                    x = torch.empty(n, dtype=torch.float64, device=dt.device)
                elif col.kind in (N.CSV_INT, N.CSV_UINT):
                    x = torch.empty(n, dtype=torch.int64 if col.width == 8 else torch.int32, device=dt.device)
                elif col.kind == N.CSV_BOOL:
                    x = torch.empty(n, dtype=torch.int32, device=dt.device)
                else:
                    x = torch.empty((n, col.width), dtype=torch.uint8, device=dt.device)
                nbytes = x.numel() * x.element_size()
                if nbytes:
                    from .server import _hip_lib

                    rc = _hip_lib().hipMemcpy(x.data_ptr(), col.data, nbytes, 3)
                    if rc != 0:
                        raise N.VdfError(N.VDF_ERR_HIP, f"hipMemcpy of a CSV column failed ({rc})")
                out.append(((col.name or b"").decode(), col.kind, x))
            return out
        finally:
            N.check(self.lib.vdf_csv_free(t))

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


class PreparedBatch:
    """A group-by batch whose C descriptors and output buffers are built once.

    run() calls vdf_groupby_batch; autotune() tunes the handle on this batch;
    capture() records the launch sequence as a HIP-graph plan and replay()
    launches it (the same tables and outputs, or set_tables() of the same
    shape first)."""

    def __init__(self, vdf: Vdf, tables, aggs, want_perm, key_type, val_types, capacities):
        self.v = vdf
        k0, v0 = tables[0]
        self.kt = key_type_of(k0) if key_type is None else key_type
        self.vts = [val_type_of(x) if val_types is None else val_types[c] for c, x in enumerate(v0)]
        self.nv = len(v0)
        self.aggs = aggs
        self.ntab = len(tables)
        self.tabs = (N.VdfTable * self.ntab)()
        self.outs = (N.VdfGroupbyOut * self.ntab)()
        self.keep = []
        self.sizes = []
        for i, (k, vs) in enumerate(tables):
            o, t = Vdf._alloc_out(k, self.vts, [x.dtype for x in vs], aggs, want_perm,
                                  capacities[i] if capacities is not None else 0)
            self.outs[i] = o
            self.keep.append(t)
            self.sizes.append(k.shape[0])
        self.set_tables(tables)
        self.vt_arr = (C.c_int * max(self.nv, 1))(*self.vts)
        self.plan = None
        track(self)

    def set_tables(self, tables):
        """Points the batch at other tables with the same row counts."""
        if len(tables) != self.ntab:
            raise ValueError("table count differs")
        self.tables = list(tables)
        for i, (k, vs) in enumerate(tables):
            if k.shape[0] != self.sizes[i]:
                raise ValueError(f"table {i}: {k.shape[0]} rows, batch has {self.sizes[i]}")
            self.tabs[i].keys = k.data_ptr() if k.shape[0] else None
            self.tabs[i].n = k.shape[0]
            for c, x in enumerate(vs):
                self.tabs[i].vals[c] = x.data_ptr() if k.shape[0] else None

    def run(self, stream=None):
        N.check(self.v.lib.vdf_groupby_batch(self.v.h, self.kt, self.tabs, self.ntab, self.vt_arr, self.nv,
                                             self.aggs, self.outs, C.c_void_p(_stream_ptr(stream))))

    def autotune(self, reps: int = 3, stream=None) -> dict:
        rep = N.VdfTuneReport()
        N.check(self.v.lib.vdf_autotune_groupby_batch(self.v.h, self.kt, self.tabs, self.ntab, self.vt_arr, self.nv,
                                                      self.aggs, self.outs, reps, C.c_void_p(_stream_ptr(stream)),
                                                      C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in N.VdfTuneReport._fields_ if k != "reserved"}

    def capture(self):
        if self.plan is None:
            p = C.c_void_p()
            N.check(self.v.lib.vdf_plan_groupby_batch(self.v.h, self.kt, self.tabs, self.ntab, self.vt_arr, self.nv,
                                                      self.aggs, C.byref(p)))
            self.plan = p
        return self

    def replay(self, stream=None):
        if self.plan is None:
            raise RuntimeError("capture() first")
        N.check(self.v.lib.vdf_plan_launch(self.plan, self.tabs, self.ntab, self.outs,
                                           C.c_void_p(_stream_ptr(stream))))

    def release_plan(self):
        if self.plan is not None:
            N.check(self.v.lib.vdf_plan_destroy(self.plan))
            self.plan = None

    def close(self):
        self.release_plan()

    def __del__(self):
        try:
            self.release_plan()
        except Exception:
            pass

    def result(self, i: int, trim: bool = True) -> GroupbyResult:
        return Vdf._result(self.keep[i], self.sizes[i], trim)


class Grouping:
    """A cached grouping of one key column (vdf_grouping_*)."""

    def __init__(self, vdf: Vdf, keys: torch.Tensor, key_type: int, stream=None):
        self.v = vdf
        self.keys = keys
        self.n = keys.shape[0]
        g = C.c_void_p()
        N.check(vdf.lib.vdf_grouping_create(vdf.h, key_type, _ptr(keys), self.n, C.c_void_p(_stream_ptr(stream)),
                                            C.byref(g)))
        self.g = g
        track(self)

    def aggregate(self, vals=(), aggs: int = N.AGG_ALL, want_perm: bool = True, stream=None, trim: bool = True,
                  val_types=None) -> GroupbyResult:
        vals = list(vals)
        vts = [val_type_of(x) if val_types is None else val_types[c] for c, x in enumerate(vals)]
        out, t = Vdf._alloc_out(self.keys, vts, [x.dtype for x in vals], aggs, want_perm)
        cols = (N.VdfValueCol * max(len(vals), 1))()
        for c, x in enumerate(vals):
            cols[c].type = vts[c]
            cols[c].data = x.data_ptr()
        N.check(self.v.lib.vdf_grouping_aggregate(self.g, cols, len(vals), aggs, C.byref(out),
                                                  C.c_void_p(_stream_ptr(stream))))
        return Vdf._result(t, self.n, trim)

    def close(self):
        if self.g:
            N.check(self.v.lib.vdf_grouping_destroy(self.g))
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
