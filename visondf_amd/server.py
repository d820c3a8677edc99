"""Group-by request server over device-resident tables (continuous batching).

Python face of the vdf_server_* entry points of libvdf_hip.so: tables are
registered once and stay in HBM; each submitted request is a group-by of one
table; the library's worker threads batch queued requests into segmented
group-by launch sequences as arena memory frees up, several batches in flight
on separate streams.  Latency stamps come from the library's steady clock
(submit, launch start, results ready on the host).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _native as N
from .device import key_type_of, track, val_type_of

_HIP_D2D = 3  # hipMemcpyDeviceToDevice
_hip = None


def _hip_lib():
    """The HIP runtime already loaded in this process (torch's or the system one)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


def _copy_out(ptr: int | None, count: int, dtype, device, shape_tail=()) -> torch.Tensor | None:
    if not ptr:
        return None
    t = torch.empty((count,) + tuple(shape_tail), dtype=dtype, device=device)
    nbytes = t.numel() * t.element_size()
    if nbytes:
        torch.cuda.synchronize(device)
        rc = _hip_lib().hipMemcpy(t.data_ptr(), ptr, nbytes, _HIP_D2D)
        if rc != 0:
            raise N.VdfError(N.VDF_ERR_HIP, f"hipMemcpy of a served result failed ({rc})")
    return t


@dataclass
class Served:
    request_id: int
    status: int
    n_groups: int
    batch_tables: int
    submit_ns: int
    start_ns: int
    done_ns: int
    out: N.VdfGroupbyOut

    @property
    def latency_s(self) -> float:
        return (self.done_ns - self.submit_ns) * 1e-9

    @property
    def queue_s(self) -> float:
        return (self.start_ns - self.submit_ns) * 1e-9


class Server:
    def __init__(self, device: int = 0, arena_bytes: int = 8 << 30, workers: int = 3,
                 max_batch_tables: int = 64, max_batch_rows: int = 0):
        self.lib = N.load()
        self.device = device
        cfg = N.VdfServerConfig(device=device, workers=workers, max_batch_tables=max_batch_tables, reserved=0,
                                arena_bytes=arena_bytes, max_batch_rows=max_batch_rows)
        h = C.c_void_p()
        N.check(self.lib.vdf_server_create(C.byref(h), C.byref(cfg)))
        self.h = h
        track(self)
        self._tables = {}  # id -> (keys, vals, key_type, val dtypes) kept alive for the server's life

    def close(self):
        if self.h:
            N.check(self.lib.vdf_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_table(self, keys: torch.Tensor, vals=(), *, key_type: int | None = None, val_types=None,
                  max_groups: int = 0) -> int:
        kt = key_type_of(keys) if key_type is None else key_type
        vals = list(vals)
        vts = [val_type_of(v) if val_types is None else val_types[c] for c, v in enumerate(vals)]
        n = keys.shape[0]
        vt_arr = (C.c_int * max(len(vals), 1))(*vts)
        vp_arr = (C.c_void_p * max(len(vals), 1))(*[v.data_ptr() for v in vals])
        tid = C.c_uint32()
        N.check(self.lib.vdf_server_add_table(self.h, kt, keys.data_ptr() if n else None, n, vt_arr, vp_arr,
                                              len(vals), max_groups, C.byref(tid)))
        self._tables[tid.value] = (keys, vals, kt, [v.dtype for v in vals])
        return tid.value

    def submit(self, table_id: int, aggs: int = N.AGG_ALL, want_perm: bool = False) -> int:
        rid = C.c_uint64()
        N.check(self.lib.vdf_server_submit(self.h, table_id, aggs, 1 if want_perm else 0, C.byref(rid)))
        return rid.value

    def wait(self, request_id: int, timeout_s: float | None = None) -> Served | None:
        """The finished request, or None when timeout_s passed first."""
        v = N.VdfServed()
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        st = self.lib.vdf_server_wait(self.h, request_id, us, C.byref(v))
        if st == N.VDF_ERR_TIMEOUT:
            return None
        N.check(st)
        return Served(request_id, v.status, v.n_groups, v.batch_tables, v.submit_ns, v.start_ns, v.done_ns, v.out)

    def release(self, request_id: int):
        N.check(self.lib.vdf_server_release(self.h, request_id))

    def stats(self) -> dict:
        s = N.VdfServerStats()
        N.check(self.lib.vdf_server_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in N.VdfServerStats._fields_}

    def tensors(self, table_id: int, r: Served, aggs: int = N.AGG_ALL, want_perm: bool = False) -> dict:
        """Copies a finished request's results into new device tensors:
        keys, offsets, counts, perm and per value column sum / mean / min / max."""
        if r.status != N.VDF_OK:
            raise N.VdfError(r.status, "request failed")
        keys, vals, kt, vdts = self._tables[table_id]
        g, n, dev = r.n_groups, keys.shape[0], keys.device
        o = r.out
        res = {"n_groups": g,
               "keys": _copy_out(o.group_keys, g, keys.dtype, dev, keys.shape[1:]),
               "offsets": _copy_out(o.offsets, g + 1, torch.int32, dev),
               "counts": _copy_out(o.counts, g, torch.int32, dev),
               "perm": _copy_out(o.perm, n, torch.int32, dev),
               "aggs": []}
        for c, dt in enumerate(vdts):
            a = o.aggs[c]
            sum_dt = torch.float64 if dt == torch.float64 else torch.int64
            res["aggs"].append({k: t for k, t in (("sum", _copy_out(a.sum, g, sum_dt, dev)),
                                                  ("mean", _copy_out(a.mean, g, torch.float64, dev)),
                                                  ("min", _copy_out(a.min, g, dt, dev)),
                                                  ("max", _copy_out(a.max, g, dt, dev))) if t is not None})
        return res
